"""Pins of the FP64 CPU oracle against golden values, closed forms, invariants
and exact brute force (CPU only)."""
import json
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp

from geometry_ref import chord_endpoints, chord_length, exact_ray_cells

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# ------------------------------------------------------------------ MATLAB built-ins
def test_mt19937_matches_matlab_documented_rand(oracle):
    g = json.load(open(os.path.join(GOLDEN, "matlab_rand_default.json")))
    got = oracle.mt_rand(g["seed"], 9)
    assert np.max(np.abs(got - np.array(g["values_column_major"]))) < g["tolerance"]


@pytest.mark.parametrize("seed", [1, 42, 2**31 - 1])
def test_mt19937_matches_numpy_legacy_generator(oracle, seed):
    ref = np.random.RandomState(seed).random_sample(1000)
    assert np.array_equal(oracle.mt_rand(seed, 1000), ref)


def test_sind_cosd_exact_at_right_angles(oracle):
    for k in range(-8, 9):
        s, c = oracle.sind(90 * k), oracle.cosd(90 * k)
        assert (s, c) == [(0.0, 1.0), (1.0, 0.0), (0.0, -1.0), (-1.0, 0.0)][k % 4]
    assert math.copysign(1.0, -oracle.sind(0.0)) == -1.0      # -sind(0) is -0


def test_sind_cosd_accuracy_and_symmetry(oracle):
    rng = np.random.default_rng(0)
    for deg in np.concatenate([rng.uniform(-720, 720, 500), np.arange(0, 360, 0.5)]):
        s, c = oracle.sind(deg), oracle.cosd(deg)
        rs, rc = math.sin(math.radians(deg)), math.cos(math.radians(deg))
        assert abs(s - rs) <= 4e-16 + 1e-13 * abs(rs) and abs(c - rc) <= 4e-16 + 1e-13 * abs(rc)
        assert oracle.sind(-deg) == -s and oracle.cosd(-deg) == c
        assert abs(s * s + c * c - 1) < 4e-16


def test_linspace_integer_spacing_exact(oracle):
    for p in (2, 5, 91, 724, 725, 1448):
        v = oracle.linspace(-(p - 1) / 2, (p - 1) / 2, p)
        assert np.array_equal(v, np.arange(p) - (p - 1) / 2)


# ------------------------------------------------------------------ paralleltomo
def par_rays(oracle, theta, p, d):
    x0 = oracle.linspace(-d / 2, d / 2, p)
    out = []
    for th in theta:
        s, c = oracle.sind(th), oracle.cosd(th)
        for j in range(p):
            out.append((c * x0[j] - s * 0.0, s * x0[j] + c * 0.0, -s, c))
    return out


def test_parallel_axis_aligned_rays(oracle):
    N, p = 8, 9
    A = oracle.paralleltomo(N, [0, 90, 180, 270], p).toarray()
    # theta = 0: vertical ray at x = -4+j covers image column j; the ray on the right edge is empty
    for j in range(p):
        row = A[j]
        if j == N:
            assert not row.any()
        else:
            assert np.array_equal(np.nonzero(row)[0], j * N + np.arange(N)) and np.all(row[row != 0] == 1.0)
    # theta = 90: horizontal ray at y = -4+j covers image row N-1-j; the top edge ray is empty
    for j in range(p):
        row = A[p + j]
        if j == N:
            assert not row.any()
        else:
            assert np.array_equal(np.nonzero(row)[0], np.arange(N) * N + (N - 1 - j))
    # 180 and 270 degrees trace the same lines in the opposite order
    assert np.array_equal(A[2 * p:3 * p][::-1], A[0:p])
    assert np.array_equal(A[3 * p:4 * p][::-1], A[p:2 * p])


def test_parallel_diagonal_ray_sqrt2_segments(oracle):
    N = 8
    A = oracle.paralleltomo(N, [45], 1, 0.0)          # one ray through the origin
    v = A.toarray()[0]
    nz = v[v != 0]
    assert len(nz) == N and np.allclose(nz, math.sqrt(2), rtol=0, atol=1e-12)
    # direction (-sind 45, cosd 45): the line y = -x, cells (fx, fy = N-1-fx),
    # column fx*N + (N-1-fy) = fx*(N+1)
    cols = np.nonzero(v)[0]
    assert np.array_equal(cols, np.arange(N) * (N + 1))


@pytest.mark.parametrize("N,p,d", [(16, 23, None), (15, 20, 25.0), (32, 45, 40.0)])
def test_parallel_row_sums_are_chord_lengths(oracle, N, p, d):
    rng = np.random.default_rng(N)
    theta = np.concatenate([rng.uniform(0, 360, 20), [0, 30, 45, 90, 135, 180]])
    d = float(p - 1) if d is None else d
    A = oracle.paralleltomo(N, theta, p, d)
    rs = np.asarray(A.sum(axis=1)).ravel()
    for r, (x0, y0, a, b) in enumerate(par_rays(oracle, theta, p, d)):
        edge = (b == 0 and abs(y0 - N / 2) < 1e-15) or (a == 0 and abs(x0 - N / 2) < 1e-15)
        ref = 0.0 if edge else chord_length(x0, y0, a, b, N / 2)
        assert abs(rs[r] - ref) <= 1e-12 * max(1.0, ref), (r, rs[r], ref)


def test_parallel_nnz_is_cells_crossed(oracle):
    N, p = 32, 45
    rng = np.random.default_rng(3)
    theta = rng.uniform(0.5, 179.5, 30)
    A = oracle.paralleltomo(N, theta, p, 40.0)
    nnz = np.diff(A.indptr)
    for r, (x0, y0, a, b) in enumerate(par_rays(oracle, theta, p, 40.0)):
        ep = chord_endpoints(x0, y0, a, b, N / 2)
        if ep is None:
            assert nnz[r] == 0
            continue
        (px, py), (qx, qy) = ep
        # generic rays only: skip lines through a lattice point (corner), where
        # two crossings coincide and one cell fewer is entered
        ys = [y0 + ((k - x0) / a) * b for k in range(-N // 2, N // 2 + 1)]
        if any(abs(y - round(y)) < 1e-9 and abs(y) <= N / 2 for y in ys):
            continue
        inner = lambda u, v: sum(1 for k in range(-N // 2 + 1, N // 2) if min(u, v) < k < max(u, v))
        assert nnz[r] == 1 + inner(px, qx) + inner(py, qy)


@pytest.mark.parametrize("N", [3, 4, 7, 8])
def test_parallel_matches_exact_rational_bruteforce(oracle, N):
    rng = np.random.default_rng(N)
    theta = np.concatenate([rng.uniform(0, 360, 12), [0, 45, 90, 135, 180, 225, 270, 315]])
    p = 2 * N + 1
    d = float(N + 2)
    A = oracle.paralleltomo(N, theta, p, d).tocsr()
    for r, (x0, y0, a, b) in enumerate(par_rays(oracle, theta, p, d)):
        edge = (b == 0 and abs(y0 - N / 2) < 1e-15) or (a == 0 and abs(x0 - N / 2) < 1e-15)
        ref = {} if edge else exact_ray_cells(N, x0, y0, a, b)
        got = dict(zip(A.indices[A.indptr[r]:A.indptr[r + 1]].tolist(), A.data[A.indptr[r]:A.indptr[r + 1]]))
        # exact brute force keeps zero-length-free cells; the reference drops points within 1e-10
        ref = {c: v for c, v in ref.items() if v > 1e-9}
        got = {c: v for c, v in got.items() if v > 1e-9}
        assert set(got) == set(ref), (r, sorted(got), sorted(ref))
        for c in ref:
            assert abs(got[c] - ref[c]) <= 1e-13, (r, c, got[c], ref[c])


def test_parallel_rays_missing_grid_give_empty_rows(oracle):
    N = 8
    A = oracle.paralleltomo(N, [0, 37], 5, 40.0)      # spacing 10: only the central ray hits
    nnz = np.diff(A.indptr)
    assert nnz[0] == 0 and nnz[1] == 0 and nnz[2] > 0 and nnz[3] == 0 and nnz[4] == 0


def test_c1_size_and_nnz(oracle):
    A = oracle.paralleltomo(64, np.arange(0, 180.0), 91)
    assert A.shape == (16380, 4096)
    assert 9.0e5 < A.nnz < 9.8e5
    assert np.all(np.diff(A.indices[A.indptr[100]:A.indptr[101]]) > 0)


# ------------------------------------------------------------------ fanlineartomo
def fan_rays(oracle, N, theta, p, R, dw, sd):
    RN, dwN, sdN = R * N, dw * N, sd * N
    dew = dwN / p
    if p % 2:
        h = (p - 1) // 2
        om = [90 - (180 / math.pi) * math.atan(sdN / (k * dew)) for k in range(1, h + 1)]
        omega = [-v for v in om[::-1]] + [0.0] + om
    else:
        h = p // 2
        om = [90 - (180 / math.pi) * math.atan(sdN / ((k - 0.5) * dew)) for k in range(1, h + 1)]
        omega = [-v for v in om[::-1]] + om
    out = []
    for th in theta:
        st, ct = oracle.sind(th), oracle.cosd(th)
        x0t, y0t = ct * 0.0 + (-st) * RN, st * 0.0 + ct * RN
        abx, aby = (-x0t) / RN, (-y0t) / RN
        for w in omega:
            so, co = oracle.sind(w), oracle.cosd(w)
            out.append((x0t, y0t, co * abx + (-so) * aby, so * abx + co * aby))
    return out


@pytest.mark.parametrize("N,p", [(16, 23), (15, 22), (24, 34)])
def test_fan_row_sums_are_chord_lengths_from_source(oracle, N, p):
    theta = np.arange(0, 360, 17.0)
    A = oracle.fanlineartomo(N, theta, p, 2.0, 2.5, 3.0)
    rs = np.asarray(A.sum(axis=1)).ravel()
    # independent geometry: source on a circle of radius R*N, ray j aimed at the
    # detector element centre deposall_j on a line at distance sd*N from the source
    dew = 2.5 * N / p
    dep = (np.arange(1, (p - 1) // 2 + 1) * dew) if p % 2 else ((np.arange(1, p // 2 + 1) - 0.5) * dew)
    deposall = np.concatenate([-dep[::-1], [0.0], dep]) if p % 2 else np.concatenate([-dep[::-1], dep])
    r = 0
    for th in theta:
        t = math.radians(th)
        S = np.array([-math.sin(t) * 2 * N, math.cos(t) * 2 * N])
        u = -S / np.linalg.norm(S)                      # towards the centre
        v = np.array([-u[1], u[0]])                      # along the detector (+omega side)
        for dp in deposall:
            target = S + 3 * N * u + dp * v
            dvec = (target - S) / np.linalg.norm(target - S)
            ref = chord_length(S[0], S[1], dvec[0], dvec[1], N / 2)
            assert abs(rs[r] - ref) <= 1e-9 * max(1.0, ref), (r, rs[r], ref)
            r += 1


@pytest.mark.parametrize("N", [4, 7])
def test_fan_matches_exact_rational_bruteforce(oracle, N):
    theta = np.array([0.0, 33.0, 90.0, 151.0, 270.0, 300.0])
    p = 2 * N + 3
    A = oracle.fanlineartomo(N, theta, p, 2.0, 2.5, 3.0).tocsr()
    for r, (x0, y0, a, b) in enumerate(fan_rays(oracle, N, theta, p, 2.0, 2.5, 3.0)):
        ref = {c: v for c, v in exact_ray_cells(N, x0, y0, a, b).items() if v > 1e-9}
        got = dict(zip(A.indices[A.indptr[r]:A.indptr[r + 1]].tolist(), A.data[A.indptr[r]:A.indptr[r + 1]]))
        got = {c: v for c, v in got.items() if v > 1e-9}
        assert set(got) == set(ref), (r, sorted(got), sorted(ref))
        for c in ref:
            assert abs(got[c] - ref[c]) <= 1e-12


def test_fan_rejects_source_inside(oracle):
    with pytest.raises(ValueError):
        oracle.fanlineartomo(8, [0], 5, 0.5)


# ------------------------------------------------------------------ fancurvedtomo
def curved_rays(oracle, N, theta, p, R, d):
    RN = R * N
    omega = [d / 2] if p == 1 else [-d / 2 + (k * d) / (p - 1) for k in range(p)]
    out = []
    for th in theta:
        st, ct = oracle.sind(th), oracle.cosd(th)
        x0t, y0t = ct * 0.0 + (-st) * RN, st * 0.0 + ct * RN
        abx, aby = (-x0t) / RN, (-y0t) / RN
        for w in omega:
            so, co = oracle.sind(w), oracle.cosd(w)
            out.append((x0t, y0t, co * abx + (-so) * aby, so * abx + co * aby))
    return out


@pytest.mark.parametrize("N,p,R,d", [(16, 23, 2.0, None), (15, 22, 1.5, 40.0), (24, 7, 3.0, 10.0)])
def test_curved_row_sums_are_chord_lengths_at_equal_angles(oracle, N, p, R, d):
    theta = np.arange(0, 360, 17.0)
    A = oracle.fancurvedtomo(N, theta, p, R, d)
    d = 2 * math.degrees(math.atan(1 / (2 * R - 1))) if d is None else d
    rs = np.asarray(A.sum(axis=1)).ravel()
    # independent geometry: source at angle theta on a circle of radius R*N,
    # rays at equal angular steps over [-d/2, d/2] about the line to the centre
    r = 0
    for th in theta:
        t = math.radians(th)
        S = np.array([-math.sin(t) * R * N, math.cos(t) * R * N])
        phi0 = math.atan2(-S[1], -S[0])
        for w in np.linspace(-d / 2, d / 2, p):
            phi = phi0 + math.radians(w)
            ref = chord_length(S[0], S[1], math.cos(phi), math.sin(phi), N / 2)
            assert abs(rs[r] - ref) <= 1e-9 * max(1.0, ref), (r, rs[r], ref)
            r += 1


def test_curved_default_span_touches_the_top_corners(oracle):
    # fancurvedtomo.m:25-28: with the default span the first and last rays from
    # the source at (0, R N) hit the corners (-N/2, N/2) and (N/2, N/2), so they
    # carry (almost) nothing while their neighbours cross the image
    N, p = 32, 45
    A = oracle.fancurvedtomo(N, [0.0], p, 2.0)
    rs = np.asarray(A.sum(axis=1)).ravel()
    assert rs[0] < 1e-6 and rs[-1] < 1e-6
    assert rs[1] > 0.1 and rs[-2] > 0.1
    B = oracle.fancurvedtomo(N, [0.0], p, 2.0, 2 * math.degrees(math.atan(1 / 3)) * 1.02)
    assert np.asarray(B.sum(axis=1)).ravel()[0] == 0.0      # a wider fan misses the image


@pytest.mark.parametrize("N,p", [(13, 17), (12, 16), (13, 16)])
def test_curved_mirror_symmetry(oracle, N, p):
    # theta = 0: ray j and ray p-1-j are mirror images in x = 0 (column fx*N+r ->
    # (N-1-fx)*N+r); N and p are chosen so that no ray runs along the grid line x = 0
    A = oracle.fancurvedtomo(N, [0.0], p, 2.0).tocsr()
    for j in range(p):
        k = p - 1 - j
        a = dict(zip(A.indices[A.indptr[j]:A.indptr[j + 1]].tolist(), A.data[A.indptr[j]:A.indptr[j + 1]]))
        b = dict(zip(A.indices[A.indptr[k]:A.indptr[k + 1]].tolist(), A.data[A.indptr[k]:A.indptr[k + 1]]))
        a = {c: v for c, v in a.items() if v > 1e-9}
        b = {(N - 1 - c // N) * N + c % N: v for c, v in b.items() if v > 1e-9}
        assert set(a) == set(b)
        for c in a:
            assert abs(a[c] - b[c]) <= 1e-11


@pytest.mark.parametrize("N", [4, 7])
def test_curved_matches_exact_rational_bruteforce(oracle, N):
    theta = np.array([0.0, 33.0, 90.0, 151.0, 270.0, 300.0])
    p = 2 * N + 3
    A = oracle.fancurvedtomo(N, theta, p, 2.0, 35.0).tocsr()
    for r, (x0, y0, a, b) in enumerate(curved_rays(oracle, N, theta, p, 2.0, 35.0)):
        ref = {c: v for c, v in exact_ray_cells(N, x0, y0, a, b).items() if v > 1e-9}
        got = dict(zip(A.indices[A.indptr[r]:A.indptr[r + 1]].tolist(), A.data[A.indptr[r]:A.indptr[r + 1]]))
        got = {c: v for c, v in got.items() if v > 1e-9}
        assert set(got) == set(ref), (r, sorted(got), sorted(ref))
        for c in ref:
            assert abs(got[c] - ref[c]) <= 1e-12


def test_curved_argument_errors(oracle):
    with pytest.raises(ValueError, match="half squareroot"):
        oracle.fancurvedtomo(8, [0], 5, 0.5)
    with pytest.raises(ValueError, match="interval"):
        oracle.fancurvedtomo(8, [0], 5, 2.0, 190.0)
    with pytest.raises(ValueError, match="interval"):
        oracle.fancurvedtomo(8, [0], 5, 2.0, -1.0)


# ------------------------------------------------------------------ seismictomo
def seismic_rays(N, s, p):
    N2 = N / 2
    lin = lambda a, b, n: [b] if n == 1 else [a + (k * (b - a)) / (n - 1) for k in range(n)]
    y0 = lin(-N2 + (N / s) / 2, N2 - (N / s) / 2, s)
    p1, p2 = (p + 1) // 2, p // 2
    xp = [-N2] * p2 + (lin(-N2 + (N / p1) / 2, N2 - (N / p1) / 2, p1) if p1 else [])
    yp = (lin(-N2 + (N / p2) / 2, N2 - (N / p2) / 2, p2) if p2 else []) + [N2] * p1
    return [(N2, y, xr - N2, yr - y) for y in y0 for xr, yr in zip(xp, yp)]


@pytest.mark.parametrize("N,s,p", [(16, 16, 32), (15, 7, 9), (24, 5, 1), (8, 1, 6)])
def test_seismic_row_sums_are_source_receiver_distances(oracle, N, s, p):
    # both end points lie on the boundary, so the chord is the whole segment
    A = oracle.seismictomo(N, s, p)
    assert A.shape == (s * p, N * N)
    rs = np.asarray(A.sum(axis=1)).ravel()
    for r, (x0, y0, a, b) in enumerate(seismic_rays(N, s, p)):
        assert abs(rs[r] - math.hypot(a, b)) <= 1e-9 * math.hypot(a, b), (r, rs[r], math.hypot(a, b))


@pytest.mark.parametrize("N", [4, 7])
def test_seismic_matches_exact_rational_bruteforce(oracle, N):
    s, p = 3, 5
    A = oracle.seismictomo(N, s, p).tocsr()
    for r, (x0, y0, a, b) in enumerate(seismic_rays(N, s, p)):
        ref = {c: v for c, v in exact_ray_cells(N, x0, y0, a, b).items() if v > 1e-9}
        got = dict(zip(A.indices[A.indptr[r]:A.indptr[r + 1]].tolist(), A.data[A.indptr[r]:A.indptr[r + 1]]))
        got = {c: v for c, v in got.items() if v > 1e-9}
        assert set(got) == set(ref), (r, sorted(got), sorted(ref))
        for c in ref:
            assert abs(got[c] - ref[c]) <= 1e-12


def test_seismic_horizontal_ray_stays_in_one_image_row(oracle):
    # N = 8, s = 8, p = 16: source 4 sits at y = -0.5 and left receiver 4 at
    # y = -0.5, so that ray runs along image row N/2 + 1 (1-based, from the top)
    N = 8
    A = oracle.seismictomo(N, 8, 16).tocsr()
    r = 3 * 16 + 3
    cols = A.indices[A.indptr[r]:A.indptr[r + 1]]
    assert len(cols) == N and np.all(cols % N == N // 2)
    np.testing.assert_array_equal(A.data[A.indptr[r]:A.indptr[r + 1]], np.ones(N))


@pytest.mark.parametrize("N", [7, 10, 20, 64, 100])
def test_tectonic_phantom_matches_oracle_restatement(oracle, N):
    import airtoolsii_amd.phantoms as P
    a, b = P.tectonic(N), oracle.tectonic(N)
    assert np.array_equal(a, b)
    assert set(np.unique(a)) <= {0.0, 0.75, 1.0}


def test_tectonic_phantom_shape_n64(oracle):
    # phantomgallery.m:425-453 at N = 64: N5 = 13, N13 = 5, N7 = 9, N20 = 3
    x = oracle.tectonic(64)
    assert np.all(x[12:22, 24:] >= 0.75)            # right plate, rows 13..22 from column 25
    assert np.all(x[11, 25:] == 0.75) and x[11, 24] == 0 and np.all(x[10, 27:] == 0.75) and x[10, 26] == 0
    assert np.all(x[12:26, :25] == 1.0)             # left plate, rows 13..26, columns 1..25
    assert np.all(x[12 + 18:26 + 18, 59] == 1.0)    # column 60: shifted by the 18 odd j in 25..60
    assert x[12 + 17, 59] == 0.0 and x[12 + 18, 60] == 0.0


# ------------------------------------------------------------------ sphericaltomo
def arc_inside(cx, cy, rho, lo, hi, samples=200000):
    """Length of the circle (cx,cy,rho) inside [lo,hi)^2 by dense angular sampling."""
    phi = (np.arange(samples) + 0.5) * (2 * np.pi / samples)
    x, y = cx + rho * np.cos(phi), cy + rho * np.sin(phi)
    inside = (x >= lo) & (x < hi) & (y >= lo) & (y < hi)
    return 2 * np.pi * rho * inside.mean()


@pytest.mark.parametrize("N,nC", [(16, 23), (17, 10), (32, 45)])
def test_spherical_row_sums_are_arc_lengths_inside_the_image(oracle, N, nC):
    angles = np.arange(0.0, 360.0, 37.0)
    A = oracle.sphericaltomo(N, angles, nC)
    assert A.shape == (len(angles) * nC, N * N)
    rs = np.asarray(A.sum(axis=1)).ravel()
    # geometry in pixel units: centres on a circle of radius N/sqrt(2) about
    # pixel ceil(N/2), circle n of radius 2n/nC * N/sqrt(2); a sample lands in
    # the image when its rounded pixel is in 1..N, i.e. inside [0.5, N+0.5)^2
    scale = N / math.sqrt(2)
    c0 = math.ceil(N / 2)
    r = 0
    for a in angles:
        cx, cy = c0 + scale * math.cos(math.radians(a)), c0 + scale * math.sin(math.radians(a))
        for n in range(1, nC + 1):
            rho = 2 * n / nC * scale
            nphi = math.ceil(4 * math.pi * scale * (2 * n / nC))
            base = 2 * math.pi * (2 * n / nC) / nphi
            ref = arc_inside(cx, cy, rho, 0.5, N + 0.5) / scale
            # up to 8 boundary crossings, each off by at most one sample, plus the repeated sample at 2 pi
            assert abs(rs[r] - ref) <= 9 * base + 1e-3 * ref, (a, n, rs[r], ref)
            r += 1


@pytest.mark.parametrize("N,nC", [(8, 11), (9, 6), (16, 23)])
def test_spherical_matches_vectorised_sampling(oracle, N, nC):
    angles = np.array([0.0, 45.0, 90.0, 133.0, 180.0, 270.0, 300.5])
    A = oracle.sphericaltomo(N, angles, nC).tocsr()
    radii = np.linspace(0, 2, nC + 1)[1:]
    dx = np.sqrt(2) / N
    for m, a in enumerate(angles):
        xi = np.array([oracle.cosd(a), oracle.sind(a)])
        for n, rad in enumerate(radii):
            nphi = int(np.ceil((4 * np.pi / dx) * rad))
            k = np.arange(nphi + 1) * (2 * np.pi / nphi)
            xx = (xi[0] + rad * np.cos(k)) / dx + np.ceil(N / 2)
            yy = (xi[1] + rad * np.sin(k)) / dx + np.ceil(N / 2)
            col = np.sign(xx) * np.floor(np.abs(xx) + 0.5)
            row = np.sign(yy) * np.floor(np.abs(yy) + 0.5)
            keep = (col > 0) & (col <= N) & (row > 0) & (row <= N)
            J = ((N - row[keep] + 1) + (col[keep] - 1) * N).astype(np.int64) - 1
            u, c = np.unique(J, return_counts=True)
            i = m * nC + n
            np.testing.assert_array_equal(A.indices[A.indptr[i]:A.indptr[i + 1]], u)
            np.testing.assert_allclose(A.data[A.indptr[i]:A.indptr[i + 1]], 2 * np.pi * rad / nphi * c, rtol=1e-15)


# ------------------------------------------------------------------ seismicwavetomo
def fresnel_dense(N, s, p, omega):
    """The sensitivity kernels of seismicwavetomo.m:188-212 with numpy arrays,
    laid out as the reference's meshgrid (rows top to bottom, S(:) column-major)."""
    N2 = N / 2
    rays = seismic_rays(N, s, p)
    xr = np.arange(-N2 + .5, N2, 1.0)
    yr = np.arange(N2 - .5, -N2, -1.0)
    xx, yy = np.meshgrid(xr, yr)
    out = []
    for x0, y0, a, b in rays:
        R = np.array([x0 + a, y0 + b]); S = np.array([x0, y0])
        tSX = np.sqrt((xx - S[0]) ** 2 + (yy - S[1]) ** 2)
        tRX = np.sqrt((xx - R[0]) ** 2 + (yy - R[1]) ** 2)
        dist = np.sqrt(np.sum((S - R) ** 2))
        dt = tSX + tRX - dist
        K = np.cos(2 * np.pi * dt * (omega / N)) * np.exp(-(10 * dt * (omega / N)) ** 2)
        out.append((dist * K / K.sum()).ravel(order="F"))
    return np.array(out)


@pytest.mark.parametrize("N,s,p,omega", [(12, 12, 24, 10.0), (9, 4, 7, 5.0)])
def test_seismicwave_matches_dense_kernels(oracle, N, s, p, omega):
    A = oracle.seismicwavetomo(N, s, p, omega).toarray()
    D = fresnel_dense(N, s, p, omega)
    keep = D >= 1e-6
    assert np.array_equal(A != 0, keep)
    np.testing.assert_allclose(A[keep], D[keep], rtol=1e-12)


def test_seismicwave_peak_on_the_source_receiver_line(oracle):
    # the kernel is 1 where delta = 0, i.e. on the segment from S to R; the
    # largest entry of each row sits at a pixel closest to that segment
    N, s, p = 16, 4, 8
    A = oracle.seismicwavetomo(N, s, p).toarray()
    for r, (x0, y0, a, b) in enumerate(seismic_rays(N, s, p)):
        q = int(np.argmax(A[r]))
        c, rr = divmod(q, N)
        px, py = c - N / 2 + 0.5, N / 2 - 0.5 - rr
        t = np.clip(((px - x0) * a + (py - y0) * b) / (a * a + b * b), 0, 1)
        d = math.hypot(px - (x0 + t * a), py - (y0 + t * b))
        assert d <= math.sqrt(0.5) + 1e-12, (r, d)


# ------------------------------------------------------------------ weights
def small_problem(oracle, N=16, p=23, d=30.0):
    return oracle.paralleltomo(N, np.arange(0, 180, 7.0), p, d)


def test_weights_closed_forms(oracle):
    A = small_problem(oracle)
    m, n = A.shape
    A2 = A.multiply(A)
    rn2 = np.asarray(A2.sum(1)).ravel()
    rab = np.asarray(abs(A).sum(1)).ravel()
    cab = np.asarray(abs(A).sum(0)).ravel()
    cnz = np.diff(A.tocsc().indptr).astype(float)
    with np.errstate(divide="ignore"):
        exp = {
            "landweber": (np.ones(m), np.ones(n)),
            "cimmino": ((1 / m) * (1 / rn2), np.ones(n)),
            "cav": (1 / (A2 @ cnz), np.ones(n)),
            "drop": (1 / rn2, 1 / cnz),
            "sart": (1 / rab, 1 / cab),
        }
    assert (rn2 == 0).any(), "problem should contain rays that miss the grid"
    for meth, (Me, De) in exp.items():
        Me, De = np.where(np.isinf(Me), 0, Me), np.where(np.isinf(De), 0, De)
        M, D = oracle.weights(meth, A)
        np.testing.assert_allclose(M, Me, rtol=1e-13, atol=0)
        np.testing.assert_allclose(D, De, rtol=1e-13, atol=0)


# ------------------------------------------------------------------ calczeta / relaxpar
def test_calczeta_roots(oracle):
    k = np.arange(2, 200)
    z = oracle.calczeta(k)
    assert abs(z[0] - 1 / 3) < 1e-15
    assert np.all((z > 0) & (z < 1)) and np.all(np.diff(z) > 0)
    for kk, zz in zip(k, z):
        res = (2 * kk - 1) * zz ** (kk - 1) - sum(zz ** i for i in range(kk - 1))
        assert abs(res) < 1e-10 * (2 * kk)
    with pytest.raises(ValueError):
        oracle.calczeta([])


def test_psi_vectors_in_stable_range(oracle):
    rho = 3.0
    for s in ("psi1", "psi1mod", "psi2", "psi2mod"):
        casel, lam, _ = oracle.relaxpar_vector(s, rho, 30)
        assert casel == 3 and lam.shape == (30,)
        assert np.all(lam > 0) and np.allclose(lam[:2], math.sqrt(2) / rho)
    _, l1, _ = oracle.relaxpar_vector("psi1", rho, 30)
    _, l1m, _ = oracle.relaxpar_vector("psi1mod", rho, 30)
    np.testing.assert_allclose(l1m[2:], 2 * l1[2:], rtol=1e-15)
    assert np.all(l1[2:] * rho < 2)


# ------------------------------------------------------------------ SIRT loop
@pytest.fixture(scope="module")
def prob(oracle):
    A = oracle.paralleltomo(24, np.arange(0, 180, 4.0), 34)
    x = oracle.phantomgallery("shepplogan", 24).ravel(order="F")
    b = A @ x
    e = np.random.default_rng(5).standard_normal(b.size)
    b = b + 0.01 * np.linalg.norm(b) * e / np.linalg.norm(e)
    return A, x, b


def test_landweber_equals_identity_custom_method(oracle, prob):
    A, x, b = prob
    X1, i1, _ = oracle.sirt("landweber", A, b, 20)
    X2, i2, _ = oracle.sirt({}, A, b, 20)
    assert np.array_equal(X1, X2) and i1["rho"] == i2["rho"]


def test_cimmino_equals_scaled_custom_M_under_default_relaxpar(oracle, prob):
    A, x, b = prob
    rn2 = np.asarray(A.multiply(A).sum(1)).ravel()
    Mc = np.where(rn2 > 0, 1 / np.where(rn2 > 0, rn2, 1), 0)
    X1, i1, _ = oracle.sirt("cimmino", A, b, 30)
    X2, i2, _ = oracle.sirt({"M": Mc}, A, b, 30)
    np.testing.assert_allclose(i2["rho"], i1["rho"] * A.shape[0], rtol=1e-9)
    np.testing.assert_allclose(X1, X2, rtol=1e-8, atol=1e-12)


def test_diagonal_system_closed_form(oracle):
    n = 50
    a = np.linspace(0.5, 2.0, n)
    xs = np.cos(np.arange(n))
    A = sp.diags(a).tocsr()
    b = a * xs
    lam = 0.3
    K = [1, 2, 5, 17]
    X, info, _ = oracle.sirt("landweber", A, b, K, None, {"relaxpar": lam})
    for col, k in enumerate(K):
        np.testing.assert_allclose(X[:, col], xs * (1 - (1 - lam * a * a) ** k), rtol=1e-12, atol=1e-14)
    # Cimmino on a diagonal system: M = 1/(n a^2) -> x_k = x*(1 - (1 - lam/n)^k)
    X, _, _ = oracle.sirt("cimmino", A, b, K, None, {"relaxpar": 40.0})
    for col, k in enumerate(K):
        np.testing.assert_allclose(X[:, col], xs * (1 - (1 - 40.0 / n) ** k), rtol=1e-12, atol=1e-14)
    # SART: M = 1/a, D = 1/a -> x_k = x*(1 - (1 - lam)^k), rho = 1
    X, info, _ = oracle.sirt("sart", A, b, K, None, {"relaxpar": 0.7})
    assert info["rho"] == 1.0
    for col, k in enumerate(K):
        np.testing.assert_allclose(X[:, col], xs * (1 - 0.3 ** k), rtol=1e-12, atol=1e-14)


def test_line_search_solves_scaled_identity_in_one_step(oracle):
    n = 30
    A = (2.5 * sp.identity(n)).tocsr()
    xs = np.sin(np.arange(n) + 1.0)
    X, info, _ = oracle.sirt("landweber", A, 2.5 * xs, 1, None, {"relaxpar": "line"})
    np.testing.assert_allclose(X[:, 0], xs, rtol=1e-14)
    np.testing.assert_allclose(info["relaxpar"], 1 / 2.5 ** 2, rtol=1e-14)


def test_default_rho_matches_power_iteration(oracle, prob):
    A, x, b = prob
    M, D = oracle.weights("drop", A)
    rho = oracle.spectral_radius(A, M, D)
    v = np.ones(A.shape[1])
    for _ in range(3000):
        w = D * (A.T @ (M * (A @ v)))
        lam = np.linalg.norm(w) / np.linalg.norm(v)
        v = w / np.linalg.norm(w)
    assert abs(rho - lam) < 1e-6 * rho


def test_bounds_projection(oracle, prob):
    A, x, b = prob
    X, _, _ = oracle.sirt("sart", A, b, 15, None, {"lbound": 0, "ubound": 1})
    assert X.min() >= 0 and X.max() <= 1
    n = A.shape[1]
    lb = np.full(n, -np.inf); lb[::3] = 0.2
    ub = np.full(n, np.inf); ub[1::3] = 0.1
    X, _, _ = oracle.sirt("sart", A, b, 15, None, {"lbound": lb, "ubound": ub})
    assert np.all(X[::3, 0] >= 0.2) and np.all(X[1::3, 0] <= 0.1)


def test_save_list_semantics(oracle, prob):
    A, x, b = prob
    Xall, _, _ = oracle.sirt("cimmino", A, b, np.arange(1, 21), None, {"rho": 0.01})
    Xk, info, _ = oracle.sirt("cimmino", A, b, [3, 7, 20], None, {"rho": 0.01})
    assert np.array_equal(Xk, Xall[:, [2, 6, 19]])
    assert list(info["itersaved"]) == [3, 7, 20] and info["finaliter"] == 20 and info["stoprule"] == 0


def test_discrepancy_principle_stop(oracle, prob):
    A, x, b = prob
    Xall, _, _, rn = oracle.sirt("sart", A, b, np.arange(1, 61), None, {}, return_resnorm=True)
    tau = rn[25] * 1.0000001                            # first k with ||r_k|| <= tau
    kstop = int(np.argmax(rn <= tau))
    X, info, _ = oracle.sirt("sart", A, b, [5, 10, 60], None, {"stoprule": {"type": "DP", "taudelta": tau}})
    assert info["stoprule"] == 2 and info["finaliter"] == kstop
    assert list(info["itersaved"]) == [k for k in (5, 10) if k < kstop] + [kstop]
    assert np.array_equal(X[:, -1], Xall[:, kstop - 1])
    # DP satisfied at x0: no iteration, empty X, itersaved = [0]
    X, info, _ = oracle.sirt("sart", A, b, 10, None, {"stoprule": {"type": "DP", "taudelta": 1e9}})
    assert X.shape[1] == 0 and info["finaliter"] == 0 and list(info["itersaved"]) == [0]


def test_monotone_error_rule(oracle, prob):
    A, x, b = prob
    kmax = 80
    Xall, _, _ = oracle.sirt("cimmino", A, b, np.arange(1, kmax + 1), None, {"rho": 0.02})
    R = [b - A @ np.zeros(A.shape[1])] + [b - A @ Xall[:, k] for k in range(kmax)]
    crit = [(R[k - 1] @ (R[k - 1] + R[k])) / 2 / np.linalg.norm(R[k - 1]) for k in range(1, kmax + 1)]
    tau = crit[40] * 1.000001
    kstop = 1 + int(np.argmax(np.array(crit) <= tau))
    X, info, _ = oracle.sirt("cimmino", A, b, kmax, None, {"rho": 0.02, "stoprule": {"type": "ME", "taudelta": tau}})
    assert info["stoprule"] == 3 and info["finaliter"] == kstop
    np.testing.assert_array_equal(X[:, -1], Xall[:, kstop - 1])


def test_train_relaxpar_on_synthetic_error_model(oracle):
    # Error curve e_k(lam) = 1 + 0.5*max(k* - k, 0)^2 + 0.001*max(k - k*, 0) with
    # k* = ceil(10/lam): every lam reaches the same floor, larger lam sooner, so the
    # search must end near the top of (0, 2/rho).  Warm starts continue the count.
    rho = 1.0
    state = {"k_end": 0}

    def run(A, b, K, x0, opts):
        lam = opts.get("relaxpar", 1.9 / rho)
        k0 = 0 if x0 is None else state["k_end"]
        ks = k0 + np.asarray(K)
        state["k_end"] = int(ks[-1])
        kstar = math.ceil(10 / lam)
        err = 1 + 0.5 * np.maximum(kstar - ks, 0) ** 2 + 0.001 * np.maximum(ks - kstar, 0)
        return err[None, :], {"rho": rho}

    lam = oracle.train_relaxpar(None, None, np.zeros(1), "sart", 100, None, run=run)
    assert 0.9 * 2 / rho <= lam < 2 / rho


def test_train_relaxpar_row_action_branch(oracle):
    # ART/CART branch (train_relaxpar.m:90-96,213-268): the first runs use
    # relaxpar 0.25, relaxparmax = 2, and the search tests the upper probe
    # first.  Error floor 1 + 10 (lam - 0.5)^2, reached at k* = ceil(10/lam):
    # meub = 1.02 * floor(0.25) admits lam in about [0.24, 0.76]; both first
    # probes (0.76, 1.24) lie above it, so the search moves left and settles at
    # the largest admissible lam (fewest iterations), near 0.76.
    state = {"k_end": 0}

    def run(A, b, K, x0, opts):
        lam = opts["relaxpar"]
        k0 = 0 if x0 is None else state["k_end"]
        ks = k0 + np.asarray(K)
        state["k_end"] = int(ks[-1])
        kstar = math.ceil(10 / lam)
        floor = 1 + 10 * (lam - 0.5) ** 2
        err = floor + 0.5 * np.maximum(kstar - ks, 0) ** 2 + 0.001 * np.maximum(ks - kstar, 0)
        return err[None, :], {}

    lam = oracle.train_relaxpar(None, None, np.zeros(1), "kaczmarz", 100, None, run=run)
    hi = 0.5 + math.sqrt((1.02 * (1 + 10 * 0.25 ** 2) - 1) / 10)
    assert hi - 0.03 <= lam <= hi + 0.02, (lam, hi)


def test_train_dpme_on_prescribed_iterates(oracle):
    # A = I and iterates x_k = x* + (k - 7)^2/100 * u (u a unit vector): the error
    # is smallest at k = 7, found in the first chunk of 10.  X keeps the start
    # vector as column 1, so the DP residual is that of iterate kopt - 1 = 6 and
    # ME also uses iterate 5 (train_dpme.m:82,99,120,126).
    n = 5
    xs = np.arange(1.0, n + 1)
    u = np.zeros(n); u[2] = 1.0
    state = {"k_end": 0}

    def run(A, b, K, x0, opts):
        k0 = 0 if not np.any(x0) else state["k_end"]
        ks = k0 + np.asarray(K)
        state["k_end"] = int(ks[-1])
        return np.stack([xs + (k - 7) ** 2 / 100 * u for k in ks], axis=1), {}

    noise = np.ones((n, 2))
    delta = 0.5
    A = sp.identity(n, format="csr")
    tau = oracle.train_dpme(A, xs, xs, "sart", "DP", delta, 2, 30, None, noise, run)
    e = delta * np.ones(n) / math.sqrt(n)
    r6 = xs + e - (xs + 1 / 100 * u)
    assert abs(tau - np.linalg.norm(r6) / delta) < 1e-14
    tau_me = oracle.train_dpme(A, xs, xs, "sart", "ME", delta, 2, 30, None, noise, run)
    r5 = xs + e - (xs + 4 / 100 * u)
    assert abs(tau_me - 0.5 * (r5 @ (r5 + r6)) / (delta * np.linalg.norm(r5))) < 1e-14


def test_train_dpme_continues_past_a_chunk_and_caps_at_kmax(oracle):
    # error falling until k = 23: two more chunks of 10 after the first; with
    # kmax = 15 the search ends at the cap (kopt = 15, iterate 14 used)
    n = 3
    xs = np.zeros(n)
    state = {"k_end": 0}

    def run(A, b, K, x0, opts):
        k0 = 0 if not np.any(x0) else state["k_end"]
        ks = k0 + np.asarray(K)
        state["k_end"] = int(ks[-1])
        return np.stack([np.full(n, abs(k - 23) / 10 + 0.01) for k in ks], axis=1), {}

    A = sp.identity(n, format="csr")
    b = np.zeros(n)
    noise = np.ones((n, 1))
    tau = oracle.train_dpme(A, b, xs, "sart", "DP", 1.0, 1, 100, None, noise, run)
    x22 = np.full(n, 1 / 10 + 0.01)
    assert abs(tau - np.linalg.norm(np.ones(n) / math.sqrt(n) - x22)) < 1e-14
    tau15 = oracle.train_dpme(A, b, xs, "sart", "DP", 1.0, 1, 15, None, noise, run)
    x14 = np.full(n, 9 / 10 + 0.01)
    assert abs(tau15 - np.linalg.norm(np.ones(n) / math.sqrt(n) - x14)) < 1e-14


def test_fbp_reconstructs_a_disk(oracle):
    # the filtered back projection inverts the Radon transform: a disk of
    # radius N/4 and value 1 comes back near 1 inside and near 0 outside
    N = 64
    theta = np.arange(0.0, 180.0)
    A = oracle.paralleltomo(N, theta, 91)
    c = (np.arange(N) - (N - 1) / 2)
    disk = ((c[:, None] ** 2 + c[None, :] ** 2) <= (N / 4) ** 2).astype(float)
    b = A @ disk.ravel(order="F")
    for filt in ("ram-lak", "shepp-logan", "cosine", "hamming", "hann"):
        x = oracle.fbp(A, b, theta, filt).reshape((N, N), order="F")
        inner = (c[:, None] ** 2 + c[None, :] ** 2) <= (N / 4 - 3) ** 2
        outer = (c[:, None] ** 2 + c[None, :] ** 2) >= (N / 4 + 3) ** 2
        assert abs(np.median(x[inner]) - 1) < 0.05, filt
        assert np.max(np.abs(x[outer])) < 0.1, filt
    np.testing.assert_allclose(oracle.fbp(A, b, theta, "none"), A.T @ b)
    with pytest.raises(ValueError):
        oracle.fbp(A, b, theta, "gaussian")


# ------------------------------------------------------------------ ART / CART
def test_splitmix64_known_answers(oracle):
    # the first outputs of SplitMix64 from state 0 (state advanced by the golden gamma)
    g = 0x9E3779B97F4A7C15
    L = oracle.lib()
    assert L.ora_splitmix64(0) == 0xE220A8397B1DCDAF
    assert L.ora_splitmix64(g) == 0x6E789E6AA1B965F4
    assert L.ora_splitmix64((2 * g) % 2 ** 64) == 0x06C45D188009454F


def _small_system(seed=1, m=30, n=20):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=0.3, random_state=seed, format="csr") + sp.eye(m, n, format="csr")
    x = rng.uniform(0, 1, n)
    return A.tocsr(), x, A @ x


def test_kaczmarz_step_projects_onto_the_row_hyperplane(oracle):
    A, x, b = _small_system()
    for i in (0, 7, 29):
        X, _ = oracle.art([i + 1], A, b + 0.3, 1, None, {"relaxpar": 1.0})
        assert abs(A[i] @ X[:, 0] - (b[i] + 0.3)) < 1e-13


def test_kaczmarz_converges_on_a_consistent_system(oracle):
    A, x, b = _small_system()
    for meth in ("kaczmarz", "randkaczmarz"):
        X, info = oracle.art(meth, A, b, 300, None, {"seed": 3})
        assert np.linalg.norm(X[:, -1] - x) < 1e-9 * np.linalg.norm(x), meth
    X, info = oracle.art("symkaczmarz", A, b, 300)
    assert info["finaliter"] == 300 and np.linalg.norm(X[:, -1] - x) < 1e-9


def test_symkaczmarz_is_kaczmarz_over_the_mirrored_order(oracle):
    A, x, b = _small_system()
    Xs, _ = oracle.art("symkaczmarz", A, b, [4, 8])
    order = list(range(1, 31)) + list(range(30, 0, -1))
    Xc, _ = oracle.art(order, A, b, [2, 4])
    np.testing.assert_array_equal(Xs, Xc)


def test_kaczmarz_bounds_match_a_literal_loop(oracle):
    A, x, b = _small_system(4)
    b = b + np.random.default_rng(5).standard_normal(b.size) * 0.1
    opts = {"lbound": 0.2, "ubound": np.full(20, 0.8), "relaxpar": 0.7, "damp": 0.1}
    X, _ = oracle.art("kaczmarz", A, b, 3, np.full(20, 2.0), opts)
    xx = np.full(20, 2.0)
    D = A.toarray()
    nrm = (D * D).sum(1)
    nrm = nrm + 0.1 * nrm.max()
    for _ in range(3):
        for i in range(30):
            xx = xx + (0.7 * (b[i] - D[i] @ xx) / nrm[i]) * D[i]
            xx = np.minimum(np.maximum(xx, 0.2), 0.8)
    np.testing.assert_allclose(X[:, -1], xx, rtol=1e-13, atol=1e-14)


def test_randkaczmarz_sweeps_are_permutations(oracle):
    A, x, b = _small_system()
    L = oracle.lib()
    for k in (1, 2, 3):
        perm = list(range(30))
        for t in range(29, 0, -1):
            j = int(L.ora_rng_u01(9, 1, k, t) * (t + 1))
            perm[t], perm[j] = perm[j], perm[t]
        assert sorted(perm) == list(range(30)) and perm != list(range(30))
    X1, _ = oracle.art("randkaczmarz", A, b, 2, None, {"seed": 9})
    X2, _ = oracle.art("randkaczmarz", A, b, 2, None, {"seed": 10})
    assert not np.array_equal(X1, X2)


def test_cart_column_step_zeroes_the_column_gradient(oracle):
    # with relaxpar 1 and no bounds a column step is exact coordinate descent:
    # afterwards a_j' r = 0, and the residual kept by the sweep is b - A x
    A, x, b = _small_system()
    Ac = A.tocsc()
    for j in (0, 5, 19):
        X, _ = oracle.cart([j + 1], A, b, 1, None, {"relaxpar": 1.0})
        r = b - A @ X[:, 0]
        assert abs(Ac[:, j].toarray().ravel() @ r) < 1e-12
    X, info = oracle.cart("columnaction", A, b, 200, None, {"relaxpar": 1.0, "Kbegin": 10 ** 6})
    assert np.linalg.norm(X[:, -1] - x) < 1e-9


def test_cart_flagging_freezes_converged_columns(oracle):
    # a huge THR flags every column at sweep Kbegin; flagged columns are skipped
    # until their unflag draw succeeds, so x stays put in the sweep after
    A, x, b = _small_system()
    X, info = oracle.cart("columnaction", A, b, [3, 4], None,
                          {"relaxpar": 1.0, "Kbegin": 3, "THR": 1e9, "Nunflag": 50})
    assert not info["flags"].any() or info["flags"].sum() < 5
    np.testing.assert_array_equal(X[:, 0], X[:, 1])


def test_art_stop_rules(oracle):
    A, x, b = _small_system()
    X, info = oracle.art("kaczmarz", A, b, 100, None, {"stoprule": {"type": "DP", "taudelta": 1e-3}})
    r = [np.linalg.norm(b - A @ oracle.art("kaczmarz", A, b, k)[0][:, 0]) for k in range(1, info["finaliter"] + 1)]
    assert info["stoprule"] == 2 and r[-1] <= 1e-3 < r[-2]
    X, info = oracle.cart("columnaction", A, b, 500, None, {"stoprule": {"type": "ME", "taudelta": 1e-2}})
    assert info["stoprule"] == 3 and info["finaliter"] < 500


# ------------------------------------------------------------------ phantoms
def test_phantoms_basic_properties(oracle):
    sl = oracle.phantomgallery("shepplogan", 64)
    assert sl.min() == 0 and abs(sl.max() - 1) < 1e-12
    assert sl[32, 32] > 0 and sl[0, 0] == 0
    g = oracle.phantomgallery("grains", 64, None, 1)
    labels = np.unique(g)
    assert g.max() == 1 and len(labels) <= round(3 * math.sqrt(64))
    t = oracle.phantomgallery("threephases", 64, None, 2)
    assert set(np.unique(t)).issubset({0.0, 0.5, 1.0})
    s = oracle.phantomgallery("smooth", 32)
    assert abs(s.max() - 1) < 1e-15 and s.min() > 0


def test_product_phantoms_match_oracle(oracle):
    from airtoolsii_amd import phantomgallery
    for N in (48, 65):
        for name, args in (("shepplogan", ()), ("smooth", (None,)), ("grains", (None, 1)), ("grains", (30, 7)),
                           ("threephases", (None, 3)), ("threephases", (40, 9))):
            a = phantomgallery(name, N, *args)
            b = oracle.phantomgallery(name, N, *args)
            if name == "smooth":
                np.testing.assert_allclose(a, b, rtol=1e-13, atol=1e-15)
            else:
                assert np.array_equal(a, b), (name, N, args)


# Shepp-Logan pixels summed by hand (phantomgallery.m:178-214).  The pixel
# nearest to a point (x, y) of [-1,1]^2 is column round((x+1)(N-1)/2) and row
# N-1-round((y+1)(N-1)/2): rot90 puts y = +1 on the top row.  Each point lies
# well inside (or outside) every ellipse, so one pixel either way does not
# change its sum; the table gives the ellipses that contain it:
#   e1  A=1    .69 x .92  at (0, 0)
#   e2  A=-.8  .6624 x .874  at (0, -.0184)
#   e3  A=-.2  .11 x .31  at (.22, 0), phi = -18 deg
#   e4  A=-.2  .16 x .41  at (-.22, 0), phi = 18 deg
#   e5  A=.1   .21 x .25  at (0, .35)
#   e6/e7 A=.1 .046 circles at (0, +-.1)
#   e8  A=.1   .046 x .023 at (-.08, -.605); e9 .023 circle at (0, -.606); e10 .023 x .046 at (.06, -.605)
# For the tilted ellipses the point sits 0.85 of the semi-major axis out along
# the tilted axis (direction (-sin phi, cos phi)), which lies outside the
# ellipse tilted the other way.
def _sl_points():
    d3 = np.array([-math.sin(math.radians(-18)), math.cos(math.radians(-18))])
    d4 = np.array([-math.sin(math.radians(18)), math.cos(math.radians(18))])
    return [
        ((0.0, 0.89), 1.0),                     # e1 only (above e2's top at .8556)
        ((0.0, -0.3), 1.0 - 0.8),               # e1 + e2
        ((-0.08, -0.605), 1.0 - 0.8 + 0.1),     # e8 (bottom left small ellipse)
        ((0.0, -0.606), 1.0 - 0.8 + 0.1),       # e9
        ((0.06, -0.605), 1.0 - 0.8 + 0.1),      # e10
        ((0.0, 0.45), 1.0 - 0.8 + 0.1),         # e5 (above e6)
        ((0.0, 0.13), 1.0 - 0.8 + 0.1 + 0.1),   # e5 (y from .1 to .6) and e6
        ((0.0, -0.1), 1.0 - 0.8 + 0.1),         # e7 (below e5)
        (tuple(np.array([0.22, 0.0]) + 0.85 * 0.31 * d3), 0.0),    # e3 tip: 1 - .8 - .2 -> clamped to 0
        (tuple(np.array([-0.22, 0.0]) + 0.85 * 0.41 * d4), 0.0),   # e4 tip
        ((0.0, -0.95), 0.0),                    # outside the head
        ((0.7, 0.7), 0.0),
    ]


def test_shepplogan_hand_summed_pixels(oracle):
    from airtoolsii_amd import phantomgallery
    N = 511
    im = oracle.phantomgallery("shepplogan", N)
    for (x, y), val in _sl_points():
        c = int(round((x + 1) * (N - 1) / 2))
        r = N - 1 - int(round((y + 1) * (N - 1) / 2))
        assert abs(im[r, c] - val) < 1e-12, ((x, y), r, c, im[r, c], val)
    # the head is not symmetric: an upside-down or mirrored image differs
    assert not np.array_equal(im, np.flipud(im)) and not np.array_equal(im, np.fliplr(im))
    assert np.array_equal(phantomgallery("shepplogan", N), im)


def test_shepplogan_against_ellipse_membership(oracle):
    """Every pixel against a direct evaluation of the ten quadratic forms on the
    pixel's (x, y) (x = -1 + 2c/(N-1), y = 1 - 2r/(N-1)), skipping pixels whose
    form is within 1e-9 of an ellipse boundary."""
    N = 96
    im = oracle.phantomgallery("shepplogan", N)
    E = [(1, .69, .92, 0, 0, 0), (-.8, .6624, .8740, 0, -.0184, 0), (-.2, .11, .31, .22, 0, -18),
         (-.2, .16, .41, -.22, 0, 18), (.1, .21, .25, 0, .35, 0), (.1, .046, .046, 0, .1, 0),
         (.1, .046, .046, 0, -.1, 0), (.1, .046, .023, -.08, -.605, 0), (.1, .023, .023, 0, -.606, 0),
         (.1, .023, .046, .06, -.605, 0)]
    c, r = np.meshgrid(np.arange(N), np.arange(N))
    X = -1 + 2.0 * c / (N - 1)
    Y = 1 - 2.0 * r / (N - 1)
    ref = np.zeros((N, N))
    near = np.zeros((N, N), bool)
    for A, a, b, x0, y0, deg in E:
        t = math.radians(deg)
        R = np.array([[math.cos(t), math.sin(t)], [-math.sin(t), math.cos(t)]])
        u = R[0, 0] * (X - x0) + R[0, 1] * (Y - y0)
        v = R[1, 0] * (X - x0) + R[1, 1] * (Y - y0)
        q = (u / a) ** 2 + (v / b) ** 2
        ref += A * (q <= 1)
        near |= np.abs(q - 1) < 1e-9
    ref = np.maximum(ref, 0)
    assert near.mean() < 0.01
    np.testing.assert_allclose(im[~near], ref[~near], atol=1e-12)


def test_grains_against_full_grid_voronoi(oracle):
    """grains, phantomgallery.m:361-389 taken literally: labels of the whole
    Nbig x Nbig grid (meshgrid(1:Nbig), column-major, first centre on ties),
    reshaped, cropped to rows and columns dN+(1:N), then scaled."""
    for N, ng_arg, seed in ((40, None, 1), (64, 20, 5), (37, 11, 2)):
        ng = int(math.floor(3 * math.sqrt(N) + 0.5)) if ng_arg is None else ng_arg
        rs = np.random.RandomState(seed)
        dN = int(math.floor(N / 10 + 0.5))
        Nbig = N + 2 * dN
        xG = np.ceil(Nbig * rs.random_sample(ng))
        yG = np.ceil(Nbig * rs.random_sample(ng))
        Xg, Yg = np.meshgrid(np.arange(1, Nbig + 1.0), np.arange(1, Nbig + 1.0))
        Xc, Yc = Xg.ravel(order="F"), Yg.ravel(order="F")
        dist = (Xc[:, None] - xG[None, :]) ** 2 + (Yc[:, None] - yG[None, :]) ** 2
        lab = np.argmin(dist, axis=1).reshape((Nbig, Nbig), order="F").astype(np.float64)
        lab = lab[dN:dN + N, dN:dN + N]
        ref = lab / lab.max()
        got = oracle.phantomgallery("grains", N, ng_arg, seed)
        assert np.array_equal(got, ref), (N, ng, seed)
        # ties exist at these sizes (integer distances), so first-index matters
        assert (np.sort(dist, axis=1)[:, 0] == np.sort(dist, axis=1)[:, 1]).any()


def _threephases_literal(N, p, seed):
    """phantomgallery.m:257-292 with MATLAB's array semantics: [I,J] = meshgrid(1:N),
    c = rand(p,2)*N filled column-major, blobs added one at a time."""
    rs = np.random.RandomState(seed)
    I, J = np.meshgrid(np.arange(1, N + 1.0), np.arange(1, N + 1.0))
    s1 = 0.025 * N
    c1 = rs.random_sample(2 * p).reshape((p, 2), order="F") * N
    im1 = np.zeros((N, N))
    for i in range(p):
        im1 = im1 + np.exp(-np.abs(I - c1[i, 0]) ** 3 / (2.5 * s1) ** 3 - np.abs(J - c1[i, 1]) ** 3 / s1 ** 3)
    im1 = np.where(im1 < 0.35, 0.0, 2.0)
    s2 = 0.025 * N
    c2 = rs.random_sample(2 * p).reshape((p, 2), order="F") * N
    im2 = np.zeros((N, N))
    for i in range(p):
        im2 = im2 + np.exp(-(I - c2[i, 0]) ** 2 / (2 * s2) ** 2 - (J - c2[i, 1]) ** 2 / s2 ** 2)
    im2 = np.where(im2 < 0.55, 0.0, 1.0)
    im = im1 + im2
    im[im == 3] = 1
    return im / im.max()


def test_threephases_and_smooth_against_literal_loops(oracle):
    for N, p, seed in ((48, 100, 3), (64, 30, 1), (33, 100, 8)):
        ref = _threephases_literal(N, p, seed)
        assert np.array_equal(oracle.phantomgallery("threephases", N, p, seed), ref), (N, p, seed)
        assert len(np.unique(ref)) == 3
    for N, p in ((40, 4), (31, 2)):
        I, J = np.meshgrid(np.arange(1, N + 1.0), np.arange(1, N + 1.0))
        sigma = 0.25 * N
        c = np.array([[0.6 * N, 0.6 * N], [0.5 * N, 0.3 * N], [0.2 * N, 0.7 * N], [0.8 * N, 0.2 * N]])
        a = [1, 0.5, 0.7, 0.9]
        im = np.zeros((N, N))
        for i in range(p):
            im = im + a[i] * np.exp(-(I - c[i, 0]) ** 2 / (1.2 * sigma) ** 2 - (J - c[i, 1]) ** 2 / sigma ** 2)
        np.testing.assert_allclose(oracle.phantomgallery("smooth", N, p), im / im.max(), rtol=1e-14, atol=0)


# ------------------------------------------------------------------ NCP statistic
def _ncp_numpy(r, res_dims):
    """check_stoprules.m:96-133 with numpy's FFT (an independent transform)."""
    r = np.ravel(r)
    if len(res_dims) == 1:
        R = r[:, None]
    else:
        R = r.reshape(res_dims[1], res_dims[0]).T          # column-major reshape(rk, res_dims)
    q = R.shape[0] // 2
    PK = np.abs(np.fft.fft(R, axis=0)[: q + 1]) ** 2
    c = np.cumsum(PK[1:], axis=0) / PK[1:].sum(axis=0)
    cw = (np.arange(1, q + 1) / q)[:, None]
    return float(np.mean(np.sqrt(((c - cw) ** 2).sum(axis=0))))


@pytest.mark.parametrize("dims", [[91, 7], [64, 5], [725, 3], [1000]])
def test_ncp_statistic_matches_fft(oracle, dims):
    rng = np.random.default_rng(11)
    r = rng.standard_normal(int(np.prod(dims))) + np.linspace(0, 3, int(np.prod(dims)))
    assert abs(oracle.ncp_stat(r, dims) - _ncp_numpy(r, dims)) <= 1e-12


def _ncp_first_rise(nccs, smooth):
    """The NCP window rule (check_stoprules.m:135-151) in closed form:
    max(dk) < max([dk(2:end); ncc]) holds exactly when ncc exceeds every entry
    of dk, so the run stops at the first k with ncc_k > max(ncc_{k-smooth..k-1}),
    the entries before k = 0 being Inf (check_inputs.m:207-209)."""
    for k, v in enumerate(nccs):
        if v > max(np.inf if j < 0 else nccs[j] for j in range(k - smooth, k)):
            return k
    return None


@pytest.mark.parametrize("meth,smooth,two_d", [("kaczmarz", 2, True), ("kaczmarz", 1, False),
                                               ("randkaczmarz", 3, True), ("columnaction", 3, True)])
def test_row_action_ncp_stop_matches_closed_form(oracle, meth, smooth, two_d):
    N, theta, p = 16, np.arange(0.0, 180.0, 10.0), 23
    A = oracle.paralleltomo(N, theta, p)
    b = A @ oracle.phantomgallery("shepplogan", N).ravel(order="F")
    e = np.random.default_rng(2).standard_normal(b.size)
    b = b + 0.02 * np.linalg.norm(b) * e / np.linalg.norm(e)
    dims = [p, len(theta)] if two_d else [A.shape[0]]
    run = oracle.cart if meth == "columnaction" else oracle.art
    kmax = 60
    X, _ = run(meth, A, b, np.arange(1, kmax + 1), None, {"seed": 5})
    nccs = [_ncp_numpy(b, dims)] + [_ncp_numpy(b - A @ X[:, k], dims) for k in range(kmax)]
    want = _ncp_first_rise(nccs, smooth)
    assert want is not None and 1 <= want < kmax
    _, info = run(meth, A, b, kmax, None, {"seed": 5, "stoprule": {"type": "NCP", "res_dims": dims,
                                                                  "ncp_smooth": smooth}})
    assert info["stoprule"] == 1 and info["finaliter"] == want


@pytest.mark.parametrize("p,f0", [(91, 5), (64, 1), (64, 32), (101, 50)])
def test_ncp_statistic_single_frequency_closed_form(oracle, p, f0):
    """A pure cosine at frequency f0 puts all of pk(1:q) at f0, so c = [f >= f0]
    and ncc = sqrt(sum_{f<f0} (f/q)^2 + sum_{f>=f0} (1 - f/q)^2)."""
    t = np.arange(p)
    r = np.cos(2 * np.pi * f0 * t / p)
    q = p // 2
    f = np.arange(1, q + 1)
    expect = math.sqrt(np.sum(np.where(f < f0, (f / q) ** 2, (1 - f / q) ** 2)))
    assert abs(oracle.ncp_stat(np.tile(r, 3), [p, 3]) - expect) <= 1e-9


def test_ncp_stop_index_follows_window_rule(oracle):
    """The oracle's NCP stop (check_stoprules.m:96-151, dk = Inf(ncp_smooth,1)) against
    the rule applied to numpy-FFT statistics of the residuals of the same run."""
    N, theta, p = 24, np.arange(0.0, 180.0, 6.0), 35
    A = oracle.paralleltomo(N, theta, p)
    x = oracle.phantomgallery("shepplogan", N).ravel(order="F")
    b = A @ x
    b = b + 0.05 * np.linalg.norm(b) * np.random.default_rng(3).standard_normal(b.size) / np.sqrt(b.size)
    rho = oracle.spectral_radius(A, *oracle.weights("cimmino", A))
    kmax = 60
    for smooth in (1, 2, 4):
        st = {"type": "NCP", "res_dims": [p, len(theta)], "ncp_smooth": smooth}
        X, info, _ = oracle.sirt("cimmino", A, b, kmax, None, {"rho": rho, "stoprule": st})
        Xall, _, _ = oracle.sirt("cimmino", A, b, list(range(1, kmax + 1)), None, {"rho": rho})
        dk = [math.inf] * smooth
        expect = kmax
        for k in range(0, kmax + 1):
            xk = np.zeros(A.shape[1]) if k == 0 else Xall[:, k - 1]
            ncc = _ncp_numpy(b - A @ xk, [p, len(theta)])
            if max(dk) < max(dk[1:] + [ncc]) or k >= kmax:
                expect = k
                break
            dk = dk[1:] + [ncc]
        assert info["finaliter"] == expect, (smooth, info["finaliter"], expect)
        assert info["stoprule"] == (1 if expect < kmax else 0)
        assert expect < kmax


# ------------------------------------------------------------------ learned relaxation network
def test_mlp_oracle_gradient_matches_finite_differences():
    """The FP64 network reference: analytic gradient against central differences."""
    from oracle import mlp
    rng = np.random.default_rng(4)
    p = mlp.init_params(1).astype(np.float64) + 0.01 * rng.standard_normal(mlp.NPARAM)
    X = rng.standard_normal((37, mlp.NI))
    T = np.stack([rng.uniform(0.1, 1.9, 37), (rng.uniform(size=37) > 0.5).astype(float)], 1)
    T[::5, 0] = -1                                   # dropped lambda terms
    L, g = mlp.loss_grad(p, X, T, stop_w=0.7)
    h = 1e-6
    for j in list(rng.choice(mlp.NPARAM, 60, replace=False)) + [1024, 5376, 5377]:
        e = np.zeros(mlp.NPARAM); e[j] = h
        fd = (mlp.loss_grad(p + e, X, T, 0.7)[0] - mlp.loss_grad(p - e, X, T, 0.7)[0]) / (2 * h)
        assert abs(fd - g[j]) <= 1e-6 * max(1.0, abs(fd)), (j, fd, g[j])


def test_mlp_bf16_rounding():
    from oracle import mlp
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5e-3, 0.0, 3.0e38])
    r = mlp.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7      # ties to even
    assert abs(r[3] - v[3]) <= 2 ** -9 * abs(v[3]) and r[4] == 0.0


# ------------------------------------------------------------------ binary phantom
def _marg2id_literal(marginal):
    """marg2id, phantomgallery.m:630-659, with its index arithmetic: 1-based
    ids, index_tmp = reshape(1:2^sN, 2^(i-1), []) (column-major), the columns
    marginal(i)+1 : 2 : end kept for each known entry."""
    sN = len(marginal)
    index = np.arange(1, 2 ** sN + 1)
    known = [i for i in range(sN) if marginal[i] >= 0]
    Index = np.zeros((2 ** sN, len(known)), np.int64)
    for c, i in enumerate(known):
        index_tmp = index.reshape((2 ** i, 2 ** sN // 2 ** i), order="F")
        index_use = index_tmp[:, int(marginal[i])::2]
        Index[index_use.ravel(order="F") - 1, c] = 1
    return np.nonzero(Index.sum(axis=1) == len(known))[0] + 1


def _id2n_literal(id_):
    """id2n, :613-628: the base-2 digits of id-1, least significant first."""
    n, q = np.zeros(9), id_ - 1
    for i in range(9):
        qi = q // 2
        n[i] = q - 2 * qi
        q = qi
    return n


def _binary_literal(N, seed, H):
    """sequential_simulation_full_clique, :511-611, with NaN for unknown pixels,
    column-major marginals and find(cdf > r, 1)."""
    rs = np.random.RandomState(seed)
    sim = np.full((N, N), np.nan)
    cdf = np.cumsum(H) / H.sum()
    id_ = int(np.nonzero(cdf > rs.random_sample())[0][0]) + 1
    sim[0:3, 0:3] = _id2n_literal(id_).reshape((3, 3), order="F")

    def draw(block):
        marginal = block.ravel(order="F").copy()
        marginal[np.isnan(marginal)] = -1
        ids = _marg2id_literal(marginal)
        pdf = H[ids - 1] / np.sum(H[ids - 1])
        c = np.cumsum(pdf) / np.sum(pdf)
        pos = int(np.nonzero(c > rs.random_sample())[0][0])
        return _id2n_literal(ids[pos]).reshape((3, 3), order="F")

    for j in range(3, N):
        sim[0:3, j] = draw(sim[0:3, j - 2:j + 1])[:, -1]
    for i in range(3, N):
        for j in range(2, N):
            n = draw(sim[i - 2:i + 1, j - 2:j + 1])
            if j == 2:
                sim[i, 0:3] = n[-1, :]
            else:
                sim[i, j] = n[-1, -1]
    return sim


def test_binary_training_image_and_histogram(oracle):
    """getTI: 81 x 84, 1957 ones (one per IJ pair of phantomgallery.m:682-2638).
    The clique histogram counts every 3 x 3 window once; its all-zero and
    all-one bins equal the window sums 0 and 9 of a box filter; each bin equals
    the count of windows matching its pattern (bits read column-major, as
    id2n/reshape define them); the mirror's histogram is the histogram with the
    window's first and last columns exchanged."""
    from scipy.ndimage import correlate
    from airtoolsii_amd import phantoms
    TI = oracle.binary_training_image()
    assert TI.shape == (81, 84) and int(TI.sum()) == 1957 and set(np.unique(TI)) == {0, 1}
    assert np.array_equal(phantoms.binary_training_image(), TI)
    H = oracle.clique_histogram(TI)
    assert H.sum() == 79 * 82
    box = correlate(TI.astype(np.int64), np.ones((3, 3), np.int64), mode="constant")[1:-1, 1:-1]
    assert H[0] == np.sum(box == 0) and H[511] == np.sum(box == 9)
    T = 2 * TI.astype(np.int64) - 1
    for id1 in (2, 3, 5, 9, 17, 33, 65, 129, 257, 300, 448, 73, 146, 292):
        P = _id2n_literal(id1).reshape((3, 3), order="F")
        score = correlate(T, (2 * P - 1).astype(np.int64), mode="constant")[1:-1, 1:-1]
        assert H[id1 - 1] == np.sum(score == 9), id1
    Hf = oracle.clique_histogram(TI[:, ::-1])
    swap = [((q & 0o007) << 6) | (q & 0o070) | ((q >> 6) & 0o007) for q in range(512)]
    assert np.array_equal(Hf, H[swap])
    assert not np.array_equal(Hf, H)
    np.testing.assert_array_equal(phantoms.clique_histogram(TI), H)


@pytest.mark.parametrize("N,seed", [(3, 1), (4, 7), (6, 2), (9, 11)])
def test_binary_against_literal_simulation(oracle, N, seed):
    from airtoolsii_amd import phantoms
    TI = oracle.binary_training_image()
    H = oracle.clique_histogram(TI) + oracle.clique_histogram(TI[:, ::-1]) + np.finfo(float).eps
    ref = _binary_literal(N, seed, H)
    assert np.array_equal(oracle.phantomgallery("binary", N, seed), ref)
    assert np.array_equal(phantoms.binary(N, seed), ref)
    assert oracle.binary_draws(N, seed) == 1 + (N - 3) * (N - 1)


def test_binary_product_matches_oracle(oracle):
    from airtoolsii_amd import phantomgallery
    for N, seed in ((64, 1), (97, 5), (128, 3)):
        a = phantomgallery("binary", N, seed)
        b = oracle.phantomgallery("binary", N, seed)
        assert np.array_equal(a, b), (N, seed)
        # the training image is about a quarter ones; the simulation follows it loosely
        assert 0.05 < a.mean() < 0.6
    assert not np.array_equal(phantomgallery("binary", 64, 1), phantomgallery("binary", 64, 2))
    with pytest.raises(ValueError):
        phantomgallery("binary", 2, 1)


# ------------------------------------------------------------------ function-handle weights
def test_weights_handle_branches_closed_forms(oracle):
    """get_mfun_dfun.m's function-handle branches against closed forms on the
    matrix behind the handle: row norms (Cimmino, DROP), column counts (DROP),
    sum_j a_ij^2 / nnz_j (CAV's handle branch, which divides where the matrix
    branch multiplies), |A 1| and |A' 1| (SART); an empty row gives weight 0."""
    import scipy.sparse as sp
    rng = np.random.default_rng(3)
    A = sp.random(30, 20, density=0.2, random_state=4, format="csr")
    A = sp.csr_matrix(A + sp.csr_matrix((np.ones(20), (np.arange(20) % 30, np.arange(20))), shape=(30, 20)))
    A = sp.lil_matrix(A); A[7, :] = 0; A = A.tocsr(); A.eliminate_zeros()
    fn = lambda v, t: A @ v if t == "notransp" else A.T @ v
    m, n = A.shape
    rsq = np.asarray(A.multiply(A).sum(axis=1)).ravel()
    cnz = np.diff(A.tocsc().indptr).astype(float)
    assert rsq[7] == 0 and (cnz > 0).all()
    with np.errstate(divide="ignore"):
        M, D = oracle.weights_handle("cimmino", fn, m, n)
        ref = np.where(rsq > 0, 1.0 / (m * rsq), 0.0)
        np.testing.assert_allclose(M, ref, rtol=1e-13); assert np.all(D == 1)
        M, D = oracle.weights_handle("drop", fn, m, n)
        np.testing.assert_allclose(M, np.where(rsq > 0, 1.0 / rsq, 0.0), rtol=1e-13)
        np.testing.assert_allclose(D, 1.0 / cnz, rtol=1e-15)
        M, D = oracle.weights_handle("cav", fn, m, n)
        w = np.asarray(A.multiply(A) @ (1.0 / cnz)).ravel()
        np.testing.assert_allclose(M, np.where(w > 0, 1.0 / w, 0.0), rtol=1e-13)
        # the matrix branch weights the squares by the counts instead
        Mm, _ = oracle.weights("cav", A)
        assert not np.allclose(M, Mm)
        M, D = oracle.weights_handle("sart", fn, m, n)
        ra, ca = np.abs(A @ np.ones(n)), np.abs(A.T @ np.ones(m))
        np.testing.assert_allclose(M, np.where(ra > 0, 1.0 / ra, 0.0), rtol=1e-14)
        np.testing.assert_allclose(D, 1.0 / ca, rtol=1e-14)
    # an empty column makes CAV's handle branch 0 * Inf = NaN in every row, as in MATLAB
    B = sp.lil_matrix(A); B[:, 3] = 0; B = B.tocsr(); B.eliminate_zeros()
    fb = lambda v, t: B @ v if t == "notransp" else B.T @ v
    M, _ = oracle.weights_handle("cav", fb, m, n)
    assert np.isnan(M).all()
