// air_tracer.hip -- K1: line-model ray tracer for paralleltomo / fanlineartomo /
// fancurvedtomo / seismictomo geometries on gfx950, emitting CSR that matches MATLAB's sparse() of the
// reference tracer entry for entry, plus the matrix-free ray-driven projector.
//
// Reference algorithm: testprobs/paralleltomo.m:180-301 and
// testprobs/fanlineartomo.m:227-357.  Per ray the reference computes the
// crossings with all 2(N+1) grid lines, sorts them by the ray parameter t
// (stable), keeps the points inside the box, drops a point that coincides
// (1e-10) with its successor, and turns each pair of consecutive points into
// one (column, length) entry.  Here each GPU lane walks one ray instead: the
// x-line and y-line crossings are each monotone in t, so a two-pointer merge
// (x-crossing first on equal t, which is the stable-sort order) replaces the
// sort, and the dedup/length/midpoint steps run on the merged stream.  All
// arithmetic is FP64 in the reference's operation order and this file is
// compiled with -ffp-contract=off, so every entry is bit-identical to the
// FP64 CPU tracer.
//
// CSR order: sparse() stores columns ascending.  Column = fx*N + (N-1-fy)
// with fx = floor(xm), fy = floor(ym) of the segment midpoint; along a ray
// fx and fy are monotone, so walking the ray in the direction of increasing
// fx and reversing each run of equal fx when fy increases yields ascending
// columns without a sort.  A separate check pass verifies the order.
#include <algorithm>
#include <cmath>
#include <vector>

#include "air_internal.h"

namespace air {

// ---------------------------------------------------------------- host geometry
// MATLAB cosd/sind: exact 0 and +-1 at multiples of 90 degrees; elsewhere the
// remainder in [-45,45] degrees is evaluated in radians.
static void sincosd(double deg, double* s, double* c) {
    if (!std::isfinite(deg)) { *s = NAN; *c = NAN; return; }
    double r = std::fmod(deg, 360.0);
    double q = std::nearbyint(r / 90.0);
    double z = r - 90.0 * q;
    int iq = ((int)q % 4 + 4) % 4;
    double zr = z * (M_PI / 180.0);
    double sz = std::sin(zr), cz = std::cos(zr);
    if (z == 0.0) { sz = 0.0; cz = 1.0; }
    double ss, cc;
    switch (iq) {
    case 0: ss = sz; cc = cz; break;
    case 1: ss = cz; cc = -sz; break;
    case 2: ss = -sz; cc = -cz; break;
    default: ss = -cz; cc = sz; break;
    }
    if (ss == 0.0) ss = 0.0;
    if (cc == 0.0) cc = 0.0;
    *s = ss;
    *c = cc;
}

// paralleltomo.m:118-122,195-200: x0 = linspace(-d/2,d/2,p), y0 = 0, rotated by
// theta; direction (a,b) = (-sind, cosd).
void host_rays_paralleltomo(int N, const double* theta, int nA, int p, double d, std::vector<RayParam>& out) {
    (void)N;
    std::vector<double> x0(p);
    double d1 = -d / 2.0, d2 = d / 2.0;
    if (p == 1) x0[0] = d2;
    else {
        double n1 = (double)(p - 1);
        for (int k = 0; k < p; ++k) x0[k] = d1 + ((double)k * (d2 - d1)) / n1;
        x0[0] = d1;
        x0[p - 1] = d2;
    }
    out.resize((size_t)nA * p);
    for (int i = 0; i < nA; ++i) {
        double s, c;
        sincosd(theta[i], &s, &c);
        for (int j = 0; j < p; ++j) {
            RayParam r;
            r.x0 = c * x0[j] - s * 0.0;
            r.y0 = s * x0[j] + c * 0.0;
            r.a = -s;
            r.b = c;
            out[(size_t)i * p + j] = r;
        }
    }
}

// fanlineartomo.m:141-169,227-257: source at Omega(theta)*[0;R]; ray j is the
// central direction rotated by omega_j = 90 - atand(sd/depos_j).
void host_rays_fanlineartomo(int N, const double* theta, int nA, int p, double R, double dw, double sd,
                             std::vector<RayParam>& out) {
    (void)N;
    double dew = dw / p;
    std::vector<double> omega(p);
    if (p % 2) {
        int h = (p - 1) / 2;
        omega[h] = 0.0;
        for (int k = 1; k <= h; ++k) {
            double om = 90.0 - (180.0 / M_PI) * std::atan(sd / ((double)k * dew));
            omega[h + k] = om;
            omega[h - k] = -om;
        }
    } else {
        int h = p / 2;
        for (int k = 1; k <= h; ++k) {
            double om = 90.0 - (180.0 / M_PI) * std::atan(sd / (((double)k - 0.5) * dew));
            omega[h - 1 + k] = om;
            omega[h - k] = -om;
        }
    }
    std::vector<double> so(p), co(p);
    for (int j = 0; j < p; ++j) sincosd(omega[j], &so[j], &co[j]);
    out.resize((size_t)nA * p);
    for (int i = 0; i < nA; ++i) {
        double st, ct;
        sincosd(theta[i], &st, &ct);
        double x0t = ct * 0.0 + (-st) * R;
        double y0t = st * 0.0 + ct * R;
        double abx = (-x0t) / R, aby = (-y0t) / R;
        for (int j = 0; j < p; ++j) {
            RayParam r;
            r.x0 = x0t;
            r.y0 = y0t;
            r.a = co[j] * abx + (-so[j]) * aby;
            r.b = so[j] * abx + co[j] * aby;
            out[(size_t)i * p + j] = r;
        }
    }
}

// fancurvedtomo.m:140-154,210-241: source at Omega(theta)*[0;R]; ray j turns
// the central direction -xytheta/R by omega_j, omega = linspace(-d/2,d/2,p)
// (equal angular steps: the curved detector).
void host_rays_fancurvedtomo(int N, const double* theta, int nA, int p, double R, double d,
                             std::vector<RayParam>& out) {
    (void)N;
    std::vector<double> omega(p);
    const double d1 = -d / 2.0, d2 = d / 2.0;
    if (p == 1) omega[0] = d2;
    else {
        const double n1 = (double)(p - 1);
        for (int k = 0; k < p; ++k) omega[k] = d1 + ((double)k * (d2 - d1)) / n1;
        omega[0] = d1;
        omega[p - 1] = d2;
    }
    std::vector<double> so(p), co(p);
    for (int j = 0; j < p; ++j) sincosd(omega[j], &so[j], &co[j]);
    out.resize((size_t)nA * p);
    for (int i = 0; i < nA; ++i) {
        double st, ct;
        sincosd(theta[i], &st, &ct);
        const double x0t = ct * 0.0 + (-st) * R;
        const double y0t = st * 0.0 + ct * R;
        const double abx = (-x0t) / R, aby = (-y0t) / R;
        for (int j = 0; j < p; ++j) {
            RayParam r;
            r.x0 = x0t;
            r.y0 = y0t;
            r.a = co[j] * abx + (-so[j]) * aby;
            r.b = so[j] * abx + co[j] * aby;
            out[(size_t)i * p + j] = r;
        }
    }
}

// linspace(d1,d2,n) as MATLAB evaluates it: d1 + (k*(d2-d1))/(n-1), both end
// points exact, n = 1 gives d2, n = 0 nothing.
static void linspace_into(double d1, double d2, int n, double* out) {
    if (n <= 0) return;
    if (n == 1) { out[0] = d2; return; }
    const double n1 = (double)(n - 1);
    for (int k = 0; k < n; ++k) out[k] = d1 + ((double)k * (d2 - d1)) / n1;
    out[0] = d1;
    out[n - 1] = d2;
}

// seismictomo.m:98-116,187-202: s sources on the right edge x = N/2, p
// receivers, floor(p/2) on the left edge and ceil(p/2) on the top edge; ray
// (i,j) starts at source i with direction (xp_j - x0_i, yp_j - y0_i), row i*p + j.
void seismic_positions(int N, int s, int p, std::vector<double>& y0, std::vector<double>& xp,
                       std::vector<double>& yp) {
    const double N2 = (double)N / 2.0;
    const double Ns = ((double)N / s) / 2.0;
    y0.assign(s, 0.0);
    linspace_into(-N2 + Ns, N2 - Ns, s, y0.data());
    const int p1 = (p + 1) / 2, p2 = p / 2;
    const double Np1 = ((double)N / p1) / 2.0, Np2 = ((double)N / p2) / 2.0;
    xp.assign(p, 0.0);
    yp.assign(p, 0.0);
    for (int j = 0; j < p2; ++j) xp[j] = -N2;
    linspace_into(-N2 + Np1, N2 - Np1, p1, xp.data() + p2);
    linspace_into(-N2 + Np2, N2 - Np2, p2, yp.data());
    for (int j = p2; j < p; ++j) yp[j] = N2;
}

void host_rays_seismictomo(int N, int s, int p, std::vector<RayParam>& out) {
    const double N2 = (double)N / 2.0;
    std::vector<double> y0, xp, yp;
    seismic_positions(N, s, p, y0, xp, yp);
    out.resize((size_t)s * p);
    for (int i = 0; i < s; ++i)
        for (int j = 0; j < p; ++j) {
            RayParam r;
            r.x0 = N2;
            r.y0 = y0[i];
            r.a = xp[j] - N2;
            r.b = yp[j] - y0[i];
            out[(size_t)i * p + j] = r;
        }
}

// ---------------------------------------------------------------- device ray walk
struct Pt {
    double t, x, y;
};

// Walk one ray's kept points in traversal order (fwd: increasing t) and call
// seg(x_u, y_u, x_v, y_v) for each pair of consecutive kept points.
template <class Seg>
__device__ __forceinline__ void walk_ray(const RayParam& r, int N, bool fwd, Seg&& seg) {
    const double half = (double)N / 2.0;
    const double lo = -half;
    // x-line crossings: t ascending along k iff a > 0; none when a is +-0.
    const bool hasXlist = (r.a > 0.0) || (r.a < 0.0);
    const bool hasYlist = (r.b > 0.0) || (r.b < 0.0);
    int kx = 0, dkx = 1, nx = hasXlist ? N + 1 : 0;
    if (hasXlist && ((r.a > 0.0) != fwd)) { kx = N; dkx = -1; }
    int ky = 0, dky = 1, ny = hasYlist ? N + 1 : 0;
    if (hasYlist && ((r.b > 0.0) != fwd)) { ky = N; dky = -1; }

    Pt X, Y;
    bool hx = false, hy = false;
    auto nextX = [&]() {
        hx = false;
        while (nx > 0) {
            double g = lo + (double)kx;
            kx += dkx;
            nx--;
            double tx = (g - r.x0) / r.a;
            double yx = r.b * tx + r.y0;
            if (yx >= lo && yx <= half) { X.t = tx; X.x = g; X.y = yx; hx = true; return; }
        }
    };
    auto nextY = [&]() {
        hy = false;
        while (ny > 0) {
            double g = lo + (double)ky;
            ky += dky;
            ny--;
            double ty = (g - r.y0) / r.b;
            double xy = r.a * ty + r.x0;
            if (xy >= lo && xy <= half) { Y.t = ty; Y.x = xy; Y.y = g; hy = true; return; }
        }
    };
    nextX();
    nextY();
    bool havePrev = false, haveQ = false;
    double px = 0, py = 0, qx = 0, qy = 0;
    auto kept = [&](double x, double y) {
        if (haveQ) seg(qx, qy, x, y);
        qx = x; qy = y; haveQ = true;
    };
    while (hx || hy) {
        bool takeX = fwd ? (hx && (!hy || X.t <= Y.t)) : (hx && (!hy || X.t > Y.t));
        double x, y;
        if (takeX) { x = X.x; y = X.y; nextX(); }
        else { x = Y.x; y = Y.y; nextY(); }
        if (!havePrev) {
            havePrev = true;
            px = x; py = y;
            if (!fwd) kept(x, y);
            continue;
        }
        bool coinc = fabs(x - px) <= 1e-10 && fabs(y - py) <= 1e-10;
        if (fwd) { if (!coinc) kept(px, py); }
        else { if (!coinc) kept(x, y); }
        px = x; py = y;
    }
    if (fwd && havePrev) kept(px, py);
}

struct SegOut {
    int col;      // 0-based column, or -1 when outside 1..N^2
    int fx;
    double val;
};

__device__ __forceinline__ SegOut seg_value(int N, double xu, double yu, double xv, double yv) {
    const double half = (double)N / 2.0;
    double dx = xv - xu, dy = yv - yu;
    SegOut o;
    o.val = sqrt(dx * dx + dy * dy);
    double xm = 0.5 * (xu + xv) + half;
    double ym = 0.5 * (yu + yv) + half;
    double fx = floor(xm), fy = floor(ym);
    double c1 = fx * (double)N + ((double)N - fy);
    double nn = (double)N * (double)N;
    o.col = (c1 >= 1.0 && c1 <= nn) ? (int)c1 - 1 : -1;
    o.fx = (int)fx;
    return o;
}

__device__ __forceinline__ bool ray_excluded(const RayParam& r, int N, int edge_check) {
    const double half = (double)N / 2.0;
    return edge_check && ((r.b == 0.0 && fabs(r.y0 - half) < 1e-15) || (r.a == 0.0 && fabs(r.x0 - half) < 1e-15));
}

// CSR traversal direction: fx non-decreasing along the walk.
__device__ __forceinline__ bool csr_forward(const RayParam& r) { return !(r.a < 0.0); }
// fy increases along the walk -> reverse each run of equal fx.
__device__ __forceinline__ bool csr_rev_groups(const RayParam& r, bool fwd) {
    return fwd ? (r.b > 0.0) : (r.b < 0.0);
}

// Pass 1: entries per ray (adjacent equal columns merged as sparse() sums them).
__global__ __launch_bounds__(256) void k_trace_count(const RayParam* __restrict__ rays, int64_t nrays, int N,
                                                     int edge_check, int64_t* __restrict__ len,
                                                     int* __restrict__ err) {
    int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= nrays) return;
    RayParam r = rays[ray];
    int64_t cnt = 0;
    if (!ray_excluded(r, N, edge_check)) {
        bool fwd = csr_forward(r);
        int last = -2;
        bool bad = false;
        walk_ray(r, N, fwd, [&](double xu, double yu, double xv, double yv) {
            SegOut o = seg_value(N, xu, yu, xv, yv);
            if (o.col < 0) bad = true;
            if (o.col != last) cnt++;
            last = o.col;
        });
        if (bad) atomicOr(err, 1);
    }
    len[ray] = cnt;
}

template <typename V>
__device__ __forceinline__ void reverse_run(int* col, V* val, int64_t a, int64_t b) {
    for (int64_t i = a, j = b - 1; i < j; ++i, --j) {
        int tc = col[i]; col[i] = col[j]; col[j] = tc;
        V tv = val[i]; val[i] = val[j]; val[j] = tv;
    }
}

// Pass 2: write entries in ascending column order.
template <typename V>
__global__ __launch_bounds__(256) void k_trace_fill(const RayParam* __restrict__ rays, int64_t nrays, int N,
                                                    int edge_check, const int64_t* __restrict__ rowptr,
                                                    int* __restrict__ col, V* __restrict__ val) {
    int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= nrays) return;
    RayParam r = rays[ray];
    if (ray_excluded(r, N, edge_check)) return;
    const int64_t base = rowptr[ray];
    const int64_t cap = rowptr[ray + 1] - base;
    bool fwd = csr_forward(r);
    bool rev = csr_rev_groups(r, fwd);
    int64_t pos = 0, gstart = 0;
    int gfx = INT32_MIN;
    int pc = -2, pfx = 0;
    double pv = 0.0;
    auto put = [&](int c, int fx, double v) {
        if (fx != gfx) {
            if (rev) reverse_run(col + base, val + base, gstart, pos < cap ? pos : cap);
            gstart = pos;
            gfx = fx;
        }
        if (pos < cap) {
            col[base + pos] = c;
            val[base + pos] = from_f64<V>(v);
        }
        pos++;
    };
    walk_ray(r, N, fwd, [&](double xu, double yu, double xv, double yv) {
        SegOut o = seg_value(N, xu, yu, xv, yv);
        if (o.col == pc) { pv += o.val; return; }
        if (pc != -2) put(pc, pfx, pv);
        pc = o.col; pfx = o.fx; pv = o.val;
    });
    if (pc != -2) put(pc, pfx, pv);
    if (rev) reverse_run(col + base, val + base, gstart, pos < cap ? pos : cap);
}

// Check pass: columns strictly ascending and inside [0, n) (wave per row).
__global__ __launch_bounds__(256) void k_check_rows(const int64_t* __restrict__ rowptr, const int* __restrict__ col,
                                                    int64_t m, int64_t n, int* __restrict__ err) {
    int64_t row = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    int lane = threadIdx.x & 63;
    if (row >= m) return;
    int64_t s = rowptr[row], e = rowptr[row + 1];
    bool bad = false;
    for (int64_t q = s + lane; q < e; q += 64) {
        int c = col[q];
        if (c < 0 || c >= n) bad = true;
        if (q + 1 < e && col[q + 1] <= c) bad = true;
    }
    if (bad) atomicOr(err, 2);
}

void trace_to_csr(const Geometry& g, int dtype, DevSparse& A, hipStream_t s) {
    const int64_t m = g.nrays;
    A.nrows = m;
    A.ncols = (int64_t)g.N * g.N;
    A.dtype = dtype;
    DevBuf len, err;
    len.alloc(sizeof(int64_t) * (m + 1));
    err.alloc(sizeof(int));
    AIR_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
    AIR_HIP(hipMemsetAsync(len.p, 0, sizeof(int64_t) * (m + 1), s));
    const RayParam* rays = g.rays.as<RayParam>();
    int blocks = (int)((m + 255) / 256);
    if (m > 0) k_trace_count<<<blocks, 256, 0, s>>>(rays, m, g.N, g.edge_check, len.as<int64_t>(), err.as<int>());
    AIR_HIP(hipGetLastError());
    A.ptr.alloc(sizeof(int64_t) * (m + 1));
    scan_exclusive_i64(len.as<int64_t>(), A.ptr.as<int64_t>(), m, s);
    int64_t nnz = 0;
    int herr = 0;
    AIR_HIP(hipMemcpyAsync(&nnz, A.ptr.as<int64_t>() + m, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    AIR_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    AIR_HIP(hipStreamSynchronize(s));
    AIR_CHECK(herr == 0, AIR_E_GEOMETRY, "Index exceeds matrix dimensions: a ray segment maps outside 1..N^2");
    AIR_CHECK(nnz < (int64_t)INT32_MAX * 64, AIR_E_ARG, "matrix too large");
    A.nnz = nnz;
    alloc_sparse_arrays(A, nnz, dtype, s);
    dispatch_dtype(dtype, [&](auto v) {
        using V = decltype(v);
        if (m > 0)
            k_trace_fill<V><<<blocks, 256, 0, s>>>(rays, m, g.N, g.edge_check, A.ptr.as<int64_t>(), A.idx.as<int>(),
                                                   A.val.as<V>());
        return 0;
    });
    AIR_HIP(hipGetLastError());
    if (m > 0) {
        int64_t threads = m * 64;
        k_check_rows<<<(unsigned)((threads + 255) / 256), 256, 0, s>>>(A.ptr.as<int64_t>(), A.idx.as<int>(), m,
                                                                      A.ncols, err.as<int>());
        AIR_HIP(hipGetLastError());
    }
    AIR_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    AIR_HIP(hipStreamSynchronize(s));
    AIR_CHECK(herr == 0, AIR_E_GEOMETRY, "ray tracer: CSR column order check failed");
}

// ---------------------------------------------------------------- matrix-free projector
// Forward: y_i = sum_s aval_s * u[col_s] (paralleltomo.m:284).  One lane per ray.
__global__ __launch_bounds__(256) void k_mf_forward(const RayParam* __restrict__ rays, int64_t nrays, int N,
                                                    int edge_check, const float* __restrict__ u,
                                                    float* __restrict__ y) {
    int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= nrays) return;
    RayParam r = rays[ray];
    float acc = 0.f;
    if (!ray_excluded(r, N, edge_check)) {
        walk_ray(r, N, true, [&](double xu, double yu, double xv, double yv) {
            SegOut o = seg_value(N, xu, yu, xv, yv);
            if (o.col >= 0) acc += (float)o.val * u[o.col];
        });
    }
    y[ray] = acc;
}

// Adjoint: g[col_s] += v_i * aval_s (paralleltomo.m:286), FP64 atomics.
__global__ __launch_bounds__(256) void k_mf_adjoint(const RayParam* __restrict__ rays, int64_t nrays, int N,
                                                    int edge_check, const float* __restrict__ v,
                                                    double* __restrict__ g) {
    int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= nrays) return;
    float vi = v[ray];
    if (vi == 0.f) return;
    RayParam r = rays[ray];
    if (ray_excluded(r, N, edge_check)) return;
    walk_ray(r, N, true, [&](double xu, double yu, double xv, double yv) {
        SegOut o = seg_value(N, xu, yu, xv, yv);
        if (o.col >= 0) unsafeAtomicAdd(&g[o.col], (double)vi * (double)(float)o.val);
    });
}

__global__ __launch_bounds__(256) void k_mf_row_stats(const RayParam* __restrict__ rays, int64_t nrays, int N,
                                                      int edge_check, const double* __restrict__ w,
                                                      double* __restrict__ sq, double* __restrict__ ab,
                                                      double* __restrict__ sqw) {
    int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= nrays) return;
    RayParam r = rays[ray];
    double s2 = 0, s1 = 0, s3 = 0;
    if (!ray_excluded(r, N, edge_check)) {
        walk_ray(r, N, true, [&](double xu, double yu, double xv, double yv) {
            SegOut o = seg_value(N, xu, yu, xv, yv);
            if (o.col < 0) return;
            double a = (double)(float)o.val;
            s2 += a * a;
            s1 += fabs(a);
            if (w) s3 += (a * a) * w[o.col];
        });
    }
    if (sq) sq[ray] = s2;
    if (ab) ab[ray] = s1;
    if (sqw) sqw[ray] = s3;
}

__global__ __launch_bounds__(256) void k_mf_col_stats(const RayParam* __restrict__ rays, int64_t nrays, int N,
                                                      int edge_check, double* __restrict__ colabs,
                                                      int* __restrict__ colcnt) {
    int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= nrays) return;
    RayParam r = rays[ray];
    if (ray_excluded(r, N, edge_check)) return;
    walk_ray(r, N, true, [&](double xu, double yu, double xv, double yv) {
        SegOut o = seg_value(N, xu, yu, xv, yv);
        if (o.col < 0) return;
        double a = (double)(float)o.val;
        if (colabs) unsafeAtomicAdd(&colabs[o.col], fabs(a));
        if (colcnt && a != 0.0) atomicAdd(&colcnt[o.col], 1);
    });
}

void mf_forward(const Geometry& g, const float* u, float* y, hipStream_t s) {
    if (g.nrays == 0) return;
    k_mf_forward<<<(unsigned)((g.nrays + 255) / 256), 256, 0, s>>>(g.rays.as<RayParam>(), g.nrays, g.N,
                                                                   g.edge_check, u, y);
    AIR_HIP(hipGetLastError());
}
void mf_adjoint(const Geometry& g, const float* v, double* gacc, hipStream_t s) {
    if (g.nrays == 0) return;
    k_mf_adjoint<<<(unsigned)((g.nrays + 255) / 256), 256, 0, s>>>(g.rays.as<RayParam>(), g.nrays, g.N,
                                                                   g.edge_check, v, gacc);
    AIR_HIP(hipGetLastError());
}
void mf_row_stats(const Geometry& g, const double* colw, double* sq, double* ab, double* sqw, hipStream_t s) {
    if (g.nrays == 0) return;
    k_mf_row_stats<<<(unsigned)((g.nrays + 255) / 256), 256, 0, s>>>(g.rays.as<RayParam>(), g.nrays, g.N,
                                                                     g.edge_check, colw, sq, ab, sqw);
    AIR_HIP(hipGetLastError());
}
void mf_col_stats(const Geometry& g, double* colabs, int* colcnt, hipStream_t s) {
    if (g.nrays == 0) return;
    k_mf_col_stats<<<(unsigned)((g.nrays + 255) / 256), 256, 0, s>>>(g.rays.as<RayParam>(), g.nrays, g.N,
                                                                     g.edge_check, colabs, colcnt);
    AIR_HIP(hipGetLastError());
}

}  // namespace air
