"""Closed-form geometry used to pin the tracers (test helper, not a tracer).

chord_length: length of the part of the line (x0,y0) + t (a,b) inside the
square [-h,h]^2 (Liang-Barsky clipping).  exact_ray_cells: brute force of one
ray with exact rational arithmetic on the binary64 ray parameters, giving the
set of crossed cells and their lengths without any rounding in the geometry.
"""
from fractions import Fraction as F
import math


def chord_length(x0, y0, a, b, h):
    t0, t1 = -math.inf, math.inf
    for p, d in ((x0, a), (y0, b)):
        if d == 0:
            if p < -h or p > h:
                return 0.0
            continue
        u, v = (-h - p) / d, (h - p) / d
        t0, t1 = max(t0, min(u, v)), min(t1, max(u, v))
    if t1 <= t0:
        return 0.0
    return (t1 - t0) * math.hypot(a, b)


def chord_endpoints(x0, y0, a, b, h):
    t0, t1 = -math.inf, math.inf
    for p, d in ((x0, a), (y0, b)):
        if d == 0:
            if p < -h or p > h:
                return None
            continue
        u, v = (-h - p) / d, (h - p) / d
        t0, t1 = max(t0, min(u, v)), min(t1, max(u, v))
    if t1 <= t0:
        return None
    return (x0 + t0 * a, y0 + t0 * b), (x0 + t1 * a, y0 + t1 * b)


def exact_ray_cells(N, x0, y0, a, b):
    """{column (0-based, x = im(:) column-major, y flipped): length} for one
    ray, from exact rational crossings.  Points coinciding exactly merge."""
    h = F(N, 2)
    X0, Y0, A, B = F(x0), F(y0), F(a), F(b)
    pts = set()
    for k in range(N + 1):
        g = -h + k
        if A != 0:
            t = (g - X0) / A
            y = B * t + Y0
            if -h <= y <= h:
                pts.add((t, g, y))
        if B != 0:
            t = (g - Y0) / B
            x = A * t + X0
            if -h <= x <= h:
                pts.add((t, x, g))
    # merge exactly coincident points (corners)
    uniq = {}
    for t, x, y in pts:
        uniq[(x, y)] = t
    seq = sorted(((t, x, y) for (x, y), t in uniq.items()))
    out = {}
    for (t1, xa, ya), (t2, xb, yb) in zip(seq, seq[1:]):
        xm = (xa + xb) / 2 + h
        ym = (ya + yb) / 2 + h
        fx, fy = math.floor(xm), math.floor(ym)
        col = fx * N + (N - 1 - fy)
        L = math.sqrt(float((xb - xa) ** 2 + (yb - ya) ** 2))
        out[col] = out.get(col, 0.0) + L
    return out
