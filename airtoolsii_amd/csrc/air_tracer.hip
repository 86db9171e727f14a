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
#include <cstdlib>
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
void parallel_detector(int p, double d, std::vector<double>& x0) {
    x0.assign((size_t)p, 0.0);
    double d1 = -d / 2.0, d2 = d / 2.0;
    if (p == 1) x0[0] = d2;
    else {
        double n1 = (double)(p - 1);
        for (int k = 0; k < p; ++k) x0[k] = d1 + ((double)k * (d2 - d1)) / n1;
        x0[0] = d1;
        x0[p - 1] = d2;
    }
}

void host_rays_paralleltomo(int N, const double* theta, int nA, int p, double d, std::vector<RayParam>& out) {
    (void)N;
    std::vector<double> x0;
    parallel_detector(p, d, x0);
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
// seg(x_u, y_u, x_v, y_v) for each pair of consecutive kept points.  The walk
// visits the x-lines kxa..kxb and the y-lines kya..kyb (line k at -N/2 + k)
// and keeps a crossing when its other coordinate lies in the window spanned by
// those lines; the whole grid (0..N, 0..N) is the reference's in-box test
// (paralleltomo.m:236-238).  A smaller window yields the points of the full
// sequence inside that rectangle, which is contiguous along the ray.
template <class Seg>
__device__ __forceinline__ void walk_window(const RayParam& r, int N, int kxa, int kxb, int kya, int kyb, bool fwd,
                                            Seg&& seg) {
    const double half = (double)N / 2.0;
    const double lo = -half;
    const double xlo = lo + (double)kxa, xhi = lo + (double)kxb;
    const double ylo = lo + (double)kya, yhi = lo + (double)kyb;
    // x-line crossings: t ascending along k iff a > 0; none when a is +-0.
    const bool hasXlist = (r.a > 0.0) || (r.a < 0.0);
    const bool hasYlist = (r.b > 0.0) || (r.b < 0.0);
    int kx = kxa, dkx = 1, nx = hasXlist ? kxb - kxa + 1 : 0;
    if (hasXlist && ((r.a > 0.0) != fwd)) { kx = kxb; dkx = -1; }
    int ky = kya, dky = 1, ny = hasYlist ? kyb - kya + 1 : 0;
    if (hasYlist && ((r.b > 0.0) != fwd)) { ky = kyb; dky = -1; }

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
            if (yx >= ylo && yx <= yhi) { X.t = tx; X.x = g; X.y = yx; hx = true; return; }
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
            if (xy >= xlo && xy <= xhi) { Y.t = ty; Y.x = xy; Y.y = g; hy = true; return; }
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

// The forward walk of walk_window (fwd = true) with one division per step and
// no inner scan loops, for the matrix-free projector.  Each list's accepted
// lines form one contiguous run (the other coordinate is monotone in k); an
// analytic estimate of that run, widened by 1 line on each side, bounds the
// candidates, and candidates outside the window are consumed in t order
// without being kept.  Merging the two sorted candidate lists and then
// dropping the rejected ones gives the same sequence as merging the accepted
// ones, so the kept points and segments equal walk_window's bit for bit (the
// crossings use the same operations).
__host__ __device__ inline void cand_range(double p0, double q0, double dp, double dq, double qlo, double qhi,
                                           double lo, int ka, int kb, int* k0, int* k1) {
    // lines p = lo + k whose crossing q = dq * ((p - p0) / dp) + q0 lies in [qlo, qhi]
    if (!(dq > 0.0 || dq < 0.0)) {                      // q constant
        if (q0 >= qlo && q0 <= qhi) { *k0 = ka; *k1 = kb; }
        else { *k0 = 1; *k1 = 0; }
        return;
    }
    const double r1 = p0 + dp * ((qlo - q0) / dq) - lo, r2 = p0 + dp * ((qhi - q0) / dq) - lo;
    double a = fmin(r1, r2), b = fmax(r1, r2);
    if (!(a == a) || !(b == b)) { *k0 = ka; *k1 = kb; return; }
    a = fmax(floor(a) - 1.0, (double)ka);
    b = fmin(ceil(b) + 1.0, (double)kb);
    *k0 = (int)a;
    *k1 = (int)b;
}

// seg(emit, x_u, y_u, x_v, y_v) is called on every step with the step's
// candidate segment; only steps with emit set carry one of the walk's
// segments.  Each step is straight-line code (selects, one division), so the
// lanes of a wave stay converged while they walk different rays.  A list's
// state is its current line k and that crossing's t; the crossing's other
// coordinate is formed when the candidate is taken (the same operations as
// walk_window).
template <class Seg>
__device__ __forceinline__ void walk_window_fwd(const RayParam& r, int N, int kxa, int kxb, int kya, int kyb,
                                                Seg&& seg) {
    const double half = (double)N / 2.0;
    const double lo = -half;
    const double xlo = lo + (double)kxa, xhi = lo + (double)kxb;
    const double ylo = lo + (double)kya, yhi = lo + (double)kyb;
    const bool hasX = (r.a > 0.0) || (r.a < 0.0);
    const bool hasY = (r.b > 0.0) || (r.b < 0.0);
    int x0k = 1, x1k = 0, y0k = 1, y1k = 0;
    if (hasX) cand_range(r.x0, r.y0, r.a, r.b, ylo, yhi, lo, kxa, kxb, &x0k, &x1k);
    if (hasY) cand_range(r.y0, r.x0, r.b, r.a, xlo, xhi, lo, kya, kyb, &y0k, &y1k);
    // t ascends with k when the direction component is positive
    int kx = r.a > 0.0 ? x0k : x1k, dkx = r.a > 0.0 ? 1 : -1, nx = x1k - x0k;
    int ky = r.b > 0.0 ? y0k : y1k, dky = r.b > 0.0 ? 1 : -1, ny = y1k - y0k;
    bool hx = nx >= 0, hy = ny >= 0;       // nx, ny: candidates left after the current one
    double Xt = hx ? (lo + (double)kx - r.x0) / r.a : 0.0;
    double Yt = hy ? (lo + (double)ky - r.y0) / r.b : 0.0;
    bool havePrev = false, haveQ = false;
    double px = 0, py = 0, qx = 0, qy = 0;
    while (hx || hy) {
        const bool sx = hx && (!hy || Xt <= Yt);
        // the taken crossing: x-line (g, b t + y0) or y-line (a t + x0, g)
        const int kc = sx ? kx : ky;
        const double tc = sx ? Xt : Yt;
        const double g = lo + (double)kc;
        const double o = (sx ? r.b : r.a) * tc + (sx ? r.y0 : r.x0);
        const bool ok = sx ? (o >= ylo && o <= yhi) : (o >= xlo && o <= xhi);
        const double x = sx ? g : o, y = sx ? o : g;
        // next candidate of that list: one division on the selected operands
        const bool more = sx ? nx > 0 : ny > 0;
        const int kn = kc + (sx ? dkx : dky);
        const double tn = (lo + (double)kn - (sx ? r.x0 : r.y0)) / (sx ? r.a : r.b);
        const bool upx = sx && more, upy = !sx && more;
        Xt = upx ? tn : Xt;
        Yt = upy ? tn : Yt;
        kx = upx ? kn : kx;
        ky = upy ? kn : ky;
        nx -= upx ? 1 : 0;
        ny -= upy ? 1 : 0;
        hx = sx ? more : hx;
        hy = sx ? hy : more;
        // point (x, y) follows (px, py): keep (px, py) unless the two coincide
        const bool coinc = fabs(x - px) <= 1e-10 && fabs(y - py) <= 1e-10;
        const bool keep = ok && havePrev && !coinc;
        seg(keep && haveQ, qx, qy, px, py);
        qx = keep ? px : qx;
        qy = keep ? py : qy;
        haveQ = haveQ || keep;
        havePrev = havePrev || ok;
        px = ok ? x : px;
        py = ok ? y : py;
    }
    seg(havePrev && haveQ, qx, qy, px, py);
}

template <class Seg>
__device__ __forceinline__ void walk_ray(const RayParam& r, int N, bool fwd, Seg&& seg) {
    walk_window(r, N, 0, N, 0, N, fwd, static_cast<Seg&&>(seg));
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
// paralleltomo.m:282-287 (forward aval'*u(col), adjoint A(col) += u_row*aval)
// with the products computed from the rays on every call.  The work is
// tile-major: the image is cut into T x T tiles and every ray is listed under
// the tiles it passes (a host binning with a half-pixel margin).  A
// work-group owns one tile; each lane walks one listed ray through the tile's
// window (the tile's grid lines plus one line on each side), which yields the
// points of the full walk near the tile and so the full walk's segments
// there, bit for bit, and keeps the segments whose midpoint pixel lies in the
// tile.  Each segment of a ray belongs to exactly one tile.  The forward
// product gathers x from the tile in LDS and writes one partial sum per
// (ray, tile) item, summed per ray afterwards in a fixed order; the adjoint
// accumulates into the tile in LDS (FP64) and stores it, so no global atomics
// are needed.

// window of tile (tx, ty): pixel columns fx in [tx T, tx T + T), image rows
// iy = N-1-fy in [ty T, ty T + T); x-lines tx T - 1 .. tx T + T + 1, y-lines
// N - ty T - T - 1 .. N - ty T + 1, clipped to the grid
struct TileWin {
    int fx0, iy0, kxa, kxb, kya, kyb;
};

__host__ __device__ inline TileWin tile_window(int tile, int nt, int T, int N) {
    TileWin w;
    const int tx = tile % nt, ty = tile / nt;
    w.fx0 = tx * T;
    w.iy0 = ty * T;
    w.kxa = w.fx0 - 1 > 0 ? w.fx0 - 1 : 0;
    w.kxb = w.fx0 + T + 1 < N ? w.fx0 + T + 1 : N;
    w.kya = N - w.iy0 - T - 1 > 0 ? N - w.iy0 - T - 1 : 0;
    w.kyb = N - w.iy0 + 1 < N ? N - w.iy0 + 1 : N;
    return w;
}

// steps of walk_window_fwd for one ray in one window: its candidate lines
__host__ __device__ inline int window_steps(const RayParam& r, int N, const TileWin& w) {
    const double lo = -(double)N / 2.0;
    const double xlo = lo + (double)w.kxa, xhi = lo + (double)w.kxb;
    const double ylo = lo + (double)w.kya, yhi = lo + (double)w.kyb;
    int x0k = 1, x1k = 0, y0k = 1, y1k = 0;
    if (r.a > 0.0 || r.a < 0.0) cand_range(r.x0, r.y0, r.a, r.b, ylo, yhi, lo, w.kxa, w.kxb, &x0k, &x1k);
    if (r.b > 0.0 || r.b < 0.0) cand_range(r.y0, r.x0, r.b, r.a, xlo, xhi, lo, w.kya, w.kyb, &y0k, &y1k);
    return (x1k >= x0k ? x1k - x0k + 1 : 0) + (y1k >= y0k ? y1k - y0k + 1 : 0);
}

// seg_value restricted to a tile: the segment's length and its slot in the
// tile (column-major, columns padded to T+1 slots so that one row of the tile
// spreads over the LDS banks), or -1 when its pixel (column
// floor(xm) N + N-1-floor(ym)) lies outside the tile or outside the image
template <int T>
__device__ __forceinline__ double tile_seg(int N, const TileWin& w, double xu, double yu, double xv, double yv,
                                           int* slot) {
    const double half = (double)N / 2.0;
    const double dx = xv - xu, dy = yv - yu;
    const double xm = 0.5 * (xu + xv) + half;
    const double ym = 0.5 * (yu + yv) + half;
    const int fx = (int)floor(xm), iy = N - 1 - (int)floor(ym);
    const int c = fx - w.fx0, rr = iy - w.iy0;
    *slot = ((unsigned)c < (unsigned)T && (unsigned)rr < (unsigned)T && fx < N && iy < N) ? c * (T + 1) + rr : -1;
    return sqrt(dx * dx + dy * dy);
}

// The DDA form of the tile walk: t steps from grid line to grid line with
// t = (line - x0) * (1/a) (no division), each segment's length is
// (t_b - t_a) |d| (no square root) and its cell comes from the midpoint
// (floor, the tracer's rule), so a segment falls in the same cell as the
// tracer's.  Segments with |dx| and |dy| <= 1e-10 are dropped as the tracer
// merges such points; the lengths then agree with the tracer's to rounding
// (about 1e-15 relative), not bit for bit.  seg(len, slot) gets every kept
// segment inside the tile.
struct RayAux {
    double ia, ib;      // 1/a, 1/b (0 when the component is 0)
    double dn, mab;     // |d|, max(|a|, |b|)
};

static RayAux ray_aux(const RayParam& r) {
    RayAux q;
    q.ia = (r.a > 0.0 || r.a < 0.0) ? 1.0 / r.a : 0.0;
    q.ib = (r.b > 0.0 || r.b < 0.0) ? 1.0 / r.b : 0.0;
    q.dn = std::sqrt(r.a * r.a + r.b * r.b);
    q.mab = std::max(std::fabs(r.a), std::fabs(r.b));
    return q;
}

template <int T, class Seg>
__device__ __forceinline__ void walk_tile_dda(const RayParam& r, const RayAux& q, int N, const TileWin& w, Seg&& seg) {
    const double half = (double)N / 2.0, lo = -half;
    const int fxe = min(w.fx0 + T, N), iye = min(w.iy0 + T, N);
    const double X0 = lo + (double)w.fx0, X1 = lo + (double)fxe;
    const double Y0 = lo + (double)(N - iye), Y1 = lo + (double)(N - w.iy0);
    const bool hx = (r.a > 0.0) || (r.a < 0.0), hy = (r.b > 0.0) || (r.b < 0.0);
    double tin = -INFINITY, tout = INFINITY;
    if (hx) {
        const double t1 = (X0 - r.x0) * q.ia, t2 = (X1 - r.x0) * q.ia;
        tin = fmax(tin, fmin(t1, t2));
        tout = fmin(tout, fmax(t1, t2));
    } else if (!(r.x0 >= X0 && r.x0 <= X1)) {
        return;
    }
    if (hy) {
        const double t1 = (Y0 - r.y0) * q.ib, t2 = (Y1 - r.y0) * q.ib;
        tin = fmax(tin, fmin(t1, t2));
        tout = fmin(tout, fmax(t1, t2));
    } else if (!(r.y0 >= Y0 && r.y0 <= Y1)) {
        return;
    }
    if (!(tin < tout)) return;
    // the first line of each family after the entry point
    int kx = 0, dkx = 0, ky = 0, dky = 0;
    double txn = INFINITY, tyn = INFINITY;
    if (hx) {
        dkx = r.a > 0.0 ? 1 : -1;
        const double xi = (r.x0 + r.a * tin) - lo;
        kx = r.a > 0.0 ? (int)floor(xi) + 1 : (int)ceil(xi) - 1;
        txn = ((lo + (double)kx) - r.x0) * q.ia;
        if (!(txn > tin)) { kx += dkx; txn = ((lo + (double)kx) - r.x0) * q.ia; }
    }
    if (hy) {
        dky = r.b > 0.0 ? 1 : -1;
        const double yi = (r.y0 + r.b * tin) - lo;
        ky = r.b > 0.0 ? (int)floor(yi) + 1 : (int)ceil(yi) - 1;
        tyn = ((lo + (double)ky) - r.y0) * q.ib;
        if (!(tyn > tin)) { ky += dky; tyn = ((lo + (double)ky) - r.y0) * q.ib; }
    }
    // A segment between consecutive crossings lies in the cell bounded by the
    // lines crossed so far: fx = kx - 1 (kx when moving left), fy likewise.
    // Its midpoint is at least half its extent along each axis, dt |a| / 2 and
    // dt |b| / 2, from every edge of that cell; with dt min(|a|,|b|) > 1e-9
    // that is 5e-10, far above the rounding of the crossing parameters (about
    // 1e-13), so the floor of the midpoint names that cell and the integer
    // cell is used.  Shorter segments (including the first one when the entry
    // lies within rounding of a line) take the midpoint rule.
    const double mib = fmin(fabs(r.a), fabs(r.b));
    double ta = tin;
    for (int guard = 0; guard < 2 * T + 8 && ta < tout; ++guard) {     // at most 2T+2 lines cross a tile
        const double tb = fmin(fmin(txn, tyn), tout);
        const double dt = tb - ta;
        if (dt * mib > 1e-9) {
            const int fx = dkx > 0 ? kx - 1 : kx;
            const int iy = N - 1 - (dky > 0 ? ky - 1 : ky);
            const int c = fx - w.fx0, rr = iy - w.iy0;
            if ((unsigned)c < (unsigned)T && (unsigned)rr < (unsigned)T && fx < N && iy < N)
                seg(dt * q.dn, c * (T + 1) + rr);
        } else if (dt * q.mab > 1e-10) {
            const double tm = 0.5 * (ta + tb);
            const int fx = (int)floor((r.x0 + r.a * tm) + half);
            const int iy = N - 1 - (int)floor((r.y0 + r.b * tm) + half);
            const int c = fx - w.fx0, rr = iy - w.iy0;
            if ((unsigned)c < (unsigned)T && (unsigned)rr < (unsigned)T && fx < N && iy < N)
                seg(dt * q.dn, c * (T + 1) + rr);
        }
        if (txn <= tb) { kx += dkx; txn = ((lo + (double)kx) - r.x0) * q.ia; }
        if (tyn <= tb) { ky += dky; tyn = ((lo + (double)ky) - r.y0) * q.ib; }
        ta = tb;
    }
}

// A tile's items are sorted by their step count (longest first) and dealt
// round-robin to `splits` work-groups: block b takes items part, part +
// splits, ... of tile b / splits (part = b % splits), so the lanes of a wave
// walk rays of about the same length and the blocks get equal shares.
__device__ __forceinline__ void tile_chunk(const int64_t* __restrict__ tile_ptr, int splits, int* tile,
                                           int64_t* q0, int64_t* q1) {
    const int t = (int)(blockIdx.x / (unsigned)splits), part = (int)(blockIdx.x % (unsigned)splits);
    *tile = t;
    *q0 = tile_ptr[t] + part;
    *q1 = tile_ptr[t + 1];
}

constexpr int MFT = 64;          // tile side in pixels
constexpr int MFT_BLOCK = 512;   // threads per work-group (6 waves per SIMD at <= 85 VGPRs)

template <bool EXACT>
__global__ __launch_bounds__(MFT_BLOCK, 6) void k_mft_forward(const RayParam* __restrict__ rays,
                                                             const RayAux* __restrict__ aux, int N, int nt,
                                                             int splits, const int64_t* __restrict__ tile_ptr,
                                                             const int* __restrict__ tile_rays,
                                                             const float* __restrict__ u,
                                                             float* __restrict__ partial) {
    __shared__ float us[MFT * (MFT + 1)];
    int tile;
    int64_t q0, q1;
    tile_chunk(tile_ptr, splits, &tile, &q0, &q1);
    const TileWin w = tile_window(tile, nt, MFT, N);
    for (int i = threadIdx.x; i < MFT * MFT; i += blockDim.x) {
        const int fx = w.fx0 + i / MFT, iy = w.iy0 + i % MFT;
        us[(i / MFT) * (MFT + 1) + i % MFT] = (fx < N && iy < N) ? u[(int64_t)fx * N + iy] : 0.f;
    }
    __syncthreads();
    const int64_t stride = (int64_t)splits * blockDim.x;
    for (int64_t q = q0 + (int64_t)splits * threadIdx.x; q < q1; q += stride) {
        const int ray = tile_rays[q];
        const RayParam r = rays[ray];
        float acc = 0.f;
        if (EXACT) {
            walk_window_fwd(r, N, w.kxa, w.kxb, w.kya, w.kyb,
                            [&](bool emit, double xu, double yu, double xv, double yv) {
                int k;
                const double val = tile_seg<MFT>(N, w, xu, yu, xv, yv, &k);
                const bool use = emit && k >= 0;
                const float xv32 = us[use ? k : 0];
                acc += use ? (float)val * xv32 : 0.f;
            });
        } else {
            walk_tile_dda<MFT>(r, aux[ray], N, w, [&](double len, int k) { acc += (float)len * us[k]; });
        }
        partial[q] = acc;
    }
}

// y[ray] = sum of the ray's partials in its item order (8 lanes per ray)
__global__ __launch_bounds__(256) void k_mft_raysum(const int64_t* __restrict__ ray_ptr,
                                                    const int* __restrict__ ray_items,
                                                    const float* __restrict__ partial, int64_t nrays,
                                                    float* __restrict__ y) {
    const int gl = threadIdx.x & 7;
    for (int64_t ray = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 3; ray < nrays;
         ray += ((int64_t)gridDim.x * blockDim.x) >> 3) {
        float a = 0.f;
        for (int64_t q = ray_ptr[ray] + gl; q < ray_ptr[ray + 1]; q += 8) a += partial[ray_items[q]];
        a += __shfl_xor(a, 4, 8);
        a += __shfl_xor(a, 2, 8);
        a += __shfl_xor(a, 1, 8);
        if (gl == 0) y[ray] = a;
    }
}

// One-time row statistics with the forward's tile walk: per (ray, tile) item
// the FP64 sums of a^2, |a| and a^2 w_j over the item's entries (a rounded
// to FP32, as the stored matrix holds it), then per ray the sum of its items
// in item order.  Entries are the forward's own (the DDA, or the crossing
// walk with AIR_MF_EXACT=1), so the weights belong to the operator applied.
template <bool EXACT>
__global__ __launch_bounds__(MFT_BLOCK, 4) void k_mft_rowstats(const RayParam* __restrict__ rays,
                                                              const RayAux* __restrict__ aux, int N, int nt,
                                                              int splits, const int64_t* __restrict__ tile_ptr,
                                                              const int* __restrict__ tile_rays,
                                                              const double* __restrict__ colw,
                                                              double* __restrict__ part, int64_t nitems) {
    __shared__ double ws[MFT * (MFT + 1)];
    int tile;
    int64_t q0, q1;
    tile_chunk(tile_ptr, splits, &tile, &q0, &q1);
    const TileWin w = tile_window(tile, nt, MFT, N);
    for (int i = threadIdx.x; i < MFT * MFT; i += blockDim.x) {
        const int fx = w.fx0 + i / MFT, iy = w.iy0 + i % MFT;
        ws[(i / MFT) * (MFT + 1) + i % MFT] = (colw && fx < N && iy < N) ? colw[(int64_t)fx * N + iy] : 0.0;
    }
    __syncthreads();
    const int64_t stride = (int64_t)splits * blockDim.x;
    for (int64_t q = q0 + (int64_t)splits * threadIdx.x; q < q1; q += stride) {
        const int ray = tile_rays[q];
        const RayParam r = rays[ray];
        double s2 = 0.0, s1 = 0.0, s3 = 0.0;
        auto add = [&](double len, int k) {
            const double a = (double)(float)len;
            s2 += a * a;
            s1 += fabs(a);
            s3 += (a * a) * ws[k];
        };
        if (EXACT) {
            walk_window_fwd(r, N, w.kxa, w.kxb, w.kya, w.kyb,
                            [&](bool emit, double xu, double yu, double xv, double yv) {
                int k;
                const double val = tile_seg<MFT>(N, w, xu, yu, xv, yv, &k);
                if (emit && k >= 0) add(val, k);
            });
        } else {
            walk_tile_dda<MFT>(r, aux[ray], N, w, add);
        }
        part[q] = s2;
        part[nitems + q] = s1;
        part[2 * nitems + q] = s3;
    }
}

__global__ __launch_bounds__(256) void k_mft_raysum3(const int64_t* __restrict__ ray_ptr,
                                                     const int* __restrict__ ray_items,
                                                     const double* __restrict__ part, int64_t nitems,
                                                     int64_t nrays, double* __restrict__ sq,
                                                     double* __restrict__ ab, double* __restrict__ sqw) {
    for (int64_t ray = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ray < nrays;
         ray += (int64_t)gridDim.x * blockDim.x) {
        double s2 = 0.0, s1 = 0.0, s3 = 0.0;
        for (int64_t q = ray_ptr[ray]; q < ray_ptr[ray + 1]; ++q) {
            const int it = ray_items[q];
            s2 += part[it];
            s1 += part[nitems + it];
            s3 += part[2 * nitems + it];
        }
        if (sq) sq[ray] = s2;
        if (ab) ab[ray] = s1;
        if (sqw) sqw[ray] = s3;
    }
}

// Adjoint of one chunk into the tile in LDS (FP64); the chunk's tile goes to
// slice (b % splits) of `out` (n values per slice), summed over the slices in
// order afterwards, or straight to g when the tile is not split.  It keeps the
// crossing walk: its cost is bound by the LDS atomic adds (one per segment),
// which a shorter walk step does not reduce.
__global__ __launch_bounds__(MFT_BLOCK, 6) void k_mft_adjoint(const RayParam* __restrict__ rays, int N, int nt,
                                                             int splits, const int64_t* __restrict__ tile_ptr,
                                                             const int* __restrict__ tile_rays,
                                                             const float* __restrict__ v,
                                                             double* __restrict__ out) {
    __shared__ double gs[MFT * (MFT + 1)];
    int tile;
    int64_t q0, q1;
    tile_chunk(tile_ptr, splits, &tile, &q0, &q1);
    const TileWin w = tile_window(tile, nt, MFT, N);
    for (int i = threadIdx.x; i < MFT * (MFT + 1); i += blockDim.x) gs[i] = 0.0;
    __syncthreads();
    const int64_t stride = (int64_t)splits * blockDim.x;
    for (int64_t q = q0 + (int64_t)splits * threadIdx.x; q < q1; q += stride) {
        const int ray = tile_rays[q];
        const float vi = v[ray];
        if (vi == 0.f) continue;
        const RayParam r = rays[ray];
        walk_window_fwd(r, N, w.kxa, w.kxb, w.kya, w.kyb,
                        [&](bool emit, double xu, double yu, double xv, double yv) {
            int k;
            const double val = tile_seg<MFT>(N, w, xu, yu, xv, yv, &k);
            if (emit && k >= 0) atomicAdd(&gs[k], (double)vi * (double)(float)val);
        });
    }
    __syncthreads();
    double* dst = out + (int64_t)(blockIdx.x % (unsigned)splits) * ((int64_t)N * N);
    for (int i = threadIdx.x; i < MFT * MFT; i += blockDim.x) {
        const int fx = w.fx0 + i / MFT, iy = w.iy0 + i % MFT;
        if (fx < N && iy < N) dst[(int64_t)fx * N + iy] = gs[(i / MFT) * (MFT + 1) + i % MFT];
    }
}

__global__ __launch_bounds__(256) void k_mft_sum_slices(const double* __restrict__ part, int splits, int64_t n,
                                                        double* __restrict__ g) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double a = part[i];
        for (int sl = 1; sl < splits; ++sl) a += part[(int64_t)sl * n + i];
        g[i] = a;
    }
}

// Pixel-driven transpose of the DDA forward for parallel beams.  Ray j of
// angle i is (c x_j, s x_j) + t (-s, c) (x_j the detector position), so it can
// meet a pixel only when |x_j - u| <= (|c| + |s|) / 2, u = c x + s y at the
// pixel's centre.  A thread owns one pixel and a chunk of angles; for each
// candidate ray it clips the pixel's box with the DDA's own crossing
// expressions ((line - x0) (1/a), (line - y0) (1/b)), so the entry and exit
// parameters are the DDA's segment ends bit for bit, and keeps the segment
// when it passes the same 1e-10 test and its midpoint cell (the same floor
// rule, as the range test fx <= z < fx + 1) is this pixel.  Each pixel sums
// its entries (rounded to FP32, as the stored matrix is) times v in FP64 in
// angle order: no atomics, and the result does not depend on scheduling.
// The candidate range is widened by 1e-6 detector units, far above the
// rounding of u and of the positions.
struct AngleAux {
    double a, b, ia, ib, dn, mab;   // direction (-s, c), its reciprocals, |d|, max(|a|, |b|)
    double mib;                     // min(|a|, |b|)
    double ci, si, k0, k1;          // candidate j in [ceil(ci x + si y + k0), floor(ci x + si y + k1)]
};

// STATS: the one-time column statistics instead (v taken as 1): out gets the
// column's sum of |a| and cnt its number of entries.
template <bool STATS>
__global__ __launch_bounds__(256) void k_mfp_adjoint(const AngleAux* __restrict__ ang,
                                                     const double* __restrict__ det, int N, int nA, int p,
                                                     int apc, int edge_check, const float* __restrict__ v,
                                                     double* __restrict__ out, int* __restrict__ cnt) {
    extern __shared__ double xs[];
    for (int j = threadIdx.x; j < p; j += blockDim.x) xs[j] = det[j];
    __syncthreads();
    const int64_t n = (int64_t)N * N;
    const int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= n) return;
    const int fx = (int)(pix / N), iy = (int)(pix - (int64_t)fx * N);
    const double half = (double)N / 2.0, lo = -half;
    const double Xa = lo + (double)fx, Xb = lo + (double)(fx + 1);
    const double Ya = lo + (double)(N - 1 - iy), Yb = lo + (double)(N - iy);
    const double cx = Xa + 0.5, cy = Ya + 0.5;
    // the midpoint cell test floor(z) == k as k <= z < k + 1
    const double mx0 = (double)fx, mx1 = (double)(fx + 1);
    const double my0 = (double)(N - 1 - iy), my1 = (double)(N - iy);
    const int i0 = (int)blockIdx.y * apc, i1 = min(nA, i0 + apc);
    double acc = 0.0;
    int nc = 0;
    for (int i = i0; i < i1; ++i) {
        const AngleAux q = ang[i];
        const double c = q.b, s = -q.a;
        const bool hx = (q.a > 0.0) || (q.a < 0.0), hy = (q.b > 0.0) || (q.b < 0.0);
        const double u = cx * q.ci + cy * q.si;
        const int j0 = max((int)ceil(u + q.k0), 0);
        const int j1 = min((int)floor(u + q.k1), p - 1);
        const float* vi = STATS ? nullptr : v + (int64_t)i * p;
        if (hx && hy) {
            // (X - x0) (1/a) is monotone in X, so min/max of the two box
            // lines are the lines taken in the order of the sign of 1/a
            const double XI = q.ia > 0.0 ? Xa : Xb, XO = q.ia > 0.0 ? Xb : Xa;
            const double YI = q.ib > 0.0 ? Ya : Yb, YO = q.ib > 0.0 ? Yb : Ya;
            for (int j = j0; j <= j1; ++j) {
                const float vj = STATS ? 1.f : vi[j];     // issued first: its latency overlaps the clip
                const double xj = xs[j];
                // the ray's start as the ray table holds it: c x_j - s 0 and s x_j + c 0
                const double rx0 = c * xj, ry0 = s * xj;
                const double tin = fmax((XI - rx0) * q.ia, (YI - ry0) * q.ib);
                const double tout = fmin((XO - rx0) * q.ia, (YO - ry0) * q.ib);
                const double dt = tout - tin;
                // The chord's ends lie on the box, so its midpoint is at least half
                // its extent along each axis, dt |a| / 2 and dt |b| / 2, from every
                // edge.  With dt min(|a|,|b|) > 1e-9 that is 5e-10, far above the
                // rounding of the DDA's midpoint (about 1e-13 here), so its cell is
                // this pixel and the 1e-10 test passes; shorter chords take both
                // tests as the DDA does (dt <= 0 fails the first).
                if (!(dt * q.mib > 1e-9)) {
                    if (!(dt * q.mab > 1e-10)) continue;
                    const double tm = 0.5 * (tin + tout);
                    const double zx = (rx0 + q.a * tm) + half, zy = (ry0 + q.b * tm) + half;
                    if (!(zx >= mx0 && zx < mx1 && zy >= my0 && zy < my1)) continue;
                }
                acc += (double)(float)(dt * q.dn) * (double)vj;
                if (STATS) ++nc;
            }
            continue;
        }
        for (int j = j0; j <= j1; ++j) {
            const double xj = xs[j];
            const double rx0 = c * xj, ry0 = s * xj;
            if (edge_check && ((!hy && fabs(ry0 - half) < 1e-15) || (!hx && fabs(rx0 - half) < 1e-15))) continue;
            double tin = -INFINITY, tout = INFINITY;
            if (hx) {
                const double t1 = (Xa - rx0) * q.ia, t2 = (Xb - rx0) * q.ia;
                tin = fmax(tin, fmin(t1, t2));
                tout = fmin(tout, fmax(t1, t2));
            } else if (!(rx0 >= Xa && rx0 <= Xb)) {
                continue;
            }
            if (hy) {
                const double t1 = (Ya - ry0) * q.ib, t2 = (Yb - ry0) * q.ib;
                tin = fmax(tin, fmin(t1, t2));
                tout = fmin(tout, fmax(t1, t2));
            } else if (!(ry0 >= Ya && ry0 <= Yb)) {
                continue;
            }
            if (!(tin < tout)) continue;
            const double dt = tout - tin;
            if (!(dt * q.mab > 1e-10)) continue;
            const double tm = 0.5 * (tin + tout);
            const double zx = (rx0 + q.a * tm) + half, zy = (ry0 + q.b * tm) + half;
            if (!(zx >= mx0 && zx < mx1 && zy >= my0 && zy < my1)) continue;
            acc += (double)(float)(dt * q.dn) * (STATS ? 1.0 : (double)vi[j]);
            if (STATS) ++nc;
        }
    }
    out[(int64_t)blockIdx.y * n + pix] = acc;
    if (STATS) cnt[(int64_t)blockIdx.y * n + pix] = nc;
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

// Host binning of the rays into image tiles: a ray is listed under every tile
// that its line passes within half a pixel of (a superset of the tiles that
// hold its segments).  Rays are visited in order, so each tile lists its rays
// ascending; the forward sums then follow each ray's own tile order.
void mf_build_tiles(Geometry& g, hipStream_t s) {
    const int N = g.N;
    const int T = MFT;
    const int nt = (N + T - 1) / T;
    const int64_t ntiles = (int64_t)nt * nt;
    const double half = (double)N / 2.0, lo = -half, mg = 0.5;
    auto fy_of = [&](double y) { return (int)std::min<double>(N - 1, std::max<double>(0, std::floor(y + half))); };
    auto fx_of = fy_of;
    auto for_tiles = [&](const RayParam& r, auto&& f) {
        const bool excl = g.edge_check && ((r.b == 0.0 && std::fabs(r.y0 - half) < 1e-15) ||
                                           (r.a == 0.0 && std::fabs(r.x0 - half) < 1e-15));
        if (excl || !(std::fabs(r.a) > 0 || std::fabs(r.b) > 0)) return;
        if (std::fabs(r.a) >= std::fabs(r.b)) {          // mostly along x: visit tile columns
            for (int tx = 0; tx < nt; ++tx) {
                const double xa = lo + (double)(tx * T) - mg, xb = lo + (double)std::min((tx + 1) * T, N) + mg;
                const double ya = r.y0 + r.b * ((xa - r.x0) / r.a), yb = r.y0 + r.b * ((xb - r.x0) / r.a);
                const double y0 = std::min(ya, yb) - mg, y1 = std::max(ya, yb) + mg;
                if (y1 < lo || y0 > half) continue;
                const int ty0 = (N - 1 - fy_of(y1)) / T, ty1 = (N - 1 - fy_of(y0)) / T;
                for (int ty = ty0; ty <= ty1; ++ty) f(ty * nt + tx);
            }
        } else {                                          // mostly along y: visit tile rows
            for (int ty = 0; ty < nt; ++ty) {
                const double ya = lo + (double)(N - std::min((ty + 1) * T, N)) - mg, yb = lo + (double)(N - ty * T) + mg;
                const double xa = r.x0 + r.a * ((ya - r.y0) / r.b), xb = r.x0 + r.a * ((yb - r.y0) / r.b);
                const double x0 = std::min(xa, xb) - mg, x1 = std::max(xa, xb) + mg;
                if (x1 < lo || x0 > half) continue;
                const int tx0 = fx_of(x0) / T, tx1 = fx_of(x1) / T;
                for (int tx = tx0; tx <= tx1; ++tx) f(ty * nt + tx);
            }
        }
    };
    std::vector<int64_t> tptr((size_t)ntiles + 1, 0), rptr((size_t)g.nrays + 1, 0);
    for (int64_t i = 0; i < g.nrays; ++i) {
        int64_t c = 0;
        for_tiles(g.host_rays[(size_t)i], [&](int64_t t) { tptr[(size_t)t + 1]++; c++; });
        rptr[(size_t)i + 1] = rptr[(size_t)i] + c;
    }
    for (int64_t t = 0; t < ntiles; ++t) tptr[(size_t)t + 1] += tptr[(size_t)t];
    const int64_t nitems = tptr[(size_t)ntiles];
    AIR_CHECK(nitems < INT32_MAX, AIR_E_UNSUPPORTED, "matrix-free projector: too many ray-tile items");
    std::vector<int> trays((size_t)nitems), ritems((size_t)nitems);
    std::vector<int64_t> fill(tptr.begin(), tptr.end() - 1);
    for (int64_t i = 0; i < g.nrays; ++i) {
        int64_t c = rptr[(size_t)i];
        for_tiles(g.host_rays[(size_t)i], [&](int64_t t) {
            const int64_t q = fill[(size_t)t]++;
            trays[(size_t)q] = (int)i;
            ritems[(size_t)c++] = (int)q;
        });
    }
    // sort each tile's items by step count, longest first (ties by ray)
    {
        std::vector<int> newpos((size_t)nitems), key;
        std::vector<int64_t> ord;
        std::vector<int> tmp;
        for (int64_t t = 0; t < ntiles; ++t) {
            const int64_t p0 = tptr[(size_t)t], len = tptr[(size_t)t + 1] - p0;
            const TileWin w = tile_window((int)t, nt, T, N);
            key.resize((size_t)len);
            ord.resize((size_t)len);
            for (int64_t j = 0; j < len; ++j) {
                key[(size_t)j] = window_steps(g.host_rays[(size_t)trays[(size_t)(p0 + j)]], N, w);
                ord[(size_t)j] = j;
            }
            std::stable_sort(ord.begin(), ord.end(), [&](int64_t u, int64_t v) { return key[(size_t)u] > key[(size_t)v]; });
            tmp.assign(trays.begin() + p0, trays.begin() + p0 + len);
            for (int64_t j = 0; j < len; ++j) {
                trays[(size_t)(p0 + j)] = tmp[(size_t)ord[(size_t)j]];
                newpos[(size_t)(p0 + ord[(size_t)j])] = (int)(p0 + j);
            }
        }
        for (auto& q : ritems) q = newpos[(size_t)q];
    }
    g.T = T;
    g.ntiles_side = nt;
    g.nitems = nitems;
    auto put = [&](DevBuf& d, const void* h, size_t bytes) {
        d.alloc(bytes);
        if (bytes) AIR_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, s));
    };
    // per-ray reciprocals, |d| and max(|a|, |b|) for the DDA walk; AIR_MF_EXACT=1
    // selects the crossing walk whose entries equal the tracer's bit for bit
    std::vector<RayAux> aux((size_t)std::max<int64_t>(g.nrays, 1));
    for (int64_t i = 0; i < g.nrays; ++i) aux[(size_t)i] = ray_aux(g.host_rays[(size_t)i]);
    put(g.raux, aux.data(), sizeof(RayAux) * aux.size());
    {
        const char* ex = std::getenv("AIR_MF_EXACT");
        g.mf_exact = ex && ex[0] == '1';
    }
    put(g.tile_ptr, tptr.data(), 8 * tptr.size());
    put(g.ray_ptr, rptr.data(), 8 * rptr.size());
    put(g.tile_rays, trays.data(), 4 * trays.size());
    put(g.ray_items, ritems.data(), 4 * ritems.size());
    g.partial.alloc(4 * (size_t)std::max<int64_t>(nitems, 1));
    // enough work-groups for 4 per CU (AIR_MF_SPLITS_PER_CU=k for k), each with at least 1024 items
    int dev = 0, cus = 0;
    AIR_HIP(hipGetDevice(&dev));
    AIR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int64_t per_cu = 4;   // paralleltomo N=512, 360 angles, p=725: 3 -> 2249, 4 -> 2303, 6 -> 2279, 8 -> 2218 iter/s
    if (const char* sp = std::getenv("AIR_MF_SPLITS_PER_CU")) per_cu = std::min(16, std::max(1, std::atoi(sp)));
    const int64_t by_cu = (per_cu * (int64_t)std::max(cus, 1) + ntiles - 1) / ntiles;
    const int64_t by_items = std::max<int64_t>(1, nitems / (ntiles * 1024));
    g.splits = (int)std::max<int64_t>(1, std::min<int64_t>({by_cu, by_items, 64}));
    if (g.splits > 1) g.gpart.alloc(8 * (size_t)g.splits * (size_t)N * (size_t)N);
    AIR_HIP(hipStreamSynchronize(s));
}

// Tables for the pixel-driven adjoint of a parallel-beam geometry: per angle
// the direction and the DDA's reciprocals (the ray table's own values), the
// detector positions, and the angle chunks (enough threads for about 32 waves
// per SIMD: 8 chunks at N = 512, measured 1.9 % faster than 4 and level with 16;
// AIR_MF_PIXEL_CHUNKS=c fixes the count, at most 16).  AIR_MF_PIXEL=0
// keeps the tile adjoint.
void mf_prepare_parallel(Geometry& g, int nA, int p, double d, hipStream_t s) {
    g.pnA = 0;
    const char* env = std::getenv("AIR_MF_PIXEL");
    if (env && env[0] == '0') return;
    if (g.mf_exact || nA < 1 || p < 2 || p > 6144 || !(d > 0.0) || g.nrays != (int64_t)nA * p) return;
    std::vector<double> x0;
    parallel_detector(p, d, x0);
    std::vector<AngleAux> ang((size_t)nA);
    for (int i = 0; i < nA; ++i) {
        const RayParam& r = g.host_rays[(size_t)i * p];
        const RayAux q = ray_aux(r);
        const double c = r.b, sn = -r.a, idx = (double)(p - 1) / (x0[(size_t)p - 1] - x0[0]);
        const double h = 0.5 * (std::fabs(c) + std::fabs(sn)) + 1e-6;
        ang[(size_t)i] = AngleAux{r.a,      r.b,       q.ia,                  q.ib,
                                  q.dn,     q.mab,     std::min(std::fabs(r.a), std::fabs(r.b)),
                                  c * idx,  sn * idx,  (-h - x0[0]) * idx,    (h - x0[0]) * idx};
    }
    g.pang.alloc(sizeof(AngleAux) * ang.size());
    g.pdet.alloc(8 * x0.size());
    AIR_HIP(hipMemcpyAsync(g.pang.p, ang.data(), sizeof(AngleAux) * ang.size(), hipMemcpyHostToDevice, s));
    AIR_HIP(hipMemcpyAsync(g.pdet.p, x0.data(), 8 * x0.size(), hipMemcpyHostToDevice, s));
    const int64_t n = (int64_t)g.N * g.N;
    const int64_t waves = (n + 63) / 64;
    int dev = 0, cus = 0;
    AIR_HIP(hipGetDevice(&dev));
    AIR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int64_t want = 32 * 4 * (int64_t)std::max(cus, 1);
    int64_t pick = (want + waves - 1) / waves;
    if (const char* ce = std::getenv("AIR_MF_PIXEL_CHUNKS")) pick = std::max(1, std::atoi(ce));
    const int chunks = (int)std::max<int64_t>(1, std::min<int64_t>({pick, 16, nA}));
    g.papc = (nA + chunks - 1) / chunks;
    g.pchunks = (nA + g.papc - 1) / g.papc;
    if (g.pchunks > 1) g.ppart.alloc(8 * (size_t)g.pchunks * (size_t)n);
    AIR_HIP(hipStreamSynchronize(s));
    g.pnA = nA;
    g.pp = p;
}

void mf_forward(const Geometry& g, const float* u, float* y, hipStream_t s) {
    if (g.model == 1) { fresnel_mf_forward(g, u, y, s); return; }
    if (g.model == 2) { circles_mf_forward(g, u, y, s); return; }
    if (g.nrays == 0) return;
    AIR_CHECK(g.T == MFT && g.splits >= 1, AIR_E_STATE, "matrix-free projector: tiles not built");
    const unsigned nblk = (unsigned)((int64_t)g.ntiles_side * g.ntiles_side * g.splits);
    auto kf = g.mf_exact ? k_mft_forward<true> : k_mft_forward<false>;
    kf<<<nblk, MFT_BLOCK, 0, s>>>(g.rays.as<RayParam>(), g.raux.as<RayAux>(), g.N, g.ntiles_side, g.splits,
                                  g.tile_ptr.as<int64_t>(), g.tile_rays.as<int>(), u, g.partial.as<float>());
    AIR_HIP(hipGetLastError());
    const unsigned grid = (unsigned)std::min<int64_t>((g.nrays * 8 + 255) / 256, 65536);
    k_mft_raysum<<<grid, 256, 0, s>>>(g.ray_ptr.as<int64_t>(), g.ray_items.as<int>(), g.partial.as<float>(),
                                      g.nrays, y);
    AIR_HIP(hipGetLastError());
}

// g = A' v (every pixel written: the tiles cover the image)
void mf_adjoint(const Geometry& g, const float* v, double* gacc, hipStream_t s) {
    if (g.model == 1) { fresnel_mf_adjoint(g, v, gacc, s); return; }
    if (g.model == 2) { circles_mf_adjoint(g, v, gacc, s); return; }
    if (g.nrays == 0) return;
    if (g.pnA > 0) {
        const int64_t n = (int64_t)g.N * g.N;
        double* out = g.pchunks > 1 ? g.ppart.as<double>() : gacc;
        dim3 grid((unsigned)((n + 255) / 256), (unsigned)g.pchunks);
        k_mfp_adjoint<false><<<grid, 256, 8 * (size_t)g.pp, s>>>(g.pang.as<AngleAux>(), g.pdet.as<double>(), g.N,
                                                                 g.pnA, g.pp, g.papc, g.edge_check, v, out, nullptr);
        AIR_HIP(hipGetLastError());
        if (g.pchunks > 1) {
            k_mft_sum_slices<<<(unsigned)std::min<int64_t>((n + 255) / 256, 8192), 256, 0, s>>>(
                g.ppart.as<double>(), g.pchunks, n, gacc);
            AIR_HIP(hipGetLastError());
        }
        return;
    }
    AIR_CHECK(g.T == MFT && g.splits >= 1, AIR_E_STATE, "matrix-free projector: tiles not built");
    const unsigned nblk = (unsigned)((int64_t)g.ntiles_side * g.ntiles_side * g.splits);
    double* out = g.splits > 1 ? g.gpart.as<double>() : gacc;
    k_mft_adjoint<<<nblk, MFT_BLOCK, 0, s>>>(g.rays.as<RayParam>(), g.N, g.ntiles_side, g.splits,
                                             g.tile_ptr.as<int64_t>(), g.tile_rays.as<int>(), v, out);
    AIR_HIP(hipGetLastError());
    if (g.splits > 1) {
        const int64_t n = (int64_t)g.N * g.N;
        k_mft_sum_slices<<<(unsigned)std::min<int64_t>((n + 255) / 256, 8192), 256, 0, s>>>(
            g.gpart.as<double>(), g.splits, n, gacc);
        AIR_HIP(hipGetLastError());
    }
}

void mf_row_stats(const Geometry& g, const double* colw, double* sq, double* ab, double* sqw, hipStream_t s) {
    if (g.model == 1) { fresnel_mf_row_stats(g, colw, sq, ab, sqw, s); return; }
    if (g.model == 2) { circles_mf_row_stats(g, colw, sq, ab, sqw, s); return; }
    if (g.nrays == 0) return;
    if (g.T == MFT && g.splits >= 1 && g.nitems > 0) {
        DevBuf part;
        part.alloc(3 * 8 * (size_t)g.nitems);
        const unsigned nblk = (unsigned)((int64_t)g.ntiles_side * g.ntiles_side * g.splits);
        auto kr = g.mf_exact ? k_mft_rowstats<true> : k_mft_rowstats<false>;
        kr<<<nblk, MFT_BLOCK, 0, s>>>(g.rays.as<RayParam>(), g.raux.as<RayAux>(), g.N, g.ntiles_side, g.splits,
                                      g.tile_ptr.as<int64_t>(), g.tile_rays.as<int>(), colw, part.as<double>(),
                                      g.nitems);
        AIR_HIP(hipGetLastError());
        k_mft_raysum3<<<(unsigned)std::min<int64_t>((g.nrays + 255) / 256, 65536), 256, 0, s>>>(
            g.ray_ptr.as<int64_t>(), g.ray_items.as<int>(), part.as<double>(), g.nitems, g.nrays, sq, ab, sqw);
        AIR_HIP(hipGetLastError());
        AIR_HIP(hipStreamSynchronize(s));   // part is freed on return
        return;
    }
    k_mf_row_stats<<<(unsigned)((g.nrays + 255) / 256), 256, 0, s>>>(g.rays.as<RayParam>(), g.nrays, g.N,
                                                                     g.edge_check, colw, sq, ab, sqw);
    AIR_HIP(hipGetLastError());
}
void mf_col_stats(const Geometry& g, double* colabs, int* colcnt, hipStream_t s) {
    if (g.model == 1) { fresnel_mf_col_stats(g, colabs, colcnt, s); return; }
    if (g.model == 2) { circles_mf_col_stats(g, colabs, colcnt, s); return; }
    if (g.nrays == 0) return;
    if (g.pnA > 0 && colabs && colcnt) {   // parallel beams: a thread per pixel, all angles, no atomics
        const int64_t n = (int64_t)g.N * g.N;
        k_mfp_adjoint<true><<<dim3((unsigned)((n + 255) / 256), 1), 256, 8 * (size_t)g.pp, s>>>(
            g.pang.as<AngleAux>(), g.pdet.as<double>(), g.N, g.pnA, g.pp, g.pnA, g.edge_check, nullptr, colabs,
            colcnt);
        AIR_HIP(hipGetLastError());
        return;
    }
    k_mf_col_stats<<<(unsigned)((g.nrays + 255) / 256), 256, 0, s>>>(g.rays.as<RayParam>(), g.nrays, g.N,
                                                                     g.edge_check, colabs, colcnt);
    AIR_HIP(hipGetLastError());
}

}  // namespace air
