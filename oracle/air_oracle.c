/*
 * air_oracle.c -- FP64 CPU restatement of the AIR Tools II SIRT path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke test and
 * bench.py's cpu_baseline leg load this library; the product never does.
 *
 * It follows the reference MATLAB sources literally (sort-based ray tracer,
 * row-by-row sparse assembly, the sirt.m loop) with the MATLAB built-in
 * semantics written out from their documented behaviour:
 *   - sort: ascending, stable, NaN last;
 *   - sparse(): duplicates summed, CSC/CSR with ascending indices;
 *   - every elementwise op rounded on its own (compile with -ffp-contract=off);
 *   - cosd/sind exact at multiples of 90 degrees;
 *   - rng(seed)+rand: MT19937 init_genrand(seed), seed 0 -> 5489, res53 doubles.
 */
#include "air_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* MATLAB built-in semantics                                           */
/* ------------------------------------------------------------------ */

/* Degree sine/cosine: reduce to the nearest multiple of 90 degrees (exact in
 * binary64), evaluate on the remainder in [-45,45] degrees, pick the quadrant.
 * Exact 0 and +-1 at multiples of 90. */
static void ora_sincosd(double deg, double* s, double* c)
{
    if (!isfinite(deg)) { *s = NAN; *c = NAN; return; }
    double r = fmod(deg, 360.0);             /* exact */
    double q = nearbyint(r / 90.0);          /* quadrant count, |q| <= 4 */
    double z = r - 90.0 * q;                 /* exact, |z| <= 45 */
    int iq = ((int)q % 4 + 4) % 4;
    double zr = z * (M_PI / 180.0);
    double sz = sin(zr), cz = cos(zr);
    if (z == 0.0) { sz = 0.0; cz = 1.0; }
    switch (iq) {
    case 0: *s = sz;  *c = cz;  break;
    case 1: *s = cz;  *c = -sz; break;
    case 2: *s = -sz; *c = -cz; break;
    default: *s = -cz; *c = sz; break;
    }
    if (*s == 0.0) *s = 0.0;                 /* exact zeros are +0 */
    if (*c == 0.0) *c = 0.0;
}
double ora_sind(double deg) { double s, c; ora_sincosd(deg, &s, &c); return s; }
double ora_cosd(double deg) { double s, c; ora_sincosd(deg, &s, &c); return c; }
double ora_atand(double x) { return (180.0 / M_PI) * atan(x); }

/* linspace(d1,d2,n): d1 + (k*(d2-d1))/(n-1), end points exact. */
void ora_linspace(double d1, double d2, int64_t n, double* out)
{
    if (n <= 0) return;
    if (n == 1) { out[0] = d2; return; }
    double n1 = (double)(n - 1);
    for (int64_t k = 0; k < n; ++k) out[k] = d1 + ((double)k * (d2 - d1)) / n1;
    out[0] = d1;
    out[n - 1] = d2;
}

double ora_round(double x) { return x < 0 ? -floor(-x + 0.5) : floor(x + 0.5); }

/* MT19937, Matsumoto & Nishimura (1998), init_genrand + genrand_res53. */
void ora_mt_seed(ora_mt* s, uint32_t matlab_seed)
{
    uint32_t seed = matlab_seed == 0 ? 5489u : matlab_seed;
    s->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        s->mt[i] = 1812433253u * (s->mt[i - 1] ^ (s->mt[i - 1] >> 30)) + (uint32_t)i;
    s->mti = 624;
}
static uint32_t ora_mt_u32(ora_mt* s)
{
    if (s->mti >= 624) {
        for (int i = 0; i < 624; ++i) {
            uint32_t y = (s->mt[i] & 0x80000000u) | (s->mt[(i + 1) % 624] & 0x7fffffffu);
            s->mt[i] = s->mt[(i + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
        }
        s->mti = 0;
    }
    uint32_t y = s->mt[s->mti++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}
double ora_mt_rand(ora_mt* s)
{
    uint32_t a = ora_mt_u32(s) >> 5, b = ora_mt_u32(s) >> 6;
    return ((double)a * 67108864.0 + (double)b) * (1.0 / 9007199254740992.0);
}

/* ------------------------------------------------------------------ */
/* Phantoms, testprobs/phantomgallery.m                                */
/* ------------------------------------------------------------------ */

/* myphantom, phantomgallery.m:155-214.  Xn(r,c) = xn(c); Yn = rot90(Xn) so
 * Yn(r,c) = xn(N-1-r) (0-based). */
void ora_phantom_shepplogan(int N, double* im)
{
    static const double e[10][6] = {
        {1, .69, .92, 0, 0, 0},         {-.8, .6624, .8740, 0, -.0184, 0},
        {-.2, .1100, .3100, .22, 0, -18}, {-.2, .1600, .4100, -.22, 0, 18},
        {.1, .2100, .2500, 0, .35, 0},  {.1, .0460, .0460, 0, .1, 0},
        {.1, .0460, .0460, 0, -.1, 0},  {.1, .0460, .0230, -.08, -.605, 0},
        {.1, .0230, .0230, 0, -.606, 0}, {.1, .0230, .0460, .06, -.605, 0}};
    double* xn = (double*)malloc(sizeof(double) * N);
    for (int k = 0; k < N; ++k) xn[k] = ((double)k - (N - 1) / 2.0) / ((N - 1) / 2.0);
    for (int64_t q = 0; q < (int64_t)N * N; ++q) im[q] = 0.0;
    for (int i = 0; i < 10; ++i) {
        double a2 = e[i][1] * e[i][1], b2 = e[i][2] * e[i][2];
        double x0 = e[i][3], y0 = e[i][4], phi = e[i][5] * M_PI / 180.0, A = e[i][0];
        double cp = cos(phi), sp = sin(phi);
        for (int c = 0; c < N; ++c)
            for (int r = 0; r < N; ++r) {
                double x = xn[c] - x0, y = xn[N - 1 - r] - y0;
                double u = x * cp + y * sp, v = y * cp - x * sp;
                if ((u * u) / a2 + (v * v) / b2 <= 1.0) im[r + (int64_t)c * N] += A;
            }
    }
    for (int64_t q = 0; q < (int64_t)N * N; ++q) if (im[q] < 0) im[q] = 0.0;
    free(xn);
}

/* smooth, phantomgallery.m:216-232.  meshgrid(1:N): I(r,c)=c+1, J(r,c)=r+1. */
void ora_phantom_smooth(int N, int p, double* im)
{
    double sigma = 0.25 * N;
    double cc[4][2] = {{0.6 * N, 0.6 * N}, {0.5 * N, 0.3 * N}, {0.2 * N, 0.7 * N}, {0.8 * N, 0.2 * N}};
    double a[4] = {1, 0.5, 0.7, 0.9};
    for (int64_t q = 0; q < (int64_t)N * N; ++q) im[q] = 0.0;
    double s1 = (1.2 * sigma) * (1.2 * sigma), s2 = sigma * sigma;
    for (int i = 0; i < p && i < 4; ++i)
        for (int c = 0; c < N; ++c)
            for (int r = 0; r < N; ++r) {
                double I = c + 1, J = r + 1;
                double dI = I - cc[i][0], dJ = J - cc[i][1];
                im[r + (int64_t)c * N] = im[r + (int64_t)c * N] + a[i] * exp(-(dI * dI) / s1 - (dJ * dJ) / s2);
            }
    double mx = -INFINITY;
    for (int64_t q = 0; q < (int64_t)N * N; ++q) if (im[q] > mx) mx = im[q];
    for (int64_t q = 0; q < (int64_t)N * N; ++q) im[q] /= mx;
}

static double cube(double v) { return v * v * v; }

/* threephases, phantomgallery.m:257-292.  rand(p,2) fills column-major. */
void ora_phantom_threephases(int N, int p, int has_seed, uint32_t seed, double* im)
{
    ora_mt st; ora_mt_seed(&st, has_seed ? seed : 0);
    int64_t NN = (int64_t)N * N;
    double* c1 = (double*)malloc(sizeof(double) * 2 * p);
    double* c2 = (double*)malloc(sizeof(double) * 2 * p);
    double* im1 = (double*)calloc(NN, sizeof(double));
    double* im2 = (double*)calloc(NN, sizeof(double));
    double sigma1 = 0.025 * N, sigma2 = 0.025 * N;
    for (int k = 0; k < 2 * p; ++k) c1[k] = ora_mt_rand(&st) * N;
    double d1 = cube(2.5 * sigma1), d2 = cube(sigma1);
    for (int i = 0; i < p; ++i)
        for (int c = 0; c < N; ++c)
            for (int r = 0; r < N; ++r) {
                double I = c + 1, J = r + 1;
                im1[r + (int64_t)c * N] += exp(-cube(fabs(I - c1[i])) / d1 - cube(fabs(J - c1[p + i])) / d2);
            }
    for (int64_t q = 0; q < NN; ++q) im1[q] = im1[q] < 0.35 ? 0.0 : 2.0;
    for (int k = 0; k < 2 * p; ++k) c2[k] = ora_mt_rand(&st) * N;
    double e1 = (2 * sigma2) * (2 * sigma2), e2 = sigma2 * sigma2;
    for (int i = 0; i < p; ++i)
        for (int c = 0; c < N; ++c)
            for (int r = 0; r < N; ++r) {
                double I = c + 1, J = r + 1;
                double u = I - c2[i], v = J - c2[p + i];
                im2[r + (int64_t)c * N] += exp(-(u * u) / e1 - (v * v) / e2);
            }
    for (int64_t q = 0; q < NN; ++q) im2[q] = im2[q] < 0.55 ? 0.0 : 1.0;
    double mx = -INFINITY;
    for (int64_t q = 0; q < NN; ++q) {
        double v = im1[q] + im2[q];
        if (v == 3.0) v = 1.0;
        im[q] = v;
        if (v > mx) mx = v;
    }
    for (int64_t q = 0; q < NN; ++q) im[q] /= mx;
    free(c1); free(c2); free(im1); free(im2);
}

/* grains, phantomgallery.m:352-389.  Voronoi labels on an Nbig^2 grid,
 * min() ties go to the first centre. */
void ora_phantom_grains(int N, int numGrains, int has_seed, uint32_t seed, double* im)
{
    ora_mt st; ora_mt_seed(&st, has_seed ? seed : 0);
    if (numGrains <= 0) numGrains = (int)ora_round(3.0 * sqrt((double)N));
    int dN = (int)ora_round(N / 10.0);
    int Nbig = N + 2 * dN;
    double* xG = (double*)malloc(sizeof(double) * numGrains);
    double* yG = (double*)malloc(sizeof(double) * numGrains);
    for (int k = 0; k < numGrains; ++k) xG[k] = ceil(Nbig * ora_mt_rand(&st));
    for (int k = 0; k < numGrains; ++k) yG[k] = ceil(Nbig * ora_mt_rand(&st));
    double mx = -INFINITY;
    for (int c = 0; c < N; ++c)
        for (int r = 0; r < N; ++r) {
            /* big-grid (0-based) row r+dN, col c+dN; meshgrid X = col+1, Y = row+1 */
            double X = c + dN + 1, Y = r + dN + 1;
            int best = 0; double bd = INFINITY;
            for (int k = 0; k < numGrains; ++k) {
                double dx = X - xG[k], dy = Y - yG[k];
                double d = dx * dx + dy * dy;
                if (d < bd) { bd = d; best = k; }
            }
            im[r + (int64_t)c * N] = (double)best;   /* minIdx - 1 */
            if (best > mx) mx = best;
        }
    for (int64_t q = 0; q < (int64_t)N * N; ++q) im[q] /= mx;
    free(xG); free(yG);
}

/* CompHistTrain (phantomgallery.m:458-509) for the full 3x3 mask and two
 * categories: every inner voxel's window, ids counted. */
void ora_clique_hist(const uint8_t* Z, int nr, int nc, double* H)
{
    for (int q = 0; q < 512; ++q) H[q] = 0.0;
    for (int j = 0; j + 2 < nc; ++j)
        for (int i = 0; i + 2 < nr; ++i) {
            int id = 0, bit = 0;
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r, ++bit)
                    if (Z[(int64_t)(i + r) + (int64_t)(j + c) * nr]) id |= 1 << bit;
            H[id] += 1.0;
        }
}

/* One draw of sequential_simulation_full_clique: the ids (0-based) whose bits
 * match the known window entries w[0..8] (-1 unknown, column-major), in
 * ascending order as marg2id returns them; pdf = H/sum, cdf = cumsum/sum, the
 * first id with cdf > rand.  Sums run left to right. */
static int binary_draw(const double* H, const int* w, ora_mt* st, int64_t* ndraw)
{
    int ids[512], n = 0;
    for (int q = 0; q < 512; ++q) {
        int ok = 1;
        for (int b = 0; b < 9 && ok; ++b)
            if (w[b] >= 0 && ((q >> b) & 1) != w[b]) ok = 0;
        if (ok) ids[n++] = q;
    }
    double sh = 0.0;
    for (int k = 0; k < n; ++k) sh += H[ids[k]];
    double pdf[512], sp = 0.0;
    for (int k = 0; k < n; ++k) { pdf[k] = H[ids[k]] / sh; sp += pdf[k]; }
    const double r = ora_mt_rand(st);
    (*ndraw)++;
    double c = 0.0;
    for (int k = 0; k < n; ++k) {
        c += pdf[k];
        if (c / sp > r) return ids[k];
    }
    return -1;
}

int64_t ora_phantom_binary(int N, int has_seed, uint32_t seed, const uint8_t* ti, int tr, int tc, double* im)
{
    if (N < 3) return -1;
    ora_mt st; ora_mt_seed(&st, has_seed ? seed : 0);
    /* H = CompHistTrain(TI) + CompHistTrain(fliplr(TI)) + eps */
    double H[512], Hf[512];
    uint8_t* fl = (uint8_t*)malloc((size_t)tr * tc);
    for (int j = 0; j < tc; ++j)
        for (int i = 0; i < tr; ++i) fl[i + (int64_t)j * tr] = ti[i + (int64_t)(tc - 1 - j) * tr];
    ora_clique_hist(ti, tr, tc, H);
    ora_clique_hist(fl, tr, tc, Hf);
    free(fl);
    for (int q = 0; q < 512; ++q) H[q] = H[q] + Hf[q] + 2.220446049250313e-16;
    int* sim = (int*)malloc(sizeof(int) * (size_t)N * N);
#define SIM(i, j) sim[(int64_t)(i) + (int64_t)(j) * N]
    for (int64_t q = 0; q < (int64_t)N * N; ++q) sim[q] = -1;
    int64_t ndraw = 0;
    int w[9];
    /* the first pattern from the whole histogram: cdf = cumsum(H)/sum(H) */
    {
        double norm = 0.0;
        for (int q = 0; q < 512; ++q) norm += H[q];
        const double r = ora_mt_rand(&st);
        ndraw++;
        double c = 0.0;
        int id = -1;
        for (int q = 0; q < 512; ++q) {
            c += H[q];
            if (c / norm > r) { id = q; break; }
        }
        for (int b = 0; b < 9; ++b) SIM(b % 3, b / 3) = (id >> b) & 1;
    }
    /* rows 1..3, columns 4..N: the window's last column is new */
    for (int j = 3; j < N; ++j) {
        for (int b = 0; b < 9; ++b) w[b] = SIM(b % 3, j - 2 + b / 3);
        const int id = binary_draw(H, w, &st, &ndraw);
        for (int r = 0; r < 3; ++r) SIM(r, j) = (id >> (6 + r)) & 1;
    }
    for (int i = 3; i < N; ++i)
        for (int j = 2; j < N; ++j) {
            for (int b = 0; b < 9; ++b) w[b] = SIM(i - 2 + b % 3, j - 2 + b / 3);
            const int id = binary_draw(H, w, &st, &ndraw);
            if (j == 2) {                           /* left side: sim(i, 1:3) = n(end, :) */
                for (int c = 0; c < 3; ++c) SIM(i, c) = (id >> (2 + 3 * c)) & 1;
            } else {                                /* sim(i, j) = n(end, end) */
                SIM(i, j) = (id >> 8) & 1;
            }
        }
    for (int64_t q = 0; q < (int64_t)N * N; ++q) im[q] = (double)sim[q];
#undef SIM
    free(sim);
    return ndraw;
}

/* ------------------------------------------------------------------ */
/* Ray tracers                                                         */
/* ------------------------------------------------------------------ */

struct ora_csr {
    int64_t m, n, nnz, cap;
    int64_t* rowptr;
    int32_t* col;
    double* val;
    int error;
};

static ora_csr* csr_new(int64_t m, int64_t n)
{
    ora_csr* A = (ora_csr*)calloc(1, sizeof(ora_csr));
    A->m = m; A->n = n;
    A->rowptr = (int64_t*)calloc(m + 1, sizeof(int64_t));
    A->cap = 1024;
    A->col = (int32_t*)malloc(sizeof(int32_t) * A->cap);
    A->val = (double*)malloc(sizeof(double) * A->cap);
    return A;
}
static void csr_push(ora_csr* A, int32_t c, double v)
{
    if (A->nnz == A->cap) {
        A->cap *= 2;
        A->col = (int32_t*)realloc(A->col, sizeof(int32_t) * A->cap);
        A->val = (double*)realloc(A->val, sizeof(double) * A->cap);
    }
    A->col[A->nnz] = c; A->val[A->nnz] = v; A->nnz++;
}

/* MATLAB sort order: ascending, NaN last, stable (ties keep index order). */
static int t_before(double a, double b)          /* strict "a sorts before b" */
{
    if (isnan(a)) return 0;
    if (isnan(b)) return 1;
    return a < b;
}
static void stable_sort_idx(const double* t, int* idx, int* tmp, int L)
{
    for (int w = 1; w < L; w *= 2) {
        for (int lo = 0; lo < L; lo += 2 * w) {
            int mid = lo + w < L ? lo + w : L, hi = lo + 2 * w < L ? lo + 2 * w : L;
            int i = lo, j = mid, o = lo;
            while (i < mid && j < hi) tmp[o++] = t_before(t[idx[j]], t[idx[i]]) ? idx[j++] : idx[i++];
            while (i < mid) tmp[o++] = idx[i++];
            while (j < hi) tmp[o++] = idx[j++];
        }
        memcpy(idx, tmp, sizeof(int) * L);
    }
}

typedef struct {
    int N;
    double *gx, *t, *xxy, *yxy, *px, *py, *aval;
    int *idx, *tmp, *colv, *ord;
} ray_ws;

static void ws_init(ray_ws* w, int N)
{
    int L = 2 * N + 2;
    w->N = N;
    w->gx = (double*)malloc(sizeof(double) * (N + 1));
    for (int k = 0; k <= N; ++k) w->gx[k] = -N / 2.0 + k;        /* (-N/2:N/2)' */
    w->t = (double*)malloc(sizeof(double) * L);
    w->xxy = (double*)malloc(sizeof(double) * L);
    w->yxy = (double*)malloc(sizeof(double) * L);
    w->px = (double*)malloc(sizeof(double) * L);
    w->py = (double*)malloc(sizeof(double) * L);
    w->aval = (double*)malloc(sizeof(double) * L);
    w->idx = (int*)malloc(sizeof(int) * L);
    w->tmp = (int*)malloc(sizeof(int) * L);
    w->colv = (int*)malloc(sizeof(int) * L);
    w->ord = (int*)malloc(sizeof(int) * L);
}
static void ws_free(ray_ws* w)
{
    free(w->gx); free(w->t); free(w->xxy); free(w->yxy); free(w->px); free(w->py);
    free(w->aval); free(w->idx); free(w->tmp); free(w->colv); free(w->ord);
}

/* One ray, paralleltomo.m:203-277 / fanlineartomo.m:247-333.  Appends the
 * ray's entries (columns ascending, duplicates summed) as one CSR row. */
static void trace_ray(ray_ws* w, ora_csr* A, double x0t, double y0t, double a, double b,
                      int edge_check)
{
    int N = w->N, L = 2 * N + 2;
    double half = N / 2.0;
    /* tx = (x - x0theta)/a; yx = b*tx + y0theta; ty = (y - y0theta)/b; xy = a*ty + x0theta */
    for (int k = 0; k <= N; ++k) {
        double tx = (w->gx[k] - x0t) / a;
        double yx = b * tx + y0t;
        double ty = (w->gx[k] - y0t) / b;
        double xy = a * ty + x0t;
        w->t[k] = tx;          w->xxy[k] = w->gx[k]; w->yxy[k] = yx;
        w->t[N + 1 + k] = ty;  w->xxy[N + 1 + k] = xy; w->yxy[N + 1 + k] = w->gx[k];
    }
    for (int q = 0; q < L; ++q) w->idx[q] = q;
    stable_sort_idx(w->t, w->idx, w->tmp, L);
    /* inclusive box filter */
    int np = 0;
    for (int q = 0; q < L; ++q) {
        double xv = w->xxy[w->idx[q]], yv = w->yxy[w->idx[q]];
        if (xv >= -half && xv <= half && yv >= -half && yv <= half) { w->px[np] = xv; w->py[np] = yv; np++; }
    }
    /* drop point q when it coincides with point q+1 (mask one shorter) */
    int nk = 0;
    for (int q = 0; q < np; ++q) {
        int drop = 0;
        if (q + 1 < np) drop = fabs(w->px[q + 1] - w->px[q]) <= 1e-10 && fabs(w->py[q + 1] - w->py[q]) <= 1e-10;
        if (!drop) { w->px[nk] = w->px[q]; w->py[nk] = w->py[q]; nk++; }
    }
    int ns = nk - 1;
    if (ns <= 0) return;
    if (edge_check && ((b == 0 && fabs(y0t - half) < 1e-15) || (a == 0 && fabs(x0t - half) < 1e-15)))
        return;
    for (int s = 0; s < ns; ++s) {
        double dx = w->px[s + 1] - w->px[s], dy = w->py[s + 1] - w->py[s];
        w->aval[s] = sqrt(dx * dx + dy * dy);
        double xm = 0.5 * (w->px[s] + w->px[s + 1]) + half;
        double ym = 0.5 * (w->py[s] + w->py[s + 1]) + half;
        double c1 = floor(xm) * N + (N - floor(ym));     /* 1-based column */
        if (!(c1 >= 1 && c1 <= (double)N * N)) { A->error = 1; return; }
        w->colv[s] = (int)c1 - 1;
        w->ord[s] = s;
    }
    /* sparse(): order by column (stable insertion sort, rows are short) */
    for (int s = 1; s < ns; ++s) {
        int v = w->ord[s], q = s - 1;
        while (q >= 0 && w->colv[w->ord[q]] > w->colv[v]) { w->ord[q + 1] = w->ord[q]; q--; }
        w->ord[q + 1] = v;
    }
    int last = -1;
    for (int s = 0; s < ns; ++s) {
        int o = w->ord[s];
        if (w->colv[o] == last) A->val[A->nnz - 1] += w->aval[o];
        else { csr_push(A, w->colv[o], w->aval[o]); last = w->colv[o]; }
    }
}

/* Concatenate per-angle row blocks (traced independently) into one CSR. */
static ora_csr* csr_concat(ora_csr** blk, int nA, int p, int64_t n)
{
    ora_csr* A = csr_new((int64_t)p * nA, n);
    int64_t tot = 0;
    for (int i = 0; i < nA; ++i) tot += blk[i]->nnz;
    free(A->col); free(A->val);
    A->cap = tot > 0 ? tot : 1;
    A->col = (int32_t*)malloc(sizeof(int32_t) * A->cap);
    A->val = (double*)malloc(sizeof(double) * A->cap);
    int64_t row = 0;
    for (int i = 0; i < nA; ++i) {
        memcpy(A->col + A->nnz, blk[i]->col, sizeof(int32_t) * blk[i]->nnz);
        memcpy(A->val + A->nnz, blk[i]->val, sizeof(double) * blk[i]->nnz);
        for (int j = 0; j < p; ++j) A->rowptr[row + j + 1] = A->nnz + blk[i]->rowptr[j + 1];
        A->nnz += blk[i]->nnz;
        row += p;
        A->error |= blk[i]->error;
        ora_csr_free(blk[i]);
    }
    return A;
}

/* paralleltomo.m:112-301 (angles traced in parallel; row order preserved) */
ora_csr* ora_paralleltomo(int N, const double* theta, int nA, int p, double d)
{
    double* x0 = (double*)malloc(sizeof(double) * p);
    ora_linspace(-d / 2.0, d / 2.0, p, x0);
    ora_csr** blk = (ora_csr**)calloc(nA > 0 ? nA : 1, sizeof(ora_csr*));
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        ray_ws w; ws_init(&w, N);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int i = 0; i < nA; ++i) {
            ora_csr* B = csr_new(p, (int64_t)N * N);
            double s = ora_sind(theta[i]), c = ora_cosd(theta[i]);
            double a = -s, b = c;
            for (int j = 0; j < p; ++j) {
                double x0t = c * x0[j] - s * 0.0;          /* y0 = zeros(p,1) */
                double y0t = s * x0[j] + c * 0.0;
                trace_ray(&w, B, x0t, y0t, a, b, 1);
                B->rowptr[j + 1] = B->nnz;
            }
            blk[i] = B;
        }
        ws_free(&w);
    }
    ora_csr* A = csr_concat(blk, nA, p, (int64_t)N * N);
    free(blk); free(x0);
    return A;
}

/* fanlineartomo.m:141-358 (R, dw, sd in absolute units) */
ora_csr* ora_fanlineartomo(int N, const double* theta, int nA, int p, double R, double dw, double sd)
{
    double dew = dw / p;
    double* omega = (double*)malloc(sizeof(double) * p);
    if (p % 2) {
        int h = (p - 1) / 2;
        omega[h] = 0.0;
        for (int k = 1; k <= h; ++k) {
            double om = 90.0 - ora_atand(sd / ((double)k * dew));
            omega[h + k] = om; omega[h - k] = -om;
        }
    } else {
        int h = p / 2;
        for (int k = 1; k <= h; ++k) {
            double om = 90.0 - ora_atand(sd / (((double)k - 0.5) * dew));
            omega[h - 1 + k] = om; omega[h - k] = -om;
        }
    }
    ora_csr** blk = (ora_csr**)calloc(nA > 0 ? nA : 1, sizeof(ora_csr*));
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        ray_ws w; ws_init(&w, N);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int i = 0; i < nA; ++i) {
            ora_csr* B = csr_new(p, (int64_t)N * N);
            double st = ora_sind(theta[i]), ct = ora_cosd(theta[i]);
            /* Omega_x(theta)*[0;R], Omega_y(theta)*[0;R] */
            double x0t = ct * 0.0 + (-st) * R;
            double y0t = st * 0.0 + ct * R;
            double abx = (-x0t) / R, aby = (-y0t) / R;
            for (int j = 0; j < p; ++j) {
                double so = ora_sind(omega[j]), co = ora_cosd(omega[j]);
                double a = co * abx + (-so) * aby;
                double b = so * abx + co * aby;
                trace_ray(&w, B, x0t, y0t, a, b, 0);
                B->rowptr[j + 1] = B->nnz;
            }
            blk[i] = B;
        }
        ws_free(&w);
    }
    ora_csr* A = csr_concat(blk, nA, p, (int64_t)N * N);
    free(blk); free(omega);
    return A;
}

/* fancurvedtomo.m:140-339 (R in absolute units, d in degrees): the rays of one
 * source position are spaced by equal angles, omega = linspace(-d/2,d/2,p)
 * (:154), each turning the central direction -xytheta/R (:233-241). */
ora_csr* ora_fancurvedtomo(int N, const double* theta, int nA, int p, double R, double d)
{
    double* omega = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
    ora_linspace(-d / 2, d / 2, p, omega);
    ora_csr** blk = (ora_csr**)calloc(nA > 0 ? nA : 1, sizeof(ora_csr*));
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        ray_ws w; ws_init(&w, N);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int i = 0; i < nA; ++i) {
            ora_csr* B = csr_new(p, (int64_t)N * N);
            double st = ora_sind(theta[i]), ct = ora_cosd(theta[i]);
            double x0t = ct * 0.0 + (-st) * R;         /* Omega_x(theta)*xy0 */
            double y0t = st * 0.0 + ct * R;            /* Omega_y(theta)*xy0 */
            double abx = (-x0t) / R, aby = (-y0t) / R; /* -xytheta/R */
            for (int j = 0; j < p; ++j) {
                double so = ora_sind(omega[j]), co = ora_cosd(omega[j]);
                trace_ray(&w, B, x0t, y0t, co * abx + (-so) * aby, so * abx + co * aby, 0);
                B->rowptr[j + 1] = B->nnz;
            }
            blk[i] = B;
        }
        ws_free(&w);
    }
    ora_csr* A = csr_concat(blk, nA, p, (int64_t)N * N);
    free(blk); free(omega);
    return A;
}

/* seismictomo.m:96-287: sources (N/2, y0_i) on the right edge, receivers
 * (xp_j, yp_j) on the left edge (first floor(p/2)) and the top edge (the
 * other ceil(p/2)); ray (i,j) is the line through both, row i*p + j. */
ora_csr* ora_seismictomo(int N, int s, int p)
{
    double N2 = N / 2.0;
    double Ns = ((double)N / s) / 2;
    int p1 = (p + 1) / 2, p2 = p / 2;                   /* ceil(p/2), floor(p/2) */
    double Np1 = ((double)N / p1) / 2, Np2 = ((double)N / p2) / 2;
    double* y0 = (double*)malloc(sizeof(double) * (s > 0 ? s : 1));
    double* xp = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
    double* yp = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
    ora_linspace(-N2 + Ns, N2 - Ns, s, y0);
    for (int j = 0; j < p2; ++j) xp[j] = -N2;
    ora_linspace(-N2 + Np1, N2 - Np1, p1, xp + p2);
    ora_linspace(-N2 + Np2, N2 - Np2, p2, yp);
    for (int j = p2; j < p; ++j) yp[j] = N2;
    ora_csr** blk = (ora_csr**)calloc(s > 0 ? s : 1, sizeof(ora_csr*));
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        ray_ws w; ws_init(&w, N);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int i = 0; i < s; ++i) {
            ora_csr* B = csr_new(p, (int64_t)N * N);
            for (int j = 0; j < p; ++j) {
                trace_ray(&w, B, N2, y0[i], xp[j] - N2, yp[j] - y0[i], 0);
                B->rowptr[j + 1] = B->nnz;
            }
            blk[i] = B;
        }
        ws_free(&w);
    }
    ora_csr* A = csr_concat(blk, s, p, (int64_t)N * N);
    free(blk); free(y0); free(xp); free(yp);
    return A;
}

static int cmp_i32(const void* a, const void* b)
{
    int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* sphericaltomo.m:108-270: circle centres xi = (cosd a, sind a) on the unit
 * circle, numCircles concentric circles of radii linspace(0,2,nC+1)(2:end),
 * each sampled at nPhi+1 points k = t*dPhi (t = 0..nPhi, both ends of the
 * period kept, :203); every sample is rounded to a pixel (round half away
 * from zero), samples outside the image are dropped, and each pixel gets
 * (2 pi r/nPhi) times the number of samples in it (unique + histc, :227-234). */
ora_csr* ora_sphericaltomo(int N, const double* angles, int nA, int nC)
{
    double* lin = (double*)malloc(sizeof(double) * (nC + 1));
    ora_linspace(0.0, 2.0, nC + 1, lin);
    double* radii = lin + 1;
    double centerImg = ceil(N / 2.0);
    double dx = sqrt(2.0) / N;
    int64_t* nPhi = (int64_t*)malloc(sizeof(int64_t) * (nC > 0 ? nC : 1));
    double* dPhi = (double*)malloc(sizeof(double) * (nC > 0 ? nC : 1));
    int64_t maxpts = 1;
    for (int n = 0; n < nC; ++n) {
        nPhi[n] = (int64_t)ceil((4 * M_PI / dx) * radii[n]);
        dPhi[n] = 2 * M_PI / (double)nPhi[n];
        if (nPhi[n] + 1 > maxpts) maxpts = nPhi[n] + 1;
    }
    ora_csr** blk = (ora_csr**)calloc(nA > 0 ? nA : 1, sizeof(ora_csr*));
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        int32_t* J = (int32_t*)malloc(sizeof(int32_t) * maxpts);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int m = 0; m < nA; ++m) {
            ora_csr* B = csr_new(nC, (int64_t)N * N);
            double xix = ora_cosd(angles[m]), xiy = ora_sind(angles[m]);
            for (int n = 0; n < nC; ++n) {
                int64_t cnt = 0;
                for (int64_t t = 0; t <= nPhi[n]; ++t) {
                    double k = (double)t * dPhi[n];
                    double xx = (xix + radii[n] * cos(k)) / dx + centerImg;
                    double yy = (xiy + radii[n] * sin(k)) / dx + centerImg;
                    double col = round(xx), row = round(yy);
                    if (col > 0 && col <= N && row > 0 && row <= N)
                        J[cnt++] = (int32_t)((N - row + 1) + (col - 1) * N);
                }
                qsort(J, (size_t)cnt, sizeof(int32_t), cmp_i32);
                double base = 2 * M_PI * radii[n] / (double)nPhi[n];
                for (int64_t q = 0; q < cnt;) {
                    int64_t e = q;
                    while (e < cnt && J[e] == J[q]) ++e;
                    csr_push(B, J[q] - 1, base * (double)(e - q));
                    q = e;
                }
                B->rowptr[n + 1] = B->nnz;
            }
            blk[m] = B;
        }
        free(J);
    }
    ora_csr* A = csr_concat(blk, nA, nC, (int64_t)N * N);
    free(blk); free(lin); free(nPhi); free(dPhi);
    return A;
}

/* seismicwavetomo.m:114-288: row (i,j) holds the Fresnel-zone kernel of the
 * wave from source i to receiver j at every pixel centre, normalised to sum
 * |S-R| over the image and cut at 1e-6, pixels in column-major order. */
ora_csr* ora_seismicwavetomo(int N, int s, int p, double omega)
{
    const double tol = 1e-6, alpha = 10;
    double N2 = N / 2.0;
    double Ns = ((double)N / s) / 2;
    int p1 = (p + 1) / 2, p2 = p / 2;
    double Np1 = ((double)N / p1) / 2, Np2 = ((double)N / p2) / 2;
    double* y0 = (double*)malloc(sizeof(double) * (s > 0 ? s : 1));
    double* xp = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
    double* yp = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
    ora_linspace(-N2 + Ns, N2 - Ns, s, y0);
    for (int j = 0; j < p2; ++j) xp[j] = -N2;
    ora_linspace(-N2 + Np1, N2 - Np1, p1, xp + p2);
    ora_linspace(-N2 + Np2, N2 - Np2, p2, yp);
    for (int j = p2; j < p; ++j) yp[j] = N2;
    double w = omega / N;
    int64_t n2 = (int64_t)N * N;
    ora_csr** blk = (ora_csr**)calloc(s > 0 ? s : 1, sizeof(ora_csr*));
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        double* K = (double*)malloc(sizeof(double) * n2);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int i = 0; i < s; ++i) {
            ora_csr* B = csr_new(p, n2);
            double S1 = N2, S2 = y0[i];
            for (int j = 0; j < p; ++j) {
                double R1 = xp[j], R2 = yp[j];
                double dist = sqrt((S1 - R1) * (S1 - R1) + (S2 - R2) * (S2 - R2));
                double sum = 0;
                for (int c = 0; c < N; ++c) {            /* column-major: S(:) */
                    double x = (-N2 + .5) + c;
                    for (int r = 0; r < N; ++r) {
                        double y = (N2 - .5) - r;
                        double tSX = sqrt((x - S1) * (x - S1) + (y - S2) * (y - S2));
                        double tRX = sqrt((x - R1) * (x - R1) + (y - R2) * (y - R2));
                        double dt = tSX + tRX - dist;
                        double a = alpha * dt * w;
                        double k = cos(2 * M_PI * dt * w) * exp(-(a * a));
                        K[(int64_t)c * N + r] = k;
                        sum += k;
                    }
                }
                for (int64_t q = 0; q < n2; ++q) {
                    double v = dist * K[q] / sum;
                    if (!(v < tol)) csr_push(B, (int32_t)q, v);
                }
                B->rowptr[j + 1] = B->nnz;
            }
            blk[i] = B;
        }
        free(K);
    }
    ora_csr* A = csr_concat(blk, s, p, n2);
    free(blk); free(y0); free(xp); free(yp);
    return A;
}

/* ---------------------------------------------------------------- ART / CART */
uint64_t ora_splitmix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
double ora_rng_u01(uint64_t seed, uint64_t stream, uint64_t k, uint64_t t)
{
    uint64_t z = ora_splitmix64(ora_splitmix64(seed ^ stream) + (k << 32) + t);
    return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static void ora_project_all(double* x, int64_t n, int lbk, double lb, const double* lbv,
                            int ubk, double ub, const double* ubv)
{
    for (int64_t j = 0; j < n; ++j) {
        double v = x[j];
        if (lbk == 1) v = fmax(v, lb); else if (lbk == 2) v = fmax(v, lbv[j]);
        if (ubk == 1) v = fmin(v, ub); else if (ubk == 2) v = fmin(v, ubv[j]);
        x[j] = v;
    }
}

/* art.m:249-285, literally: every row step updates x along the row and then
 * projects the whole vector (max with lbound, min with ubound).  rows are
 * 0-based; lamv (or the scalar lam) per step. */
void ora_art_sweep(const int64_t* rp, const int32_t* ci, const double* v, const int32_t* rows, int64_t L,
                   const double* lamv, double lam, const double* b, const double* nrm, double* x, int64_t n,
                   int lbk, double lb, const double* lbv, int ubk, double ub, const double* ubv)
{
    for (int64_t t = 0; t < L; ++t) {
        int r = rows[t];
        double dot = 0;
        for (int64_t q = rp[r]; q < rp[r + 1]; ++q) dot += v[q] * x[ci[q]];
        double lk = lamv ? lamv[t] : lam;
        double coef = (lk * (b[r] - dot)) / nrm[r];
        for (int64_t q = rp[r]; q < rp[r + 1]; ++q) x[ci[q]] = x[ci[q]] + coef * v[q];
        if (lbk || ubk) ora_project_all(x, n, lbk, lb, lbv, ubk, ub, ubv);
    }
}

/* cart.m:226-287, literally (mm = max(abs(xk)) recomputed at every step,
 * NaN ignored); cp/ri/v the CSC of A; flag/nflag per column. */
void ora_cart_sweep(const int64_t* cp, const int32_t* ri, const double* v, const int32_t* cols, int64_t L,
                    const double* nrm, double* x, double* r, int64_t n, unsigned char* flag, int32_t* nflag,
                    int lbk, double lb, const double* lbv, int ubk, double ub, const double* ubv,
                    double lam, double thr, int k, int kbegin, int nunflag, uint64_t seed)
{
    for (int64_t t = 0; t < L; ++t) {
        int j = cols[t];
        double mm = NAN;
        for (int64_t q = 0; q < n; ++q) {
            double a = fabs(x[q]);
            if (!isnan(a) && (isnan(mm) || a > mm)) mm = a;
        }
        if (flag[j]) {
            double dot = 0;
            for (int64_t q = cp[j]; q < cp[j + 1]; ++q) dot += v[q] * r[ri[q]];
            double delta = dot / nrm[j];
            double od = lam * delta;
            double xkj = x[j];
            if (lbk) { double lbj = lbk == 1 ? lb : lbv[j]; if (od < lbj - xkj) od = lbj - xkj; }
            if (ubk) { double ubj = ubk == 1 ? ub : ubv[j]; if (od > ubj - xkj) od = ubj - xkj; }
            x[j] = xkj + od;
            if (k >= kbegin && fabs(od) < thr * mm) { flag[j] = 0; nflag[j] = 1; }
            for (int64_t q = cp[j]; q < cp[j + 1]; ++q) r[ri[q]] = r[ri[q]] - od * v[q];
        } else {
            int draw = 1 + (int)floor(ora_rng_u01(seed, 3, (uint64_t)k, (uint64_t)j) * nunflag);
            if (nflag[j] < draw) nflag[j] = nflag[j] + 1;
            else flag[j] = 1;
        }
    }
}

int64_t ora_csr_m(const ora_csr* A) { return A->m; }
int64_t ora_csr_n(const ora_csr* A) { return A->n; }
int64_t ora_csr_nnz(const ora_csr* A) { return A->nnz; }
int ora_csr_error(const ora_csr* A) { return A->error; }
void ora_csr_copy(const ora_csr* A, int64_t* rowptr, int32_t* col, double* val)
{
    memcpy(rowptr, A->rowptr, sizeof(int64_t) * (A->m + 1));
    memcpy(col, A->col, sizeof(int32_t) * A->nnz);
    memcpy(val, A->val, sizeof(double) * A->nnz);
}
void ora_csr_free(ora_csr* A)
{
    if (!A) return;
    free(A->rowptr); free(A->col); free(A->val); free(A);
}

void ora_csr_to_csc(int64_t m, int64_t n, const int64_t* rowptr, const int32_t* col,
                    const double* val, int64_t* colptr, int32_t* row, double* valc)
{
    memset(colptr, 0, sizeof(int64_t) * (n + 1));
    for (int64_t e = 0; e < rowptr[m]; ++e) colptr[col[e] + 1]++;
    for (int64_t j = 0; j < n; ++j) colptr[j + 1] += colptr[j];
    int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * n);
    memcpy(pos, colptr, sizeof(int64_t) * n);
    for (int64_t i = 0; i < m; ++i)
        for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) {
            int64_t q = pos[col[e]]++;
            row[q] = (int32_t)i; valc[q] = val[e];
        }
    free(pos);
}

/* ------------------------------------------------------------------ */
/* Weights, auxil/get_mfun_dfun.m (matrix branch)                      */
/* ------------------------------------------------------------------ */
int ora_weights(int method, int64_t m, int64_t n,
                const int64_t* rowptr, const int32_t* col, const double* val,
                double* M, double* D)
{
    int64_t nnz = rowptr[m];
    switch (method) {
    case 0: /* landweber :50-54 */
        for (int64_t i = 0; i < m; ++i) M[i] = 1.0;
        for (int64_t j = 0; j < n; ++j) D[j] = 1.0;
        return 0;
    case 1: /* cimmino :56-88: M = 1/m*(1./normAi), Inf -> 0 */
    case 3: /* drop :140-198: M = 1./normAi */
        for (int64_t i = 0; i < m; ++i) {
            double s = 0;
            for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) s += val[e] * val[e];
            double v = method == 1 ? (1.0 / (double)m) * (1.0 / s) : 1.0 / s;
            M[i] = isinf(v) ? 0.0 : v;
        }
        if (method == 1) {
            for (int64_t j = 0; j < n; ++j) D[j] = 1.0;
        } else {   /* s_j = 1./nnz(A(:,j)), Inf -> 0 */
            int64_t* cnt = (int64_t*)calloc(n, sizeof(int64_t));
            for (int64_t e = 0; e < nnz; ++e) if (val[e] != 0) cnt[col[e]]++;
            for (int64_t j = 0; j < n; ++j) { double v = 1.0 / (double)cnt[j]; D[j] = isinf(v) ? 0.0 : v; }
            free(cnt);
        }
        return 0;
    case 2: { /* cav :90-138: s_j = nnz(col j); M = 1./sum(A.^2 * diag(s), 2) */
        int64_t* cnt = (int64_t*)calloc(n, sizeof(int64_t));
        for (int64_t e = 0; e < nnz; ++e) if (val[e] != 0) cnt[col[e]]++;
        for (int64_t i = 0; i < m; ++i) {
            double s = 0;
            for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) s += (val[e] * val[e]) * (double)cnt[col[e]];
            double v = 1.0 / s;
            M[i] = isinf(v) ? 0.0 : v;
        }
        for (int64_t j = 0; j < n; ++j) D[j] = 1.0;
        free(cnt);
        return 0;
    }
    case 4: { /* sart :200-241: M = 1./rowabssum, D = 1./colabssum */
        double* cs = (double*)calloc(n, sizeof(double));
        for (int64_t i = 0; i < m; ++i) {
            double s = 0;
            for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) { s += fabs(val[e]); cs[col[e]] += fabs(val[e]); }
            double v = 1.0 / s;
            M[i] = isinf(v) ? 0.0 : v;
        }
        for (int64_t j = 0; j < n; ++j) { double v = 1.0 / cs[j]; D[j] = isinf(v) ? 0.0 : v; }
        free(cs);
        return 0;
    }
    default:
        return -1;
    }
}

/* ------------------------------------------------------------------ */
/* calczeta, calc_relaxpar.m:179-244: 6 Newton steps from z0 = 1 on      */
/* g(z) = (2k-1) z^(k-1) - (z^(k-2) + ... + 1), Horner for g and g'.     */
/* ------------------------------------------------------------------ */
static void fung(double x, int k, double* f, double* df)
{
    /* C = [0, 2k-1, -ones(1,k-1)] */
    double ff = 0.0, d = 0.0;
    for (int i = 1; i <= k; ++i) {
        double Ci = i == 1 ? (double)(2 * k - 1) : -1.0;
        d = d * x + ff;
        ff = ff * x + Ci;
    }
    *f = ff; *df = d;
}
double ora_calczeta(int k)
{
    double x = 1.0, f, df;
    fung(x, k, &f, &df);
    double h = f / df;
    for (int it = 0; it < 6; ++it) {
        x = x - h;
        fung(x, k, &f, &df);
        h = f / df;
    }
    return x;
}

/* ------------------------------------------------------------------ */
/* Products                                                            */
/* ------------------------------------------------------------------ */
void ora_spmv(int64_t m, const int64_t* rowptr, const int32_t* col, const double* val,
              const double* x, double* y, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (int64_t i = 0; i < m; ++i) {
        double s = 0;
        for (int64_t e = rowptr[i]; e < rowptr[i + 1]; ++e) s += val[e] * x[col[e]];
        y[i] = s;
    }
}

static double vnorm(const double* v, int64_t m)
{
    double s = 0;
    for (int64_t i = 0; i < m; ++i) s += v[i] * v[i];
    return sqrt(s);
}

/* NCP statistic, check_stoprules.m:96-133 (direct DFT of each column) */
double ora_ncp_stat(const double* rk, int64_t m, const int* res_dims)
{
    int p = res_dims[1] > 0 ? res_dims[0] : (int)m;
    int ncols = res_dims[1] > 0 ? res_dims[1] : 1;
    int q = p / 2;
    double* pk = (double*)malloc(sizeof(double) * (q + 1));
    double* cw = (double*)malloc(sizeof(double) * (q + 1));
    double acc = 0;
    for (int c = 0; c < ncols; ++c) {
        const double* r = rk + (int64_t)c * p;
        for (int f = 0; f <= q; ++f) {
            double re = 0, im = 0;
            for (int t = 0; t < p; ++t) {
                int64_t ph = ((int64_t)f * t) % p;
                double ang = -2.0 * M_PI * (double)ph / (double)p;
                re += r[t] * cos(ang); im += r[t] * sin(ang);
            }
            pk[f] = re * re + im * im;
        }
        double tot = 0;
        for (int f = 1; f <= q; ++f) tot += pk[f];
        double cum = 0, ss = 0;
        for (int f = 1; f <= q; ++f) {
            cum += pk[f];
            double dv = cum / tot - (double)f / (double)q;
            ss += dv * dv;
        }
        acc += sqrt(ss);
    }
    free(pk); free(cw);
    return acc / ncols;
}

/* ------------------------------------------------------------------ */
/* SIRT loop, sirt/sirt.m:150-281; stop rules check_stoprules.m:49-164  */
/* ------------------------------------------------------------------ */
typedef struct { double* rkm1; double* dk; int ndk; } stop_state;

static int check_stop(const ora_sirt_opts* o, const double* rk, int64_t m, double relaxpar,
                      int k, int kmax, stop_state* ss, ora_sirt_info* info, double* nrm_out)
{
    double nrk = vnorm(rk, m);
    *nrm_out = nrk;
    switch (o->stoprule_type) {
    case 2: /* DP :56-72 */
        if (nrk <= o->taudelta || k >= kmax) {
            info->stoprule = k != kmax ? 2 : 0; info->finaliter = k; info->relaxpar = relaxpar;
            return 1;
        }
        return 0;
    case 3: { /* ME :74-94 */
        double nr1 = vnorm(ss->rkm1, m), dme = 0;
        for (int64_t i = 0; i < m; ++i) dme += ss->rkm1[i] * (ss->rkm1[i] + rk[i]);
        dme = dme / 2;
        if (dme / nr1 <= o->taudelta) {
            info->stoprule = 3; info->finaliter = k; info->relaxpar = relaxpar; return 1;
        } else if (k >= kmax) {
            info->stoprule = 0; info->finaliter = k; info->relaxpar = relaxpar; return 1;
        }
        memcpy(ss->rkm1, rk, sizeof(double) * m);
        return 0;
    }
    case 1: { /* NCP :96-151 */
        double ncc = ora_ncp_stat(rk, m, o->res_dims);
        double mx_old = -INFINITY, mx_new = ncc;
        for (int q = 0; q < ss->ndk; ++q) if (ss->dk[q] > mx_old) mx_old = ss->dk[q];
        for (int q = 1; q < ss->ndk; ++q) if (ss->dk[q] > mx_new) mx_new = ss->dk[q];
        if (mx_old < mx_new || k >= kmax) {
            info->stoprule = k != kmax ? 1 : 0; info->finaliter = k; info->relaxpar = relaxpar;
            return 1;
        }
        for (int q = 0; q + 1 < ss->ndk; ++q) ss->dk[q] = ss->dk[q + 1];
        ss->dk[ss->ndk - 1] = ncc;
        return 0;
    }
    default: /* NONE :153-161 */
        if (k >= kmax) { info->stoprule = 0; info->finaliter = k; info->relaxpar = relaxpar; return 1; }
        return 0;
    }
}

static void spmv_t(int64_t n, const int64_t* colptr, const int32_t* rowc, const double* valc,
                   const double* y, double* g)
{
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
    for (int64_t j = 0; j < n; ++j) {
        double s = 0;
        for (int64_t e = colptr[j]; e < colptr[j + 1]; ++e) s += valc[e] * y[rowc[e]];
        g[j] = s;
    }
}

int ora_sirt(int64_t m, int64_t n,
             const int64_t* rowptr, const int32_t* col, const double* val,
             const int64_t* colptr, const int32_t* rowc, const double* valc,
             const double* M, const double* D,
             const double* b, const double* x0,
             const int32_t* K, int nK, const ora_sirt_opts* o,
             double* X, int32_t* itersaved, double* resnorm, ora_sirt_info* info)
{
    int kmax = K[nK - 1];
#ifdef _OPENMP
    if (o->nthreads > 0) omp_set_num_threads(o->nthreads);
#endif
    double* xk = (double*)malloc(sizeof(double) * n);
    double* rk = (double*)malloc(sizeof(double) * m);
    double* Mr = (double*)malloc(sizeof(double) * m);
    double* g = (double*)malloc(sizeof(double) * n);
    stop_state ss = {0};
    if (o->stoprule_type == 3) {
        ss.rkm1 = (double*)malloc(sizeof(double) * m);
        for (int64_t i = 0; i < m; ++i) ss.rkm1[i] = NAN;
    }
    if (o->stoprule_type == 1) {
        ss.ndk = o->ncp_smooth > 0 ? o->ncp_smooth : 2;
        ss.dk = (double*)malloc(sizeof(double) * ss.ndk);
        for (int q = 0; q < ss.ndk; ++q) ss.dk[q] = INFINITY;
    }
    memcpy(xk, x0, sizeof(double) * n);
    /* rk = b - A*x0 */
    ora_spmv(m, rowptr, col, val, xk, rk, o->nthreads);
    for (int64_t i = 0; i < m; ++i) rk[i] = b[i] - rk[i];
    int k = 0, l = 0;
    double nrm;
    info->stoprule = 0; info->finaliter = 0; info->relaxpar = NAN;
    int stop = check_stop(o, rk, m, NAN, k, kmax, &ss, info, &nrm);
    if (resnorm) resnorm[0] = nrm;
    while (!stop) {
        k++;
        for (int64_t i = 0; i < m; ++i) Mr[i] = M ? M[i] * rk[i] : rk[i];
        spmv_t(n, colptr, rowc, valc, Mr, g);
        double lam;
        if (o->line_search) {
            double den = 0, num = 0;
            for (int64_t j = 0; j < n; ++j) den += D ? D[j] * (g[j] * g[j]) : g[j] * g[j];
            for (int64_t i = 0; i < m; ++i) num += rk[i] * Mr[i];
            lam = num / den;
        } else {
            lam = o->lambda[k - 1];
        }
        for (int64_t j = 0; j < n; ++j) {
            double v = xk[j] + lam * (D ? D[j] * g[j] : g[j]);
            if (o->lb_kind == 1) v = fmax(v, o->lb);
            else if (o->lb_kind == 2) v = fmax(v, o->lb_vec[j]);
            if (o->ub_kind == 1) v = fmin(v, o->ub);
            else if (o->ub_kind == 2) v = fmin(v, o->ub_vec[j]);
            xk[j] = v;
        }
        ora_spmv(m, rowptr, col, val, xk, rk, o->nthreads);
        for (int64_t i = 0; i < m; ++i) rk[i] = b[i] - rk[i];
        stop = check_stop(o, rk, m, lam, k, kmax, &ss, info, &nrm);
        if (resnorm) resnorm[k] = nrm;
        if ((l < nK && k == K[l]) || stop) {
            memcpy(X + (int64_t)l * n, xk, sizeof(double) * n);
            l++;
        }
    }
    /* itersaved = [K(K < finaliter); finaliter] */
    int ns = 0;
    for (int q = 0; q < nK; ++q) if (K[q] < info->finaliter) itersaved[ns++] = K[q];
    itersaved[ns++] = info->finaliter;
    info->nsaved = l;
    free(xk); free(rk); free(Mr); free(g);
    free(ss.rkm1); free(ss.dk);
    return 0;
}
