/*
 * air_oracle.h -- FP64 CPU restatement of the AIR Tools II SIRT path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (airtoolsii_amd / libairhip.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the AIR Tools II source tree).
 */
#ifndef AIR_ORACLE_H
#define AIR_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- MATLAB built-in semantics (documented behaviour, written from the maths) */
double ora_sind(double deg);
double ora_cosd(double deg);
double ora_atand(double x);
void   ora_linspace(double d1, double d2, int64_t n, double* out);
double ora_round(double x);                 /* MATLAB round: half away from zero */

/* MT19937 (init_genrand / genrand_res53).  MATLAB rng(0) == init_genrand(5489). */
typedef struct { uint32_t mt[624]; int mti; } ora_mt;
void   ora_mt_seed(ora_mt* s, uint32_t matlab_seed);
double ora_mt_rand(ora_mt* s);

/* ---- phantoms, testprobs/phantomgallery.m (column-major N x N output) */
void ora_phantom_shepplogan(int N, double* im);                 /* :155-214 */
void ora_phantom_smooth(int N, int p, double* im);              /* :216-232 */
void ora_phantom_threephases(int N, int p, int has_seed, uint32_t seed, double* im); /* :257-292 */
void ora_phantom_grains(int N, int numGrains, int has_seed, uint32_t seed, double* im); /* :352-389 */
/* CompHistTrain(Z, ones(3), 1), :458-509: H[id-1] for the 3x3 windows of the
 * nr x nc column-major 0/1 image Z (id = 1 + sum n_i 2^(i-1), window read column-major) */
void ora_clique_hist(const uint8_t* Z, int nr, int nc, double* H);
/* binary, :236-253 + sequential_simulation_full_clique :511-611 on the training
 * image ti (tr x tc column-major, getTI :677-2641); returns the number of rand draws */
int64_t ora_phantom_binary(int N, int has_seed, uint32_t seed, const uint8_t* ti, int tr, int tc, double* im);

/* ---- ray tracers -> CSR (0-based, columns ascending, duplicates summed) */
typedef struct ora_csr ora_csr;
/* testprobs/paralleltomo.m:112-301 */
ora_csr* ora_paralleltomo(int N, const double* theta, int nA, int p, double d);
/* testprobs/fanlineartomo.m:141-358  (R, dw, sd already multiplied by N) */
ora_csr* ora_fanlineartomo(int N, const double* theta, int nA, int p,
                           double R, double dw, double sd);
/* testprobs/fancurvedtomo.m:140-339  (R already multiplied by N, d in degrees) */
ora_csr* ora_fancurvedtomo(int N, const double* theta, int nA, int p, double R, double d);
/* testprobs/seismictomo.m:96-287 */
ora_csr* ora_seismictomo(int N, int s, int p);
/* testprobs/sphericaltomo.m:108-270 */
ora_csr* ora_sphericaltomo(int N, const double* angles, int nA, int nC);
/* testprobs/seismicwavetomo.m:114-288 */
ora_csr* ora_seismicwavetomo(int N, int s, int p, double omega);
/* art/art.m:249-285 and cart/cart.m:226-287 sweeps, and the counter-based
 * random stream shared with the engine (SplitMix64) */
uint64_t ora_splitmix64(uint64_t z);
double ora_rng_u01(uint64_t seed, uint64_t stream, uint64_t k, uint64_t t);
void ora_art_sweep(const int64_t* rp, const int32_t* ci, const double* v, const int32_t* rows, int64_t L,
                   const double* lamv, double lam, const double* b, const double* nrm, double* x, int64_t n,
                   int lbk, double lb, const double* lbv, int ubk, double ub, const double* ubv);
void ora_cart_sweep(const int64_t* cp, const int32_t* ri, const double* v, const int32_t* cols, int64_t L,
                    const double* nrm, double* x, double* r, int64_t n, unsigned char* flag, int32_t* nflag,
                    int lbk, double lb, const double* lbv, int ubk, double ub, const double* ubv,
                    double lam, double thr, int k, int kbegin, int nunflag, uint64_t seed);
int64_t ora_csr_m(const ora_csr*);
int64_t ora_csr_n(const ora_csr*);
int64_t ora_csr_nnz(const ora_csr*);
int     ora_csr_error(const ora_csr*);     /* nonzero: a column index fell outside 1..N^2 */
void    ora_csr_copy(const ora_csr*, int64_t* rowptr, int32_t* col, double* val);
void    ora_csr_free(ora_csr*);

/* ---- transpose CSR -> CSC (rows ascending within a column) */
void ora_csr_to_csc(int64_t m, int64_t n, const int64_t* rowptr, const int32_t* col,
                    const double* val, int64_t* colptr, int32_t* row, double* valc);

/* ---- M/D weights, auxil/get_mfun_dfun.m (matrix branch).  method:
 * 0 landweber, 1 cimmino, 2 cav, 3 drop, 4 sart */
int ora_weights(int method, int64_t m, int64_t n,
                const int64_t* rowptr, const int32_t* col, const double* val,
                double* M, double* D);

/* ---- calczeta, auxil/calc_relaxpar.m:179-244 */
double ora_calczeta(int k);
/* NCP statistic of a residual, check_stoprules.m:96-133: res_dims = {p, ncols} or {m, 0} */
double ora_ncp_stat(const double* rk, int64_t m, const int* res_dims);

/* ---- products */
void ora_spmv(int64_t m, const int64_t* rowptr, const int32_t* col, const double* val,
              const double* x, double* y, int nthreads);

/* ---- SIRT loop, sirt/sirt.m:150-281 + auxil/check_stoprules.m:49-164 */
typedef struct {
    int     stoprule_type;   /* 0 NONE, 1 NCP, 2 DP, 3 ME */
    double  taudelta;
    int     res_dims[2];     /* NCP: res_dims(1) (and res_dims(2) or 0 for 1-D) */
    int     ncp_smooth;
    int     lb_kind;         /* 0 none, 1 scalar, 2 vector */
    double  lb;
    const double* lb_vec;
    int     ub_kind;
    double  ub;
    const double* ub_vec;
    int     line_search;     /* relaxpar 'line' (casel 2) */
    const double* lambda;    /* kmax values (casel 1 or 3) */
    int     nthreads;
} ora_sirt_opts;

typedef struct {
    int     stoprule;        /* 0 max iter, 1 NCP, 2 DP, 3 ME */
    int     finaliter;
    double  relaxpar;        /* relaxpar of the stopping iteration */
    int     nsaved;
} ora_sirt_info;

int ora_sirt(int64_t m, int64_t n,
             const int64_t* rowptr, const int32_t* col, const double* val,
             const int64_t* colptr, const int32_t* rowc, const double* valc,
             const double* M, const double* D,
             const double* b, const double* x0,
             const int32_t* K, int nK, const ora_sirt_opts* o,
             double* X, int32_t* itersaved, double* resnorm, ora_sirt_info* info);

#ifdef __cplusplus
}
#endif
#endif
