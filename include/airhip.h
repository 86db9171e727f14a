/*
 * airhip.h -- C ABI of the MI355X SIRT engine (libairhip.so).
 *
 * Plain C types only: host pointers, sizes, POD structs.  One engine handle
 * owns one HIP device, one stream and one operator (a stored sparse matrix or
 * a matrix-free ray geometry).  Calls on one handle are not re-entrant;
 * separate handles may be used from separate threads.  Every function returns
 * AIR_OK (0) or a negative AIR_E_* code; air_last_error() gives the message
 * of the calling thread's last failure.
 *
 * Reference interfaces (AIR Tools II, MATLAB) each entry point replaces are
 * cited per declaration as file:line.
 */
#ifndef AIRHIP_H
#define AIRHIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AIR_OK             0
#define AIR_E_ARG         -1   /* invalid argument (MATLAB error(...) in check_inputs) */
#define AIR_E_NOMEM       -2   /* device or host allocation failed */
#define AIR_E_HIP         -3   /* HIP runtime error */
#define AIR_E_UNSUPPORTED -4   /* option not implemented for this operator */
#define AIR_E_STATE       -5   /* call out of order (e.g. no operator set) */
#define AIR_E_GEOMETRY    -6   /* tracer produced a column outside 1..N^2 */
#define AIR_E_CALLBACK    -7   /* a caller's operator function (air_set_afun) returned non-zero */

/* storage type of the matrix values; accumulation is FP32 (FP64 for AIR_F64) */
#define AIR_F32   0
#define AIR_BF16  1
#define AIR_F16   2
#define AIR_F64   3

/* SIRT methods, auxil/get_mfun_dfun.m:47-250 */
#define AIR_LANDWEBER 0
#define AIR_CIMMINO   1
#define AIR_CAV       2
#define AIR_DROP      3
#define AIR_SART      4
#define AIR_CUSTOM    5   /* struct('M',m-vector,'D',n-vector), get_mfun_dfun.m:252-307 */

/* relaxation strategies, auxil/calc_relaxpar.m:105-171 */
#define AIR_RELAX_DEFAULT 0   /* 1.9/rho */
#define AIR_RELAX_SCALAR  1
#define AIR_RELAX_LINE    2
#define AIR_RELAX_PSI1    3
#define AIR_RELAX_PSI1MOD 4
#define AIR_RELAX_PSI2    5
#define AIR_RELAX_PSI2MOD 6
#define AIR_RELAX_VECTOR  7   /* caller-supplied per-iteration lambda_k */
/* No counterpart in the reference (learned relaxation, for callers that want
 * the network's lambda_k and stop index instead of calc_relaxpar/check_stoprules): */
#define AIR_RELAX_LEARNED 8   /* lambda_k and a stop decision from the network (opts.mlp_params) */
#define AIR_RELAX_GREEDY  9   /* training trajectories: lambda_k minimising ||x_k - x_true|| along the step */

/* stopping rules, auxil/check_stoprules.m:53-161 (info.stoprule codes match) */
#define AIR_STOP_NONE 0
#define AIR_STOP_NCP  1
#define AIR_STOP_DP   2
#define AIR_STOP_ME   3
#define AIR_STOP_LEARNED 4    /* info.stoprule only: the network's stop logit fired (AIR_RELAX_LEARNED) */

/* bound kinds, auxil/check_inputs.m:108-132 */
#define AIR_BOUND_NONE   0
#define AIR_BOUND_SCALAR 1
#define AIR_BOUND_VECTOR 2

typedef struct air_engine air_engine;

/* options struct of sirt.m:36-83 as parsed by auxil/check_inputs.m:101-260 */
typedef struct {
    int           method;        /* AIR_LANDWEBER..AIR_CUSTOM */
    const double* M;             /* AIR_CUSTOM: m-vector diagonal or NULL (identity) */
    const double* D;             /* AIR_CUSTOM: n-vector diagonal or NULL (identity) */
    int           relax_kind;    /* AIR_RELAX_* */
    double        relaxpar;      /* AIR_RELAX_SCALAR value */
    const double* relax_vec;     /* AIR_RELAX_VECTOR: kmax values */
    double        rho;           /* spectral radius; NaN = compute (calc_relaxpar.m:91-103) */
    int           lb_kind;  double lb;  const double* lb_vec;
    int           ub_kind;  double ub;  const double* ub_vec;
    int           stoprule;      /* AIR_STOP_* */
    double        taudelta;      /* DP / ME */
    int           res_dims[2];   /* NCP: [p, nA] or [m, 0] */
    int           ncp_smooth;    /* NCP filter length, default 2 */
    int           verbose;       /* sirt.m:256-260 print cadence */
    /* learned relaxation (no counterpart in the reference) */
    const float*  mlp_params;    /* AIR_RELAX_LEARNED: AIR_MLP_PARAMS network parameters; with
                                    AIR_RELAX_GREEDY optional: the network's lambda drives the run
                                    and the greedy target is only recorded */
    double        noise_norm;    /* ||e|| if known (a feature of the network), else <= 0 */
    const double* x_true;        /* AIR_RELAX_GREEDY: the exact solution (n values); with
                                    AIR_RELAX_LEARNED optional, for the error record */
} air_sirt_options;

/* info struct of sirt.m:60-67,272-281 */
typedef struct {
    int    stoprule;     /* 0 max iter, 1 NCP, 2 DP, 3 ME */
    int    finaliter;
    double relaxpar;     /* relaxpar of the last iteration (scalar strategies) */
    double rho;
    int    nsaved;       /* columns written to X */
    double timetaken;    /* seconds, whole call */
    double setup_time;   /* seconds: weights + rho + first residual */
    double iter_time;    /* seconds: the K-loop on the device */
} air_sirt_info;

/* ---- lifetime -------------------------------------------------------- */
const char* air_last_error(void);
int  air_create(int device, air_engine** out);
void air_destroy(air_engine* e);
int  air_synchronize(air_engine* e);

/* ---- operators -------------------------------------------------------
 * A stored matrix: CSR with 0-based int32 column indices (ascending per row). */
int air_set_matrix_csr(air_engine* e, int64_t m, int64_t n, const int64_t* rowptr,
                       const int32_t* colind, const double* val, int dtype);
/* A MATLAB sparse matrix as handed over by mxGetJc/mxGetIr/mxGetPr (CSC,
 * 0-based row indices): the operand of auxil/afun_matrix.m:44-53. */
int air_set_matrix_csc(air_engine* e, int64_t m, int64_t n, const int64_t* colptr,
                       const int64_t* rowind, const double* val, int dtype);
/* A caller's operator: the function-handle afun of sirt/sirt.m:104-115 and
 * auxil/check_inputs.m:67-75.  fn(ctx, 0, x, y) must write y = A x (x: n
 * values, y: m values), fn(ctx, 1, w, z) z = A' w; host FP64 arrays owned by
 * the engine; a non-zero return fails the calling entry point with
 * AIR_E_CALLBACK.  Every product of the SIRT loop, of the spectral radius and
 * of air_apply calls fn on the calling thread; the rest of each iteration runs
 * on the device.  The weights follow get_mfun_dfun.m's function-handle
 * branches: Cimmino and DROP take the row norms from m products A'e_i
 * (:68-73, :150-155), CAV and DROP the column counts from n products A e_j
 * (:108-116, :179-184; CAV then sums (A'e_i)^2 / count, :117-121), SART the
 * products A 1 and A' 1 (:208, :229). */
typedef int (*air_afun_fn)(void* ctx, int transp, const double* in, double* out);
int air_set_afun(air_engine* e, int64_t m, int64_t n, air_afun_fn fn, void* ctx);
/* testprobs/paralleltomo.m:1-110 (defaults resolved by the caller: p, d).
 * matrix_free=0: trace on the GPU into CSR (+CSC); 1: keep the geometry
 * only and project on every product (paralleltomo.m:94-96): A x walks each
 * ray through the image tiles it crosses, A' y clips each pixel's candidate
 * rays; both use the same entries, equal to the tracer's to FP64 rounding. */
int air_build_paralleltomo(air_engine* e, int32_t N, const double* theta, int32_t nA,
                           int32_t p, double d, int matrix_free, int dtype);
/* testprobs/fanlineartomo.m:1-136; R, dw, sd are in units of N as in the
 * reference signature (the engine multiplies by N, fanlineartomo.m:75-90). */
int air_build_fanlineartomo(air_engine* e, int32_t N, const double* theta, int32_t nA,
                            int32_t p, double R, double dw, double sd,
                            int matrix_free, int dtype);
/* testprobs/fancurvedtomo.m:1-339; R in units of N (the engine multiplies by
 * N, fancurvedtomo.m:108), d the angular span in degrees (default resolved by
 * the caller: 2*atand(1/(2R-1)), fancurvedtomo.m:79-83). */
int air_build_fancurvedtomo(air_engine* e, int32_t N, const double* theta, int32_t nA,
                            int32_t p, double R, double d, int matrix_free, int dtype);
/* testprobs/seismictomo.m:1-287 (defaults resolved by the caller: s = N,
 * p = 2N): s sources on the right edge, p receivers on the left and top
 * edges, one row per source-receiver ray (row = source*p + receiver). */
int air_build_seismictomo(air_engine* e, int32_t N, int32_t s, int32_t p, int matrix_free, int dtype);
/* testprobs/sphericaltomo.m:1-270 (defaults resolved by the caller: numCircles
 * = round(sqrt(2) N)): row m*numCircles + n integrates circle n about centre m
 * (angles in degrees).  matrix_free = 1: no stored matrix; every product
 * re-forms the rows' circle samples (sphericaltomo.m:90-93,243-254) in FP64,
 * equal to the reference's function-handle products up to FP64 rounding; the
 * SIRT weights come from a temporary CSR (equal to the stored matrix's). */
int air_build_sphericaltomo(air_engine* e, int32_t N, const double* angles, int32_t nA,
                            int32_t numCircles, int matrix_free, int dtype);
/* testprobs/seismicwavetomo.m:1-288 (defaults resolved by the caller: s = N,
 * p = 2N, omega = 10): Fresnel-zone sensitivity of each source-receiver
 * pair, row = source*p + receiver.  matrix_free = 1: no stored matrix, the
 * kernels are re-evaluated on every product (seismicwavetomo.m:94-99). */
int air_build_seismicwavetomo(air_engine* e, int32_t N, int32_t s, int32_t p, double omega,
                              int matrix_free, int dtype);
/* ---- filtered back projection (auxil/fbp.m:31-80) ---------------------
 * x = A' * filter(b) * pi / (2 na), b the column-major sinogram (m/na rays x
 * na angles).  AIR_FBP_NONE gives the plain back projection x = A' b (:37). */
#define AIR_FBP_NONE        0
#define AIR_FBP_RAMLAK      1
#define AIR_FBP_SHEPPLOGAN  2
#define AIR_FBP_COSINE      3
#define AIR_FBP_HAMMING     4
#define AIR_FBP_HANN        5
int air_fbp(air_engine* e, const double* b, int32_t na, int filter, double* x);
/* A(v,'size') of auxil/check_inputs.m:74 (nnz = -1 for matrix-free) */
int air_size(air_engine* e, int64_t* m, int64_t* n, int64_t* nnz);
/* copy the stored CSR back to the host (values widened to double) */
int air_get_csr(air_engine* e, int64_t* rowptr, int32_t* colind, double* val);
/* afun convention, auxil/afun_matrix.m:47-50 / paralleltomo.m:282-287:
 * transp=0 -> out = A*in (m), transp=1 -> out = A'*in (n).  Host FP64. */
int air_apply(air_engine* e, int transp, const double* in, double* out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the product of
 * air_apply on device memory, calling no hipMalloc/hipFree.  in, out: device
 * FP32 vectors in the user's column order (n and m values for transp=0, m and n
 * for transp=1).  The call enqueues the product on the engine's stream and
 * returns without waiting (air_synchronize waits).  Matrix-free adjoints and
 * batch operators work in a fixed scratch array of AIR_APPLY_SCRATCH_BYTES per
 * GPU (8 bytes per column, resp. 4 per padded column); an operator that needs
 * more is refused with AIR_E_ARG.  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (two engines on two streams) race. */
#define AIR_APPLY_SCRATCH_BYTES (32u << 20)
int air_apply_device(air_engine* e, int transp, const float* in, float* out);

/* ---- SIRT building blocks ---------------------------------------------
 * auxil/get_mfun_dfun.m:47-250: diagonals of M (m) and D (n), FP64. */
int air_weights(air_engine* e, int method, double* M_out, double* D_out);
/* largest eigenvalue of D A' M A, calc_relaxpar.m:91-103 (Lanczos) */
int air_spectral_radius(air_engine* e, int method, const double* Mcustom,
                        const double* Dcustom, double* rho_out);
/* calczeta, calc_relaxpar.m:179-244, for k = 2..kmax-1 -> z[k-2] */
int air_calczeta(int32_t kmax, double* z_out);

/* ---- SIRT driver, sirt/sirt.m:1-281 ------------------------------------
 * K: nK strictly increasing iteration numbers (kmax = K[nK-1]).
 * X_out: n x nK column-major FP64 (first info.nsaved columns written).
 * itersaved: nK+1 ints (info.itersaved of sirt.m:279), returns count in
 * *n_itersaved.  resnorm (optional, kmax+1 doubles): ||r_k||, k=0..finaliter. */
int air_sirt(air_engine* e, const air_sirt_options* opts, const double* b, const double* x0,
             const int32_t* K, int32_t nK, double* X_out, int32_t* itersaved,
             int32_t* n_itersaved, double* resnorm, air_sirt_info* info);

/* Device-resident form of the same loop, for callers that time the
 * iterations alone.  prepare = sirt.m:143-169 (weights, rho, lambda, r0,
 * the k=0 stop check); iterate enqueues n iterations of sirt.m:199-261 and
 * waits for them; fetch copies x_k and the info back.  kmax bounds the
 * total number of iterations (it feeds the stop rule as K(end)). */
int air_sirt_prepare(air_engine* e, const air_sirt_options* opts, const double* b,
                     const double* x0, int32_t kmax);
int air_sirt_iterate(air_engine* e, int32_t niter, double* gpu_ms);
int air_sirt_fetch(air_engine* e, double* x_out, double* resnorm, air_sirt_info* info);

/* Per-iteration records of the last AIR_RELAX_LEARNED / AIR_RELAX_GREEDY run
 * (no counterpart in the reference; the training data of the network): for
 * the update of iteration k = 1..finaliter, row k-1 holds the 16 features of
 * the state x_{k-1} (feat), the greedy target lambda* rho (target), the raw
 * network outputs (out, learned runs) and ||x_{k-1} - x_true|| (err, when
 * x_true was given).  Any pointer may be NULL; rows beyond finaliter are not written. */
int air_sirt_records(air_engine* e, float* feat, float* target, float* out, double* err);

/* per-kernel HIP-event timing of the SIRT kernels: on = 0 turns it off,
 * on = s >= 1 brackets the kernels of every s-th iteration with events (each
 * event record costs the stream a few microseconds, so s > 1 keeps that cost
 * out of most iterations of a timed run).  Times are accumulated over the
 * bracketed iterations of iterate() calls while on (launches counts them);
 * reset by enabling again.  cols: the A'(M r) + update pass; rows: the A x
 * pass; reduce: the residual epilogue kernel of the tiled layout (0 for the
 * compressed-row layout). */
int air_set_kernel_timing(air_engine* e, int on);
int air_get_kernel_timing(air_engine* e, double* cols_ms, double* rows_ms, double* reduce_ms, int64_t* launches);

/* Graph mode (no counterpart in the reference, which has no launch
 * mechanism to speak of): iterations of the device-resident loop are captured
 * once per prepared run into a HIP graph of iters_per_graph iterations and
 * replayed, which removes the per-kernel launch cost for small problems and
 * batches.  It applies to the tiled layout with a host lambda schedule
 * (default, scalar, psi, vector) and the stop rules none/DP/ME; other runs
 * launch directly.  0 (the default) turns it off.  Takes effect at the next
 * air_sirt / air_sirt_prepare. */
int air_set_graph(air_engine* e, int32_t iters_per_graph);
/* number of graph launches on this engine so far */
int air_graph_launches(air_engine* e, int64_t* n);

/* Storage of the operator used by the SIRT loop. */
typedef struct {
    int     tiled;        /* 1: LDS-tiled layout, 0: CSR + CSC */
    int     value_bytes;  /* 4 (FP32), 2 (BF16/FP16), 8 (FP64) */
    int64_t m, n, nnz;
    int64_t n_internal;   /* padded column count of the tiled layout */
    int64_t pieces;       /* forward row segments */
    int64_t runs;         /* backward staging runs */
    int64_t tiles_fwd, tiles_bwd;
    int64_t staged_rows;  /* sum over backward tiles of their distinct rows */
    int64_t lds_bytes;    /* dynamic LDS of the backward kernel */
    int     fwd_stream;             /* forward kernel: 0 per-piece lane groups, 1 flat stream of whole waves */
    int     fwd_lanes, fwd_loads;   /* autotuned shape: lanes per piece, vector loads per round
                                       (stream form: entries per lane, 0 for the lane form, and 1 if the
                                       next round is prefetched) */
    int     bwd_lanes, bwd_loads;   /* lanes per column, vector loads per round (stream form as forward) */
    int     fwd_pieces_per_block;   /* pieces per block (stream form: target entries per wave) */
    int64_t fwd_entries, bwd_entries;  /* stored entries of the two streams: every row piece and every
                                          column padded to a multiple of 8 (the lane form) */
    float   tune_fwd_us[64];        /* autotuner timings: per-piece shape c at 256/512/1024 pieces per
                                       block -> [3c+i], c < 5; stream form v (v < 4: 8 or 16 entries per
                                       lane, without/with prefetch; v = 4, 5: the lane form without/with
                                       prefetch; v = 6, 7: the step-coded lane form without/with prefetch)
                                       at 2048/4096/8192/16384 entries per wave -> [15+4v+i]; the
                                       step-coded form v = 6, 7 with 16 waves per block -> [47+4(v-6)+i];
                                       0: not run */
    float   tune_bwd_us[16];        /* backward shapes: per-column 0-3, stream form 4-10 (8 entries per
                                       lane with prefetch, 16 without, 16 with, lane form without, with,
                                       step-coded lane form without, with) */
    int     fwd_step;               /* 1: the forward runs the step-coded lane form (one 32-bit word of
                                       unit moves per 8 entries instead of 16-bit column codes) */
    int     fwd_step_ok;            /* 1: every row piece is a unit-step staircase, so the step-coded
                                       form is available to the autotuner */
    int     bwd_step;               /* 1: the backward runs the step-coded lane form (rows staged with a
                                       fixed stride per run, one 32-bit word of moves per 8 entries) */
    int     bwd_step_ok;            /* 1: every column's moves could be encoded and the strided
                                       staging fits LDS */
    int64_t bwd_step_entries;       /* entries of the step-coded backward stream (bridging zeros and
                                       lane padding included) */
    int     fwd_waves_per_block;    /* stream forward: waves sharing one staged tile of x (4, or 16 for
                                       the step-coded form) */
    int     bwd_step_base_bits;     /* step-coded backward: bits of a word's base slot, 14 (tiles stage
                                       fewer than 16384 slots) or 15 (fewer than 32768) */
    int64_t bwd_step_lds_bytes;     /* dynamic LDS of the step-coded backward's strided staging */
} air_layout;
int air_layout_info(air_engine* e, air_layout* out);

/* ---- ART and CART: art/art.m:1-330, cart/cart.m:1-323 ----------------------
 * Row action (Kaczmarz) and column action sweeps over a stored matrix, in the
 * reference's step order.  Shared options (check_inputs.m:101-250): */
#define AIR_ART_KACZMARZ     0
#define AIR_ART_SYMKACZMARZ  1   /* top -> bottom -> top, one sweep counts as 2 iterations */
#define AIR_ART_RANDKACZMARZ 2
#define AIR_ART_CUSTOM       3   /* row order given by the caller (1-based), art.m:198-203 */
typedef struct {
    int           relax_kind;    /* AIR_RELAX_DEFAULT (ART 1, CART 0.25), AIR_RELAX_SCALAR, or
                                    AIR_RELAX_VECTOR (ART only: the function handle's values,
                                    relax_vec[(k-1)*m + i-1] for sweep k and row i, art.m:273-275) */
    double        relaxpar;
    const double* relax_vec;
    int           lb_kind;  double lb;  const double* lb_vec;
    int           ub_kind;  double ub;  const double* ub_vec;
    int           stoprule;      /* AIR_STOP_NONE, AIR_STOP_NCP, AIR_STOP_DP or AIR_STOP_ME */
    double        taudelta;
    double        damp;          /* >= 0, default 0 */
    double        THR;           /* CART flagging threshold, default 1e-4 */
    int           Kbegin;        /* CART: flagging from this sweep on, default 10 */
    int           Nunflag;       /* CART unflagging parameter, <= 0: round(kmax/4) */
    uint64_t      seed;          /* randkaczmarz orders and CART unflagging draws (SplitMix64 stream;
                                    MATLAB's rand is not reproduced) */
    int           res_dims[2];   /* NCP: [p, nA] or [m, 0], as in air_sirt_options */
    int           ncp_smooth;    /* NCP filter length, default 2 */
} air_row_options;
/* method: AIR_ART_*; order/norder: AIR_ART_CUSTOM row order.  K: nK strictly
 * increasing iteration numbers (even for symkaczmarz).  X_out: n x nK FP64;
 * itersaved: nK+1 ints; resnorm (optional): ||r_k|| for k = 0..sweeps. */
int air_art(air_engine* e, int method, const int32_t* order, int64_t norder, const air_row_options* opts,
            const double* b, const double* x0, const int32_t* K, int32_t nK, double* X_out,
            int32_t* itersaved, int32_t* n_itersaved, double* resnorm, air_sirt_info* info);
/* columnaction (order NULL) or a custom column order (1-based), cart.m:177-190 */
int air_cart(air_engine* e, const int32_t* order, int64_t norder, const air_row_options* opts,
             const double* b, const double* x0, const int32_t* K, int32_t nK, double* X_out,
             int32_t* itersaved, int32_t* n_itersaved, double* resnorm, air_sirt_info* info);

/* ---- batches of independent problems -----------------------------------
 * No counterpart in the reference (which solves one system per call): many
 * paralleltomo problems of different sizes solved together.  The engine
 * builds each problem (paralleltomo.m:112-301 with d = p-1) and concatenates
 * them into the block-diagonal operator diag(A_1, ..., A_P), held only in the
 * tiled layout, so every SIRT pass is one launch over all problems.  Per
 * problem it computes the weights of `method` (get_mfun_dfun.m) and rho
 * (calc_relaxpar.m:91-103); air_batch_info returns M with each problem's
 * default lambda = 1.9/rho_p folded into its rows, so that air_sirt with
 * AIR_CUSTOM(M, D), relaxpar 1 and rho 1 performs the independent default-
 * lambda runs.  theta holds the nA[0] + ... + nA[P-1] angles back to back. */
int air_build_paralleltomo_batch(air_engine* e, int32_t nprob, const int32_t* N, const int32_t* nA,
                                 const double* theta, const int32_t* p, int method, int dtype);
/* row_off, col_off: nprob+1 offsets; rho: nprob; M: m values (lambda-scaled); D: n values */
int air_batch_info(air_engine* e, int64_t* row_off, int64_t* col_off, double* rho, double* M, double* D);

/* ---- many right-hand sides on one operator --------------------------------
 * No counterpart in the reference (one sirt call per right-hand side): SIRT
 * (sirt.m:199-261 with the stop rule 'none') for nrhs <= 64 right-hand sides
 * sharing a stored A, each pass reading A once for all of them (SpMM).
 * B: m x nrhs column-major; X0 (optional) and X_out: n x nrhs column-major;
 * resnorm (optional): ||b_j - A x_j|| after K iterations; gpu_ms: device time
 * of the K iterations.  Options: method or custom M/D, default or scalar
 * relaxation, scalar bounds. */
int air_sirt_multi(air_engine* e, const air_sirt_options* opts, int32_t nrhs, const double* B, const double* X0,
                   int32_t K, double* X_out, double* resnorm, double* gpu_ms);

/* ---- learned relaxation network ---------------------------------------
 * No counterpart in the reference: the learned lambda_k / stop rule that
 * replaces the fixed relaxation of calc_relaxpar.m:105-171 and the stop
 * rules of check_stoprules.m when AIR_RELAX_LEARNED is chosen.  A 16-64-64-2
 * ReLU perceptron; parameters are AIR_MLP_PARAMS floats laid out as
 * W1[64][16], b1[64], W2[64][64], b2[64], W3[2][64], b3[2] (row = output).
 * Outputs (o0, o1): lambda_k * rho = AIR_MLP_LAMBDA_MAX sigmoid(o0); o1 is the stop logit. */
#define AIR_MLP_IN     16
#define AIR_MLP_HID    64
#define AIR_MLP_OUT    2
#define AIR_MLP_PARAMS 5378
#define AIR_MLP_LAMBDA_MAX 12     /* lambda_k rho ranges over (0, 12) */
/* raw outputs of n samples: X is n x 16 row-major, out n x 2 */
int air_mlp_forward(air_engine* e, const float* params, const float* X, int64_t n, float* out);
/* mean loss over n samples and its gradient (BF16 MFMA, FP32 accumulation).
 * Targets T (n x 2): t0 = lambda* rho in (0, 12) (t0 < 0 drops the term), t1 in {0, 1}.
 * Loss = mean[(12 sigmoid(o0) - t0)^2 + stop_weight * BCE(o1, t1)]. */
int air_mlp_grad(air_engine* e, const float* params, const float* X, const float* T, int64_t n,
                 double stop_weight, float* grad, double* loss);
/* Adam training in place: `epochs` passes over the n samples in seeded random
 * minibatches of `batch`; epoch_loss (optional, epochs values) gets the mean loss. */
int air_mlp_train(air_engine* e, float* params, const float* X, const float* T, int64_t n, int epochs,
                  int batch, double lr, double stop_weight, uint64_t seed, double* epoch_loss);

#ifdef __cplusplus
}
#endif
#endif
