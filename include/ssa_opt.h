/*
 * ssa_opt.h -- drop-in C API of SSA-Opt's decomposition hot path, backed by
 * an MI355X (gfx950) device path.
 *
 * Every declaration below replaces the function of the same name in the
 * reference header MKL/ssa_opt.h (declarations at MKL/ssa_opt.h:318-430,
 * definitions cited per entry) and keeps its C ABI: cdecl, LP64, int = 32
 * bits, double data, caller-allocated outputs, 0 / -1 returns, getters that
 * return a count, double-returning calls that return 0.0 on error.  The
 * reference's Python ctypes wrapper (MKL/py/ssa_wrapper.py:125-291) binds
 * these symbols unchanged; it treats the context as an opaque 1024-byte
 * block, so sizeof(SSA_Opt) and sizeof(MSSA_Opt) stay below 1024.
 *
 * The context keeps the reference's public result fields (dimensions,
 * flags, U, V, sigma, eigenvalues, n_components, total_variance) valid on
 * the host after each decomposition; the device state hangs off `dev`.
 *
 * Differences from the reference, all deliberate:
 *   - decompose_randomized performs q = 2 subspace (power) iterations by
 *     default (the reference has none); ssa_amd_set_options changes q.
 *   - the Gaussian test matrix comes from a counter-based Philox4x32-10
 *     stream keyed by (seed, decomposition number) instead of MKL's MT2203;
 *     it can be injected with ssa_amd_inject_omega for parity checks.
 *   - fft_len is the device correlation length (a power of two >= the
 *     segment length), not next_pow2(N + K - 1).
 *   - ssa_opt_decompose / _decompose_block / _extend run the same device
 *     subspace iteration with q derived from max_iter (documented below).
 */
#ifndef SSA_OPT_H
#define SSA_OPT_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double ssa_real;

/* MKL/ssa_opt.h:145-225 */
typedef struct {
    int N, L, K;
    int fft_len, r2c_len, batch_size;
    /* host views kept for C callers that check them (MKL/test/ssa_opt_test.c:179-181) */
    ssa_real *fft_x;            /* first segment spectrum, r2c_len complex (mirror) */
    ssa_real *ws_real, *ws_complex, *ws_real2, *ws_proj, *ws_batch_real, *ws_batch_complex;
    /* results (host mirrors, refreshed by every decomposition) */
    ssa_real *U;                /* L x n_components, column-major */
    ssa_real *V;                /* K x n_components, column-major */
    ssa_real *sigma;
    ssa_real *eigenvalues;
    int n_components;
    ssa_real *inv_diag_count;   /* 1/min(t+1, L, K, N-t) */
    ssa_real *U_fft, *V_fft;    /* unused (spectra stay on the device) */
    bool fft_cached;
    int prepared_kp, prepared_lwork;
    ssa_real *decomp_Omega, *decomp_Y;   /* unused host views (device resident) */
    bool prepared, initialized, decomposed;
    ssa_real total_variance;    /* sum of the k computed eigenvalues (MKL/ssa_opt.h:1741-1746) */
    /* device path */
    void *dev;
    int q;                      /* power iterations of decompose_randomized */
    int precision;              /* 0: plain FP64, 1: mixed (FP32 bulk, FP64 small solves) */
    unsigned long long seed;    /* Philox key of the test matrix */
    unsigned long long calls;   /* decompositions performed (Philox counter word) */
} SSA_Opt;

/* MKL/ssa_opt.h:227-249 */
typedef struct {
    int n;
    ssa_real *singular_values;
    ssa_real *log_sv;
    ssa_real *gaps;
    ssa_real *cumulative_var;
    ssa_real *second_diff;
    int suggested_signal;
    ssa_real gap_threshold;
} SSA_ComponentStats;

/* MKL/ssa_opt.h:251-268 */
typedef struct {
    ssa_real *R;
    int L;
    ssa_real verticality;
    bool valid;
} SSA_LRF;

/* MKL/ssa_opt.h:271-315 */
typedef struct {
    int M, N, L, K;
    int fft_len, r2c_len;
    ssa_real *U;                /* (M L) x n_components */
    ssa_real *V;                /* K x n_components */
    ssa_real *sigma;
    ssa_real *eigenvalues;
    int n_components;
    ssa_real *inv_diag_count;
    bool initialized, decomposed;
    ssa_real total_variance;
    void *dev;
    int q, precision;
    unsigned long long seed, calls;
} MSSA_Opt;

/* ---- lifecycle (MKL/ssa_opt.h:1026-1558) ---- */
int ssa_opt_init(SSA_Opt *ssa, const ssa_real *x, int N, int L);                 /* :1026-1284 */
void ssa_opt_free(SSA_Opt *ssa);                                                  /* :1286-1350 */
int ssa_opt_prepare(SSA_Opt *ssa, int max_k, int oversampling);                   /* :1414-1543 */
void ssa_opt_free_prepared(SSA_Opt *ssa);                                         /* :1352-1379 */
int ssa_opt_update_signal(SSA_Opt *ssa, const ssa_real *new_x);                   /* :1545-1558 */

/* ---- decomposition ---- */
int ssa_opt_decompose_randomized(SSA_Opt *ssa, int k, int oversampling);         /* :1685-1763 */
/* sequential power iteration in the reference (:1560-1678); here the device
 * subspace iteration with b = min(k + 8, L, K) and q = clamp(max_iter / 4, 2, 30) */
int ssa_opt_decompose(SSA_Opt *ssa, int k, int max_iter);
/* block power iteration in the reference (:1765-1933); here the device
 * subspace iteration with b = min(max(k + 8, block_size), L, K), q as above */
int ssa_opt_decompose_block(SSA_Opt *ssa, int k, int block_size, int max_iter);
int ssa_opt_extend(SSA_Opt *ssa, int additional_k, int max_iter);               /* :1935-2056 */

/* Hankel products y = H v (v: K, y: L) and z = H^T u (u: L, z: K); static
 * in the reference (:871-903) and called by its tests (MKL/test/ssa_opt_test.c:1484, :1533) */
void ssa_opt_hankel_matvec(SSA_Opt *ssa, const ssa_real *v, ssa_real *y);
void ssa_opt_hankel_matvec_T(SSA_Opt *ssa, const ssa_real *u, ssa_real *y);

/* ---- reconstruction and analysis ---- */
int ssa_opt_reconstruct(const SSA_Opt *ssa, const int *group, int n_group, ssa_real *output);   /* :2325-2400 */
int ssa_opt_cache_ffts(SSA_Opt *ssa);                                             /* :2110-2323 */
void ssa_opt_free_cached_ffts(SSA_Opt *ssa);                                      /* :2058-2078 */
int ssa_opt_wcorr_matrix(const SSA_Opt *ssa, ssa_real *W);                        /* :2407-2477 */
int ssa_opt_wcorr_matrix_fast(const SSA_Opt *ssa, ssa_real *W);                   /* :2492-2548 */
ssa_real ssa_opt_wcorr_pair(const SSA_Opt *ssa, int i, int j);                    /* :2551-2625 */
int ssa_opt_component_stats(const SSA_Opt *ssa, SSA_ComponentStats *stats);      /* :2627-2676 */
void ssa_opt_component_stats_free(SSA_ComponentStats *stats);                     /* :2678-2688 */
int ssa_opt_find_periodic_pairs(const SSA_Opt *ssa, int *pairs, int max_pairs,
                                ssa_real sv_tol, ssa_real wcorr_thresh);          /* :2690-2756 */
int ssa_opt_get_trend(const SSA_Opt *ssa, ssa_real *output);                      /* :3086-3090 */
int ssa_opt_get_noise(const SSA_Opt *ssa, int noise_start, ssa_real *output);     /* :3091-3110 */
ssa_real ssa_opt_variance_explained(const SSA_Opt *ssa, int start, int end);      /* :3111-3121 */
int ssa_opt_get_singular_values(const SSA_Opt *ssa, ssa_real *output, int max_n); /* :3123-3130 */
int ssa_opt_get_eigenvalues(const SSA_Opt *ssa, ssa_real *output, int max_n);     /* :3132-3139 */
ssa_real ssa_opt_get_total_variance(const SSA_Opt *ssa);                          /* :3141-3145 */

/* ---- forecasting (MKL/ssa_opt.h:2758-2885) ---- */
int ssa_opt_compute_lrf(const SSA_Opt *ssa, const int *group, int n_group, SSA_LRF *lrf);
void ssa_opt_lrf_free(SSA_LRF *lrf);
int ssa_opt_forecast_with_lrf(const SSA_LRF *lrf, const ssa_real *base_signal, int base_len, int n_forecast,
                              ssa_real *output);
int ssa_opt_forecast(const SSA_Opt *ssa, const int *group, int n_group, int n_forecast, ssa_real *output);
int ssa_opt_forecast_full(const SSA_Opt *ssa, const int *group, int n_group, int n_forecast, ssa_real *output);

/* ---- V-forecast (MKL/ssa_opt.h:2887-3084, declared at :352-355) ----
 * One projection per eigenvector and step, x = sum_g pi_g (U_g[0:L-1] . z) / (1 - nu^2),
 * on the device.  -1 for an index outside the computed components, nu^2 >= 1 - 1e-10,
 * n_forecast outside 1..65536, or (fast) base_len < L - 1. */
int ssa_opt_vforecast(const SSA_Opt *ssa, const int *group, int n_group, int n_forecast, ssa_real *output);
int ssa_opt_vforecast_full(const SSA_Opt *ssa, const int *group, int n_group, int n_forecast, ssa_real *output);
int ssa_opt_vforecast_fast(const SSA_Opt *ssa, const int *group, int n_group, const ssa_real *base_signal,
                           int base_len, int n_forecast, ssa_real *output);

/* ---- Cadzow iterations and gap filling (MKL/ssa_opt.h:3147-3878, declared at :372-430) ----
 * Host loops over this API: one device context is re-targeted with update_signal
 * instead of a new context per iteration, and its test-matrix counter is reset
 * so that every iteration draws the matrix a fresh context would (the reference
 * re-initialises, restarting its stream). */
typedef struct {
    int iterations;       /* iterations performed */
    ssa_real final_diff;  /* final relative difference */
    ssa_real converged;   /* 1.0 if converged, 0.0 if max_iter was reached */
} SSA_CadzowResult;

typedef struct {
    int iterations;
    ssa_real final_diff;  /* final relative change of the gap values */
    int converged;
    int n_gaps;           /* NaN positions filled */
} SSA_GapFillResult;

/* :3165-3292: returns the iterations performed, -1 on bad arguments or a failed step */
int ssa_opt_cadzow(const ssa_real *x, int N, int L, int rank, int max_iter, ssa_real tol, ssa_real *output,
                   SSA_CadzowResult *result);
/* :3294-3315: output = alpha * cadzow + (1 - alpha) * x, 0 < alpha <= 1 */
int ssa_opt_cadzow_weighted(const ssa_real *x, int N, int L, int rank, int max_iter, ssa_real tol, ssa_real alpha,
                            ssa_real *output, SSA_CadzowResult *result);
/* :3319-3440: iterates on the context's own signal (update_signal between steps) */
int ssa_opt_cadzow_inplace(SSA_Opt *ssa, int rank, int max_iter, ssa_real tol, SSA_CadzowResult *result);
/* :3498-3683: NaN positions of x filled in place by iterated rank-`rank` reconstruction */
int ssa_opt_gapfill(ssa_real *x, int N, int L, int rank, int max_iter, ssa_real tol, SSA_GapFillResult *result);
/* :3685-3877: each gap filled from a forecast of the data before it and a backcast of the data after it */
int ssa_opt_gapfill_simple(ssa_real *x, int N, int L, int rank, SSA_GapFillResult *result);

/* ---- ESPRIT parameter estimation (MKL/ssa_opt.h:4289-4558, declared at :387-403) ----
 * Host FP64 on the mirrored U: least squares U[0:L-1] Z = U[1:L] by Householder QR,
 * eigenvalues of Z by Francis double-shift QR.  The poles come in this solver's
 * order (the reference's follows dgeev). */
typedef struct {
    ssa_real *periods;      /* samples per cycle, infinite for |frequency| <= 1e-12 */
    ssa_real *frequencies;  /* cycles per sample, arg(pole) / (2 pi) */
    ssa_real *moduli;       /* |pole| */
    ssa_real *rates;        /* log |pole|, -30 below 1e-12 */
    int n_components;
} SSA_ParEstimate;

int ssa_opt_parestimate(const SSA_Opt *ssa, const int *group, int n_group, SSA_ParEstimate *result);
void ssa_opt_parestimate_free(SSA_ParEstimate *result);
/* No counterpart in the reference: the ESPRIT core of ssa_opt_parestimate on k
 * columns U (L x k, column-major): poles wr + i wi.  For callers that hold
 * eigenvectors themselves, and for host-only checks. */
int ssa_amd_esprit_poles(const ssa_real *U, int L, int k, ssa_real *wr, ssa_real *wi);

/* ---- MKL configuration hooks (MKL/py/ssa_wrapper.c:47-124) ----
 * Exported for the reference's benches; the device path has no MKL, so they do
 * nothing (init returns 0, get_threads 1, get_cpu_info reports the online cores). */
int ssa_mkl_init(int verbose);
int ssa_mkl_init_14900kf(int verbose);
int ssa_mkl_init_generic(int verbose);
void ssa_mkl_set_threads(int n);
int ssa_mkl_get_threads(void);
void ssa_mkl_get_cpu_info(int *p_cores, int *e_cores, int *is_hybrid, int *has_avx512);
void ssa_mkl_disable_affinity(void);

/* ---- MSSA (MKL/ssa_opt.h:3881-4283) ---- */
int mssa_opt_init(MSSA_Opt *mssa, const ssa_real *X, int M, int N, int L);
int mssa_opt_decompose(MSSA_Opt *mssa, int k, int oversampling);
int mssa_opt_reconstruct(const MSSA_Opt *mssa, int series_idx, const int *group, int n_group, ssa_real *output);
int mssa_opt_reconstruct_all(const MSSA_Opt *mssa, const int *group, int n_group, ssa_real *output);
int mssa_opt_series_contributions(const MSSA_Opt *mssa, ssa_real *contributions);
ssa_real mssa_opt_variance_explained(const MSSA_Opt *mssa, int start, int end);
void mssa_opt_free(MSSA_Opt *mssa);

/* ---- extensions of the device path (no counterpart in the reference) ---- */
/* q power iterations, precision (0 plain FP64, 1 mixed), Philox seed; call
 * after init.  A precision change rebuilds the device state from the
 * host copy of the signal. */
int ssa_amd_set_options(SSA_Opt *ssa, int q, int precision, unsigned long long seed);
int mssa_amd_set_options(MSSA_Opt *mssa, int q, int precision, unsigned long long seed);
/* use this K x b column-major Gaussian matrix (b = min(k + p, L, K)) for the
 * next decompose_randomized instead of drawing one */
int ssa_amd_inject_omega(SSA_Opt *ssa, const ssa_real *Omega, int K, int b);
/* reconstruct n_groups groups in one launch; groups concatenated in glist,
 * sizes in gcount; output n_groups * N values */
int ssa_amd_reconstruct_many(const SSA_Opt *ssa, const int *glist, const int *gcount, int n_groups,
                             ssa_real *output);
/* the device correlation plan of the context: out[0..5] = n, CL, CK, AL, SK, Jacobi sweeps */
int ssa_amd_plan_info(const SSA_Opt *ssa, int *out6);
/* streaming tick: slide the window by m samples (1 <= m <= N), x <- x[m:N]
 * followed by x_new[0:m], and refresh the device spectra; the same state as
 * ssa_opt_update_signal (MKL/ssa_opt.h:1545-1558) with the shifted window,
 * without the caller holding the window.  Clears decomposed. */
int ssa_amd_push_samples(SSA_Opt *ssa, const ssa_real *x_new, int m);
/* host mirrors of U and V: on (the default, as the reference's fields) every
 * decomposition copies them to the host; off, a decomposition leaves them
 * stale and they are copied when host code needs them (compute_lrf,
 * parestimate), by ssa_amd_sync_mirrors, or when mirroring is turned on
 * again.  sigma, eigenvalues and total_variance are always current. */
int ssa_amd_set_mirror(SSA_Opt *ssa, int on);
int ssa_amd_sync_mirrors(const SSA_Opt *ssa);

/* autotuner: time the block matvec at every FFT length 2^6..2^13 on this
 * device for the precision (0 plain FP64, 1 mixed) and plan every context
 * and batch created afterwards with the measured transform costs instead of
 * the n log n model; ns8 (may be NULL) receives the 8 per-transform costs */
int ssa_amd_autotune(int precision, double *ns8);

/* batched path: many uneven series in device engines (parallel graph
 * branches, SSA_AMD_BRANCHES, default 2) */
typedef struct ssa_amd_batch ssa_amd_batch;
ssa_amd_batch *ssa_amd_batch_create(int count, const int *N, const int *L, const int *k, int p, int q,
                                    int precision, unsigned long long seed);
/* as above, with the group {0..g[i]-1} reconstructed by each step (g NULL: all
 * k) and, when h > 0, an h-step LRF forecast of that group appended */
ssa_amd_batch *ssa_amd_batch_create_ex(int count, const int *N, const int *L, const int *k, const int *g, int h,
                                       int p, int q, int precision, unsigned long long seed);
/* No counterpart in the reference: as ssa_amd_batch_create_ex with the number
 * of parallel graph branches given (1..4; 0: SSA_AMD_BRANCHES or 2).  A
 * serving loop that keeps several batches in flight takes 1 branch each, so
 * its streams do not outnumber the device's hardware queues; a series'
 * result does not depend on the branch count.  NULL for branches outside
 * 0..4. */
ssa_amd_batch *ssa_amd_batch_create_g(int count, const int *N, const int *L, const int *k, const int *g, int h,
                                      int p, int q, int precision, unsigned long long seed, int branches);
/* re-target an idle batch to a new set of series (same p, q, h, precision,
 * seed); device buffers are reused when large enough, so a serving loop that
 * rebinds a few batches allocates nothing per wave */
int ssa_amd_batch_rebind(ssa_amd_batch *b, int count, const int *N, const int *L, const int *k, const int *g);
/* 1 when the batch's queued work has finished, 0 if still running, -1 on error */
int ssa_amd_batch_query(ssa_amd_batch *b);
/* device status of series i after a step (0 ok, 2 forecast verticality >= 1) */
int ssa_amd_batch_status(ssa_amd_batch *b, int i);
/* signals concatenated as float32 (offsets[i] = first sample of series i) */
int ssa_amd_batch_set_signals_f32(ssa_amd_batch *b, const float *cat, const long long *offsets);
/* Philox key of series i (default: seed ^ 0x9E3779B97F4A7C15 * (i + 1)); a
 * request keyed by its own id draws the same test matrix in any batch */
int ssa_amd_batch_set_key(ssa_amd_batch *b, int i, unsigned long long key);
/* No counterpart in the reference: block-matvec launches of at least n items
 * (1: every launch) take the wide transform geometry, which rounds
 * differently from the narrow one (FP32 at n = 8192, FP64 at 1024..4096).  A
 * server pins n so that a request's result does not depend on the size of
 * the wave it lands in; the default chooses by launch size.  Returns -1 for
 * n < 1. */
int ssa_amd_batch_set_wide_min(ssa_amd_batch *b, int n);
/* one SSA of every series: spectra, decomposition, reconstruction of all
 * components; use_graph captures the launch sequence in a HIP graph */
int ssa_amd_batch_run(ssa_amd_batch *b, int use_graph);
int ssa_amd_batch_sync(ssa_amd_batch *b);
int ssa_amd_batch_sigma(ssa_amd_batch *b, int i, ssa_real *sigma);       /* returns k_i */
int ssa_amd_batch_output(ssa_amd_batch *b, int i, ssa_real *out);         /* N_i (+ h) values */
/* every series' output at out + offsets[i] (N_i (+ h) values each) and, when
 * status is not NULL, its device status word; one synchronisation per branch */
int ssa_amd_batch_outputs(ssa_amd_batch *b, ssa_real *out, const long long *offsets, int *status);
/* device bytes the batch's arenas use (the sum of ssa_amd_series_bytes over its series) */
long long ssa_amd_batch_bytes(const ssa_amd_batch *b);
/* No counterpart in the reference: device bytes one series (N, L, k) takes in a batch of
 * oversampling p, q power iterations, precision (0 plain, 1 mixed) and horizon h; host
 * arithmetic only.  For callers that admit work against a device-memory budget (serve.py). */
long long ssa_amd_series_bytes(int N, int L, int k, int p, int q, int precision, int h);
/* one eager step with HIP events around every kernel launch: per launch the
 * phase (0 spectra, 1 matvec H, 2 matvec H^T, 3 Gram, 4 Cholesky, 5 apply,
 * 6 small SVD, 7 U/V, 8 reconstruction, 9 overlap-add, 10 forecast, 11 test
 * matrix), the FFT length log2 (or panel id / column blocks) and the
 * milliseconds; branch after branch; returns the number of launches */
int ssa_amd_batch_profile(ssa_amd_batch *b, int max_rec, int *phase, int *lgn, float *ms);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the reconstruction of ssa_opt_reconstruct (FP64) from
 * factors the caller holds on the device -- U (L x k, column-major), V
 * (K x k, column-major), sigma (k) -- into the device array output (N
 * values).  group is a host array (indices outside 0..k-1 are skipped, as in
 * ssa_opt_reconstruct).  Work space comes from a static 32 MiB device array
 * per GPU; nothing is allocated.  The work is queued on stream (NULL: the null
 * stream) and the call returns without waiting.  Calls that overlap on one GPU
 * share that array and race.  Returns 0; -1 on invalid arguments (including
 * n_group < 1 and pointers not 8-byte aligned); -2 when the series needs more
 * work space than the array holds (N above about 210000 at L = N/4, 180000
 * at L = N/2) or n_group exceeds 256; a HIP error code otherwise. */
int ssa_amd_reconstruct_noalloc(const ssa_real *U, const ssa_real *V, const ssa_real *sigma, int N, int L, int k,
                                const int *group, int n_group, ssa_real *output, void *stream);
/* plan of series i: n, CL, CK, AL, SK, b, k, log2 n */
int ssa_amd_batch_plan(const ssa_amd_batch *b, int i, int *out8);
void ssa_amd_batch_free(ssa_amd_batch *b);

#ifdef __cplusplus
}
#endif
#endif
