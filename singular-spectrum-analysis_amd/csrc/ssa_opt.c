/*
 * ssa_opt.c -- the reference's C API (include/ssa_opt.h) on the device engine.
 *
 * Every entry point keeps the argument meaning, validation and return
 * convention of the reference function it replaces (cited per function);
 * the arithmetic runs in the HIP kernels through ssa_engine.  Host code here
 * only validates, mirrors results for C callers and does the O(k) / O(L)
 * bookkeeping the reference also does on the host (getters, statistics,
 * LRF coefficients handed back to the caller).
 */
#define _DEFAULT_SOURCE   /* sysconf(_SC_NPROCESSORS_ONLN) */
#include "../../include/ssa_opt.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "ssa_engine.h"

#define SSA_MAX_H 65536        /* forecast horizon kept in the output slot */

typedef struct {
    ssa_engine *eng;
    double *x;                 /* host copy of the signal(s) */
    double *hostblk;           /* host views: fft_x and ws_* */
    int injected;              /* width b of an Omega injected for the next decomposition (0: none) */
    int pending_q;             /* q override of the next decomposition (-1: ctx q) */
    int no_mirror;             /* 1: decompositions leave the host U/V mirrors stale */
    int uv_stale;              /* the host U/V mirrors lag the device's */
} ssa_ctx;

static int imin(int a, int b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }

static int env_int(const char *name, int dflt)
{
    const char *v = getenv(name);
    return (v && *v) ? atoi(v) : dflt;
}

static int env_precision(void)
{
    const char *v = getenv("SSA_AMD_PRECISION");
    if (!v) return 0;
    return (!strcmp(v, "mixed") || !strcmp(v, "1") || !strcmp(v, "fp32")) ? 1 : 0;
}

static void weights(double *w, int N, int L, int K)
{
    for (int t = 0; t < N; t++) {
        int c = imin(imin(t + 1, L), imin(K, N - t));
        w[t] = c > 0 ? 1.0 / c : 0.0;
    }
}

/* ------------------------------------------------------------------------ */
/* SSA lifecycle                                                             */
/* ------------------------------------------------------------------------ */

static int build_engine(SSA_Opt *ssa, ssa_ctx *c)
{
    int err = 0;
    int N = ssa->N, L = ssa->L, M = 1;
    c->eng = ssa_engine_create(ssa->precision, 1, &N, &L, &M, NULL, NULL, ssa->q, &err);
    if (!c->eng) return -1;
    if (ssa_engine_upload_signal(c->eng, 0, c->x) || ssa_engine_spectra(c->eng)) return -1;
    const ssa_desc *d = ssa_engine_desc(c->eng, 0);
    ssa->fft_len = d->n;
    ssa->r2c_len = d->n / 2 + 1;
    ssa->batch_size = 2;
    if (ssa_engine_read_spectrum(c->eng, 0, ssa->fft_x, ssa->r2c_len)) return -1;
    return 0;
}

/* MKL/ssa_opt.h:1026-1284 */
int ssa_opt_init(SSA_Opt *ssa, const ssa_real *x, int N, int L)
{
    if (!ssa || !x || N < 4 || L < 2 || L > N - 1) return -1;
    memset(ssa, 0, sizeof(SSA_Opt));
    ssa->N = N;
    ssa->L = L;
    ssa->K = N - L + 1;
    ssa->q = env_int("SSA_AMD_Q", 2);
    if (ssa->q < 0) ssa->q = 0;
    ssa->precision = env_precision();
    ssa->seed = 42;
    ssa_ctx *c = (ssa_ctx *)calloc(1, sizeof(ssa_ctx));
    if (!c) return -1;
    ssa->dev = c;
    c->pending_q = -1;
    c->x = (double *)malloc(sizeof(double) * N);
    ssa->inv_diag_count = (double *)malloc(sizeof(double) * N);
    /* host views: fft_x (up to 4097 complex) + 6 small scratch views */
    c->hostblk = (double *)calloc(2 * 4097 + 6 * 8, sizeof(double));
    if (!c->x || !ssa->inv_diag_count || !c->hostblk) { ssa_opt_free(ssa); return -1; }
    memcpy(c->x, x, sizeof(double) * N);
    weights(ssa->inv_diag_count, N, L, ssa->K);
    ssa->fft_x = c->hostblk;
    double *w = c->hostblk + 2 * 4097;
    ssa->ws_real = w; ssa->ws_complex = w + 8; ssa->ws_real2 = w + 16;
    ssa->ws_proj = w + 24; ssa->ws_batch_real = w + 32; ssa->ws_batch_complex = w + 40;
    if (build_engine(ssa, c)) { ssa_opt_free(ssa); return -1; }
    ssa->initialized = true;
    return 0;
}

static void free_results(SSA_Opt *ssa)
{
    ssa_engine_host_free(ssa->U); ssa_engine_host_free(ssa->V); free(ssa->sigma); free(ssa->eigenvalues);
    ssa->U = ssa->V = ssa->sigma = ssa->eigenvalues = NULL;
}

/* MKL/ssa_opt.h:1286-1350: safe on a zeroed or failed context */
void ssa_opt_free(SSA_Opt *ssa)
{
    if (!ssa) return;
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    if (c) {
        if (c->eng) ssa_engine_destroy(c->eng);
        free(c->x);
        free(c->hostblk);
        free(c);
    }
    free_results(ssa);
    free(ssa->inv_diag_count);
    memset(ssa, 0, sizeof(SSA_Opt));
}

/* MKL/ssa_opt.h:1352-1379 */
void ssa_opt_free_prepared(SSA_Opt *ssa)
{
    if (!ssa) return;
    ssa->prepared_kp = 0;
    ssa->prepared_lwork = 0;
    ssa->prepared = false;
}

static int prepare_caps(SSA_Opt *ssa, int kp, int k)
{
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    if (kp > SSA_BMAX || kp < 1) return -1;
    int b = imax(kp, 2);
    if (ssa_engine_set_caps(c->eng, 0, b, imin(k, b))) return -1;
    if (ssa_engine_prepare(c->eng, imin(k, b), SSA_MAX_H)) return -1;
    free_results(ssa);
    /* page-locked host mirrors: every decomposition copies U and V into them */
    ssa->U = ssa_engine_host_alloc((size_t)ssa->L * k);
    ssa->V = ssa_engine_host_alloc((size_t)ssa->K * k);
    ssa->sigma = (double *)calloc(k, sizeof(double));
    ssa->eigenvalues = (double *)calloc(k, sizeof(double));
    if (!ssa->U || !ssa->V || !ssa->sigma || !ssa->eigenvalues) return -1;
    c->uv_stale = 0;
    ssa->n_components = k;
    ssa->prepared_kp = kp;
    ssa->prepared = true;
    ssa->decomposed = false;
    return 0;
}

/* MKL/ssa_opt.h:1414-1543: kp = min(max_k + p, L, K), p <= 0 -> 8 */
int ssa_opt_prepare(SSA_Opt *ssa, int max_k, int oversampling)
{
    if (!ssa || !ssa->initialized || max_k < 1) return -1;
    ssa_opt_free_prepared(ssa);
    int p = oversampling <= 0 ? 8 : oversampling;
    int kp = imin(max_k + p, imin(ssa->L, ssa->K));
    int k = imin(max_k, kp);
    return prepare_caps(ssa, kp, k);
}

/* MKL/ssa_opt.h:1545-1558 */
int ssa_opt_update_signal(SSA_Opt *ssa, const ssa_real *new_x)
{
    if (!ssa || !ssa->initialized || !new_x) return -1;
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    memcpy(c->x, new_x, sizeof(double) * ssa->N);
    ssa_opt_free_cached_ffts(ssa);
    ssa->decomposed = false;
    if (ssa_engine_upload_signal(c->eng, 0, c->x) || ssa_engine_spectra(c->eng)) return -1;
    if (ssa_engine_read_spectrum(c->eng, 0, ssa->fft_x, ssa->r2c_len)) return -1;
    return 0;
}

/* the context holds the window (c->x); a tick moves it by m samples */
int ssa_amd_push_samples(SSA_Opt *ssa, const ssa_real *x_new, int m)
{
    if (!ssa || !ssa->initialized || !x_new || m < 1 || m > ssa->N) return -1;
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    const int N = ssa->N;
    memmove(c->x, c->x + m, sizeof(double) * (size_t)(N - m));
    memcpy(c->x + (N - m), x_new, sizeof(double) * (size_t)m);
    ssa_opt_free_cached_ffts(ssa);
    ssa->decomposed = false;
    if (ssa_engine_upload_signal(c->eng, 0, c->x) || ssa_engine_spectra(c->eng)) return -1;
    if (ssa_engine_read_spectrum(c->eng, 0, ssa->fft_x, ssa->r2c_len)) return -1;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* decomposition                                                             */
/* ------------------------------------------------------------------------ */

static int run_decomposition(SSA_Opt *ssa, int kp, int k, int q)
{
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    if (c->injected && c->injected != kp) { c->injected = 0; return -1; }
    if (ssa_engine_set_bk(c->eng, 0, imax(kp, 2), k)) return -1;
    if (ssa_engine_set_rng(c->eng, 0, ssa->seed, ssa->calls)) return -1;
    int rc = ssa_engine_decompose(c->eng, q, c->injected);
    c->injected = 0;
    ssa->calls++;
    if (rc) return -1;
    double tv = 0;
    const int mirror = !c->no_mirror;
    if (ssa_engine_results(c->eng, 0, ssa->sigma, ssa->eigenvalues, &tv, mirror ? ssa->U : NULL,
                           mirror ? ssa->V : NULL, NULL) != k)
        return -1;
    c->uv_stale = !mirror;
    ssa->total_variance = tv;
    ssa->n_components = k;
    ssa->decomposed = true;
    return 0;
}

/* MKL/ssa_opt.h:1685-1763; requires prepare, kp <= prepared_kp */
int ssa_opt_decompose_randomized(SSA_Opt *ssa, int k, int oversampling)
{
    if (!ssa || !ssa->initialized || k < 1) return -1;
    ssa_opt_free_cached_ffts(ssa);
    int p = oversampling <= 0 ? 8 : oversampling;
    int kp = imin(k + p, imin(ssa->L, ssa->K));
    k = imin(k, kp);
    if (!ssa->prepared || kp > ssa->prepared_kp) return -1;
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    int q = c->pending_q >= 0 ? c->pending_q : ssa->q;
    c->pending_q = -1;
    return run_decomposition(ssa, kp, k, q);
}

static int subspace_decompose(SSA_Opt *ssa, int k, int kp, int max_iter)
{
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    if (!ssa->prepared || kp > ssa->prepared_kp || k > ssa->n_components) {
        if (prepare_caps(ssa, kp, k)) return -1;
    }
    int q = max_iter / 4;
    if (q < 2) q = 2;
    if (q > 30) q = 30;
    c->pending_q = q;
    int rc = run_decomposition(ssa, kp, k, q);
    c->pending_q = -1;
    return rc;
}

/* MKL/ssa_opt.h:1560-1678 (argument checks); device subspace iteration */
int ssa_opt_decompose(SSA_Opt *ssa, int k, int max_iter)
{
    if (!ssa || !ssa->initialized || k < 1) return -1;
    ssa_opt_free_cached_ffts(ssa);
    int r = imin(ssa->L, ssa->K);
    k = imin(k, r);
    int kp = imin(k + 8, r);
    if (kp > SSA_BMAX) return -1;
    return subspace_decompose(ssa, k, imax(kp, imin(2, r)), max_iter);
}

/* MKL/ssa_opt.h:1765-1933 (argument checks); device subspace iteration */
int ssa_opt_decompose_block(SSA_Opt *ssa, int k, int block_size, int max_iter)
{
    if (!ssa || !ssa->initialized || k < 1 || block_size < 1) return -1;
    ssa_opt_free_cached_ffts(ssa);
    int r = imin(ssa->L, ssa->K);
    k = imin(k, r);
    int kp = imin(imax(k + 8, block_size), r);
    if (kp > SSA_BMAX) kp = imin(imax(k + 8, k), imin(SSA_BMAX, r));
    if (kp > SSA_BMAX) return -1;
    return subspace_decompose(ssa, k, imax(kp, imin(2, r)), max_iter);
}

/* MKL/ssa_opt.h:1935-2056: grow to n_components + additional_k */
int ssa_opt_extend(SSA_Opt *ssa, int additional_k, int max_iter)
{
    if (!ssa || !ssa->decomposed || additional_k < 1) return -1;
    return ssa_opt_decompose(ssa, ssa->n_components + additional_k, max_iter);
}

static int ensure_prepared_minimal(SSA_Opt *ssa)
{
    if (ssa->prepared) return 0;
    return prepare_caps(ssa, 2, 1);
}

void ssa_opt_hankel_matvec(SSA_Opt *ssa, const ssa_real *v, ssa_real *y)
{
    if (!ssa || !ssa->initialized || !v || !y) return;
    if (ensure_prepared_minimal(ssa)) return;
    ssa_engine_matvec(((ssa_ctx *)ssa->dev)->eng, 0, 0, v, y);
}

void ssa_opt_hankel_matvec_T(SSA_Opt *ssa, const ssa_real *u, ssa_real *y)
{
    if (!ssa || !ssa->initialized || !u || !y) return;
    if (ensure_prepared_minimal(ssa)) return;
    ssa_engine_matvec(((ssa_ctx *)ssa->dev)->eng, 0, 1, u, y);
}

/* ------------------------------------------------------------------------ */
/* reconstruction and analysis                                               */
/* ------------------------------------------------------------------------ */

/* MKL/ssa_opt.h:2325-2400: out-of-range indices skipped; n_group < 1 fails */
int ssa_opt_reconstruct(const SSA_Opt *ssa, const int *group, int n_group, ssa_real *output)
{
    if (!ssa || !ssa->decomposed || !group || !output || n_group < 1) return -1;
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    return ssa_engine_reconstruct(c->eng, 0, 0, group, &n_group, 1, 1, 0, output);
}

int ssa_amd_reconstruct_many(const SSA_Opt *ssa, const int *glist, const int *gcount, int n_groups,
                             ssa_real *output)
{
    if (!ssa || !ssa->decomposed || !glist || !gcount || !output || n_groups < 1) return -1;
    for (int g = 0; g < n_groups; g++) if (gcount[g] < 0) return -1;
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    return ssa_engine_reconstruct(c->eng, 0, 0, glist, gcount, n_groups, 1, 0, output);
}

/* MKL/ssa_opt.h:2110-2323: the transforms of every component's U and V
 * chunks, kept in HBM (ssa_engine_cache_build) until the next decomposition,
 * update_signal or free_cached_ffts.  While it is valid, reconstructions and
 * wcorr_matrix_fast read the cached spectra instead of transforming U and V
 * again.  A second call rebuilds it (MKL/ssa_opt.h:2115-2116). */
int ssa_opt_cache_ffts(SSA_Opt *ssa)
{
    if (!ssa || !ssa->decomposed) return -1;
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    ssa_opt_free_cached_ffts(ssa);
    if (ssa_engine_cache_build(c->eng)) {
        ssa_engine_cache_free(c->eng);
        return -1;
    }
    ssa->fft_cached = true;
    return 0;
}

/* MKL/ssa_opt.h:2058-2078; safe without a cache */
void ssa_opt_free_cached_ffts(SSA_Opt *ssa)
{
    if (!ssa) return;
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    if (c && c->eng && ssa->fft_cached) ssa_engine_cache_free(c->eng);
    ssa->fft_cached = false;
}

/* MKL/ssa_opt.h:2407-2477: per-component transforms formed on the fly */
int ssa_opt_wcorr_matrix(const SSA_Opt *ssa, ssa_real *W)
{
    if (!ssa || !ssa->decomposed || !W) return -1;
    return ssa_engine_wcorr(((ssa_ctx *)ssa->dev)->eng, 0, 0, W);
}

/* MKL/ssa_opt.h:2492-2548: from the cached component spectra; without a
 * cache, the slow path (as the reference falls back, :2497-2498) */
int ssa_opt_wcorr_matrix_fast(const SSA_Opt *ssa, ssa_real *W)
{
    if (!ssa || !ssa->decomposed || !W) return -1;
    return ssa_engine_wcorr(((ssa_ctx *)ssa->dev)->eng, 0, ssa->fft_cached ? 1 : 0, W);
}

/* MKL/ssa_opt.h:2551-2625 */
ssa_real ssa_opt_wcorr_pair(const SSA_Opt *ssa, int i, int j)
{
    if (!ssa || !ssa->decomposed || i < 0 || i >= ssa->n_components || j < 0 || j >= ssa->n_components) return 0.0;
    int k = ssa->n_components;
    double *W = (double *)malloc(sizeof(double) * k * k);
    if (!W) return 0.0;
    double r = 0.0;
    if (ssa_opt_wcorr_matrix(ssa, W) == 0) r = W[(size_t)i * k + j];
    free(W);
    return r;
}

/* log(v) with the reference's guard against log(0) (MKL/ssa_opt.h:2647) */
static double guarded_log(double v) { return log(v + 1e-300); }

/* index of the first largest of n values (n >= 1) */
static int first_max(const double *v, int n)
{
    int at = 0;
    for (int i = 1; i < n; i++) at = v[i] > v[at] ? i : at;
    return at;
}

/* MKL/ssa_opt.h:2627-2676.  The five arrays share one allocation owned by
 * singular_values (released by ssa_opt_component_stats_free):
 *   log_sv[i] = log(sigma_i), gaps[i] = sigma_i / sigma_{i+1} (i < n-1),
 *   cumulative_var[i] = sum_{j<=i} lambda_j / total_variance,
 *   second_diff[i] = log_sv[i-1] - 2 log_sv[i] + log_sv[i+1] inside, 0 at both ends;
 * the suggested signal rank is one past the first largest gap. */
int ssa_opt_component_stats(const SSA_Opt *ssa, SSA_ComponentStats *stats)
{
    if (!ssa || !ssa->decomposed || !stats) return -1;
    const int n = ssa->n_components;
    if (n < 2) return -1;
    memset(stats, 0, sizeof(SSA_ComponentStats));
    double *blk = (double *)calloc(5 * (size_t)n, sizeof(double));
    if (!blk) return -1;
    stats->n = n;
    stats->singular_values = blk;
    stats->log_sv = blk + n;
    stats->gaps = blk + 2 * n;
    stats->cumulative_var = blk + 3 * n;
    stats->second_diff = blk + 4 * n;
    memcpy(stats->singular_values, ssa->sigma, sizeof(double) * (size_t)n);
    double run = 0.0;
    for (int i = 0; i < n; i++) {
        stats->log_sv[i] = guarded_log(ssa->sigma[i]);
        run += ssa->eigenvalues[i];
        stats->cumulative_var[i] = run / ssa->total_variance;
        if (i + 1 < n) stats->gaps[i] = ssa->sigma[i] / (ssa->sigma[i + 1] + 1e-300);
    }
    for (int i = 1; i + 1 < n; i++)
        stats->second_diff[i] = (stats->log_sv[i - 1] + stats->log_sv[i + 1]) - 2.0 * stats->log_sv[i];
    const int widest = first_max(stats->gaps, n - 1);
    stats->gap_threshold = stats->gaps[widest];
    stats->suggested_signal = widest + 1;
    return 0;
}

void ssa_opt_component_stats_free(SSA_ComponentStats *stats)
{
    if (!stats) return;
    free(stats->singular_values);   /* the block holding all five arrays */
    memset(stats, 0, sizeof(SSA_ComponentStats));
}

/* MKL/ssa_opt.h:2690-2756: greedy pairing by singular-value ratio and |W| */
int ssa_opt_find_periodic_pairs(const SSA_Opt *ssa, int *pairs, int max_pairs, ssa_real sv_tol,
                                ssa_real wcorr_thresh)
{
    if (!ssa || !ssa->decomposed || !pairs || max_pairs < 1) return 0;
    if (sv_tol <= 0) sv_tol = 0.1;
    if (wcorr_thresh <= 0) wcorr_thresh = 0.5;
    int n = ssa->n_components, np = 0;
    double *W = (double *)malloc(sizeof(double) * n * n);
    char *used = (char *)calloc(n, 1);
    if (!W || !used || ssa_opt_wcorr_matrix(ssa, W)) { free(W); free(used); return 0; }
    for (int i = 0; i < n - 1 && np < max_pairs; i++) {
        if (used[i]) continue;
        for (int j = i + 1; j < n && np < max_pairs; j++) {
            if (used[j]) continue;
            double ratio = ssa->sigma[j] / (ssa->sigma[i] + 1e-300);
            if (fabs(1.0 - ratio) > sv_tol) continue;
            if (fabs(W[(size_t)i * n + j]) < wcorr_thresh) continue;
            pairs[2 * np] = i;
            pairs[2 * np + 1] = j;
            used[i] = used[j] = 1;
            np++;
            break;
        }
    }
    free(W);
    free(used);
    return np;
}

/* MKL/ssa_opt.h:3086-3090 */
int ssa_opt_get_trend(const SSA_Opt *ssa, ssa_real *output)
{
    int g[1] = {0};
    return ssa_opt_reconstruct(ssa, g, 1, output);
}

/* MKL/ssa_opt.h:3091-3110 */
int ssa_opt_get_noise(const SSA_Opt *ssa, int noise_start, ssa_real *output)
{
    if (!ssa || !ssa->decomposed || noise_start < 0) return -1;
    int nn = ssa->n_components - noise_start;
    if (nn <= 0) {
        memset(output, 0, sizeof(double) * ssa->N);
        return 0;
    }
    int *g = (int *)malloc(sizeof(int) * nn);
    if (!g) return -1;
    for (int i = 0; i < nn; i++) g[i] = noise_start + i;
    int rc = ssa_opt_reconstruct(ssa, g, nn, output);
    free(g);
    return rc;
}

/* MKL/ssa_opt.h:3111-3121 */
ssa_real ssa_opt_variance_explained(const SSA_Opt *ssa, int start, int end)
{
    if (!ssa || !ssa->decomposed || start < 0 || ssa->total_variance <= 0) return 0.0;
    if (end < 0 || end >= ssa->n_components) end = ssa->n_components - 1;
    double s = 0;
    for (int i = start; i <= end; i++) s += ssa->eigenvalues[i];
    return s / ssa->total_variance;
}

int ssa_opt_get_singular_values(const SSA_Opt *ssa, ssa_real *output, int max_n)   /* :3123-3130 */
{
    if (!ssa || !ssa->decomposed || !output) return -1;
    int n = imin(max_n, ssa->n_components);
    if (n > 0) memcpy(output, ssa->sigma, sizeof(double) * n);
    return n;
}

int ssa_opt_get_eigenvalues(const SSA_Opt *ssa, ssa_real *output, int max_n)       /* :3132-3139 */
{
    if (!ssa || !ssa->decomposed || !output) return -1;
    int n = imin(max_n, ssa->n_components);
    if (n > 0) memcpy(output, ssa->eigenvalues, sizeof(double) * n);
    return n;
}

ssa_real ssa_opt_get_total_variance(const SSA_Opt *ssa)                            /* :3141-3145 */
{
    if (!ssa || !ssa->decomposed) return 0.0;
    return ssa->total_variance;
}

/* ------------------------------------------------------------------------ */
/* forecasting                                                               */
/* ------------------------------------------------------------------------ */

/* MKL/ssa_opt.h:2766-2816: coefficients handed back to the caller, from the
 * host mirror of U */
int ssa_opt_compute_lrf(const SSA_Opt *ssa, const int *group, int n_group, SSA_LRF *lrf)
{
    if (!ssa || !ssa->decomposed || !group || !lrf || n_group < 1) return -1;
    if (ssa_amd_sync_mirrors(ssa)) return -1;
    int L = ssa->L;
    memset(lrf, 0, sizeof(SSA_LRF));
    lrf->L = L;
    double nu = 0.0;
    for (int g = 0; g < n_group; g++) {
        int idx = group[g];
        if (idx < 0 || idx >= ssa->n_components) return -1;
        double pi = ssa->U[(size_t)idx * L + L - 1];
        nu += pi * pi;
    }
    lrf->verticality = nu;
    if (nu >= 1.0 - 1e-10) return -1;
    lrf->R = (double *)malloc(sizeof(double) * (L > 1 ? L - 1 : 1));
    if (!lrf->R) return -1;
    double sc = 1.0 / (1.0 - nu);
    for (int j = 0; j < L - 1; j++) {
        double s = 0.0;
        for (int g = 0; g < n_group; g++) {
            int idx = group[g];
            s += ssa->U[(size_t)idx * L + L - 1] * ssa->U[(size_t)idx * L + j];
        }
        lrf->R[j] = s * sc;
    }
    lrf->valid = true;
    return 0;
}

void ssa_opt_lrf_free(SSA_LRF *lrf)   /* MKL/ssa_opt.h:2758-2764 */
{
    if (!lrf) return;
    free(lrf->R);
    memset(lrf, 0, sizeof(SSA_LRF));
}

/* MKL/ssa_opt.h:2818-2841 */
int ssa_opt_forecast_with_lrf(const SSA_LRF *lrf, const ssa_real *base_signal, int base_len, int n_forecast,
                              ssa_real *output)
{
    if (!lrf || !lrf->valid || !lrf->R || !base_signal || !output) return -1;
    int L = lrf->L, w = L - 1;
    if (base_len < w || n_forecast < 1) return -1;
    double *buf = (double *)malloc(sizeof(double) * ((size_t)w + n_forecast));
    if (!buf) return -1;
    for (int i = 0; i < w; i++) buf[i] = base_signal[base_len - w + i];
    for (int h = 0; h < n_forecast; h++) {
        double f = 0.0;
        for (int j = 0; j < w; j++) f += lrf->R[j] * buf[h + j];
        buf[w + h] = f;
        output[h] = f;
    }
    free(buf);
    return 0;
}

/* MKL/ssa_opt.h:2843-2867: device-resident reconstruction, LRF and recurrence */
int ssa_opt_forecast(const SSA_Opt *ssa, const int *group, int n_group, int n_forecast, ssa_real *output)
{
    if (!ssa || !ssa->decomposed || !group || !output || n_group < 1 || n_forecast < 1) return -1;
    if (n_forecast > SSA_MAX_H) return -1;
    return ssa_engine_forecast(((ssa_ctx *)ssa->dev)->eng, 0, group, n_group, n_forecast, output, 0, 0, NULL, 0);
}

/* MKL/ssa_opt.h:2869-2885 */
int ssa_opt_forecast_full(const SSA_Opt *ssa, const int *group, int n_group, int n_forecast, ssa_real *output)
{
    if (!ssa || !ssa->decomposed || !group || !output || n_group < 1 || n_forecast < 1) return -1;
    if (n_forecast > SSA_MAX_H) return -1;
    return ssa_engine_forecast(((ssa_ctx *)ssa->dev)->eng, 0, group, n_group, n_forecast, output, 1, 0, NULL, 0);
}

/* V-forecast (MKL/ssa_opt.h:2906-2965): the same recurrence as the R-forecast,
 * computed as it is defined there, one projection per eigenvector and step,
 * x = (1/(1 - nu^2)) sum_g pi_g (U_g[0:L-1] . z).  Every index of the group must
 * be a computed component and nu^2 < 1 - 1e-10, else -1. */
int ssa_opt_vforecast(const SSA_Opt *ssa, const int *group, int n_group, int n_forecast, ssa_real *output)
{
    if (!ssa || !ssa->decomposed || !group || !output || n_group < 1 || n_forecast < 1) return -1;
    if (n_forecast > SSA_MAX_H) return -1;
    return ssa_engine_forecast(((ssa_ctx *)ssa->dev)->eng, 0, group, n_group, n_forecast, output, 0, 1, NULL, 0);
}

/* MKL/ssa_opt.h:2967-3014: reconstruction followed by the V-forecast, N + h values */
int ssa_opt_vforecast_full(const SSA_Opt *ssa, const int *group, int n_group, int n_forecast, ssa_real *output)
{
    if (!ssa || !ssa->decomposed || !group || !output || n_group < 1 || n_forecast < 1) return -1;
    if (n_forecast > SSA_MAX_H) return -1;
    return ssa_engine_forecast(((ssa_ctx *)ssa->dev)->eng, 0, group, n_group, n_forecast, output, 1, 1, NULL, 0);
}

/* MKL/ssa_opt.h:3016-3078: V-forecast continuing the caller's base signal (its
 * last L-1 values; base_len >= L-1) instead of the group's reconstruction */
int ssa_opt_vforecast_fast(const SSA_Opt *ssa, const int *group, int n_group, const ssa_real *base_signal,
                           int base_len, int n_forecast, ssa_real *output)
{
    if (!ssa || !ssa->decomposed || !group || !base_signal || !output) return -1;
    if (n_group < 1 || n_forecast < 1 || base_len < ssa->L - 1) return -1;
    if (n_forecast > SSA_MAX_H) return -1;
    return ssa_engine_forecast(((ssa_ctx *)ssa->dev)->eng, 0, group, n_group, n_forecast, output, 0, 1, base_signal,
                               base_len);
}

/* ------------------------------------------------------------------------ */
/* Cadzow iterations and gap filling (MKL/ssa_opt.h:3147-3878)               */
/* ------------------------------------------------------------------------ */

/* A rank-`rank` decomposition as the reference's iterated callers make it
 * (MKL/ssa_opt.h:3206-3232): randomized with oversampling 8 when
 * rank + 8 < min(L, K) / 2, else the block method.  `fresh` restarts the
 * test-matrix counter: the reference builds a new context (and so restarts its
 * random stream) for every iteration.  `may_prepare` = 0 is the in-place
 * variant, which takes the randomized method only on an already prepared
 * context wide enough for it (:3370-3377). */
static int iter_decompose(SSA_Opt *ssa, int rank, int fresh, int may_prepare)
{
    int r = imin(ssa->L, ssa->K);
    int use_rand = rank + 8 < r / 2;
    if (use_rand && !may_prepare && !(ssa->prepared && rank + 8 <= ssa->prepared_kp)) use_rand = 0;
    if (fresh) ssa->calls = 0;
    if (use_rand && ((may_prepare && ssa_opt_prepare(ssa, rank, 8) != 0) ||
                     ssa_opt_decompose_randomized(ssa, rank, 8) != 0))
        use_rand = 0;
    if (!use_rand) {
        if (fresh) ssa->calls = 0;
        if (ssa_opt_decompose_block(ssa, rank, imin(rank, 16), 3) != 0) return -1;
    }
    return 0;
}

static double norm2(const double *x, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i] * x[i];
    return sqrt(s);
}

/* MKL/ssa_opt.h:3165-3292 */
int ssa_opt_cadzow(const ssa_real *x, int N, int L, int rank, int max_iter, ssa_real tol, ssa_real *output,
                   SSA_CadzowResult *result)
{
    if (!x || !output || N < 4 || L < 2 || L > N - 1 || rank < 1 || max_iter < 1) return -1;
    double *y = (double *)malloc(sizeof(double) * N), *y_new = (double *)malloc(sizeof(double) * N);
    int *group = (int *)malloc(sizeof(int) * rank);
    SSA_Opt ssa;
    memset(&ssa, 0, sizeof(ssa));
    int iter = -1;
    if (!y || !y_new || !group) goto out;
    for (int i = 0; i < rank; i++) group[i] = i;
    memcpy(y, x, sizeof(double) * N);
    double y_norm = norm2(y, N);
    if (y_norm < 1e-15) y_norm = 1.0;
    if (ssa_opt_init(&ssa, y, N, L) != 0) goto out;
    double diff = 0.0;
    bool converged = false;
    int it;
    for (it = 0; it < max_iter; it++) {
        if (it > 0 && ssa_opt_update_signal(&ssa, y) != 0) goto out;
        if (iter_decompose(&ssa, rank, 1, 1) != 0 || ssa_opt_reconstruct(&ssa, group, rank, y_new) != 0) goto out;
        double d2 = 0.0;
        for (int i = 0; i < N; i++) {
            double d = y_new[i] - y[i];
            d2 += d * d;
        }
        diff = sqrt(d2) / y_norm;
        double *t = y;
        y = y_new;
        y_new = t;
        if (diff < tol) {
            converged = true;
            it++;
            break;
        }
    }
    memcpy(output, y, sizeof(double) * N);
    if (result) {
        result->iterations = it;
        result->final_diff = diff;
        result->converged = converged ? 1.0 : 0.0;
    }
    iter = it;
out:
    ssa_opt_free(&ssa);
    free(y); free(y_new); free(group);
    return iter;
}

/* MKL/ssa_opt.h:3294-3315 */
int ssa_opt_cadzow_weighted(const ssa_real *x, int N, int L, int rank, int max_iter, ssa_real tol, ssa_real alpha,
                            ssa_real *output, SSA_CadzowResult *result)
{
    if (!(alpha > 0.0) || alpha > 1.0) return -1;
    int ret = ssa_opt_cadzow(x, N, L, rank, max_iter, tol, output, result);
    if (ret < 0) return ret;
    if (alpha < 1.0) {
        double beta = 1.0 - alpha;
        for (int i = 0; i < N; i++) output[i] = alpha * output[i] + beta * x[i];
    }
    return ret;
}

/* MKL/ssa_opt.h:3319-3440: the context's own signal, re-decomposed in place.
 * The first reconstruction sets the scale; later ones are compared with the
 * previous reconstruction and become the signal for the next step. */
int ssa_opt_cadzow_inplace(SSA_Opt *ssa, int rank, int max_iter, ssa_real tol, SSA_CadzowResult *result)
{
    if (!ssa || !ssa->initialized || rank < 1 || max_iter < 1) return -1;
    int N = ssa->N;
    double *y_new = (double *)malloc(sizeof(double) * N), *y_cur = (double *)malloc(sizeof(double) * N);
    int *group = (int *)malloc(sizeof(int) * rank);
    int iter = -1;
    if (!y_new || !y_cur || !group) goto out;
    for (int i = 0; i < rank; i++) group[i] = i;
    double y_norm = 1.0, diff = 0.0;
    bool converged = false;
    int it;
    for (it = 0; it < max_iter; it++) {
        ssa->decomposed = false;
        ssa->n_components = 0;
        ssa->total_variance = 0.0;
        ssa_opt_free_cached_ffts(ssa);
        if (iter_decompose(ssa, rank, 0, 0) != 0) goto out;
        if (it > 0) memcpy(y_cur, y_new, sizeof(double) * N);
        if (ssa_opt_reconstruct(ssa, group, rank, y_new) != 0) goto out;
        if (it == 0) {
            y_norm = norm2(y_new, N);
            if (y_norm < 1e-15) y_norm = 1.0;
        } else {
            double d2 = 0.0;
            for (int i = 0; i < N; i++) {
                double d = y_new[i] - y_cur[i];
                d2 += d * d;
            }
            diff = sqrt(d2) / y_norm;
            if (diff < tol) {
                converged = true;
                it++;
                break;
            }
        }
        if (it < max_iter - 1 && ssa_opt_update_signal(ssa, y_new) != 0) goto out;
    }
    if (result) {
        result->iterations = it;
        result->final_diff = diff;
        result->converged = converged ? 1.0 : 0.0;
    }
    iter = it;
out:
    free(y_new); free(y_cur); free(group);
    return iter;
}

/* MKL/ssa_opt.h:3455-3496: gaps start as the straight line between their
 * neighbours (a gap at an end takes the value of its one neighbour) */
static void interp_gaps(double *x, const unsigned char *gap, int N)
{
    int i = 0;
    while (i < N) {
        if (!gap[i]) {
            i++;
            continue;
        }
        int a = i;
        while (i < N && gap[i]) i++;
        int b = i, len = b - a;
        double lv, rv;
        if (a > 0) lv = x[a - 1];
        else lv = (b < N) ? x[b] : 0.0;
        rv = (b < N) ? x[b] : lv;
        for (int j = 0; j < len; j++) {
            double t = (len > 1) ? (double)(j + 1) / (len + 1) : 0.5;
            x[a + j] = lv + t * (rv - lv);
        }
    }
}

/* MKL/ssa_opt.h:3498-3683 */
int ssa_opt_gapfill(ssa_real *x, int N, int L, int rank, int max_iter, ssa_real tol, SSA_GapFillResult *result)
{
    if (!x || N < 4 || L < 2 || L > N - 1 || rank < 1) return -1;
    unsigned char *gap = (unsigned char *)calloc(N, 1);
    if (!gap) return -1;
    int n_gaps = 0;
    for (int i = 0; i < N; i++)
        if (isnan(x[i])) {
            gap[i] = 1;
            n_gaps++;
        }
    if (n_gaps == 0) {
        free(gap);
        if (result) {
            result->iterations = 0;
            result->final_diff = 0.0;
            result->converged = 1;
            result->n_gaps = 0;
        }
        return 0;
    }
    if (N - n_gaps < L + rank) {
        free(gap);
        return -1;
    }
    interp_gaps(x, gap, N);
    double *old_v = (double *)malloc(sizeof(double) * n_gaps), *new_v = (double *)malloc(sizeof(double) * n_gaps);
    double *rec = (double *)malloc(sizeof(double) * N);
    int *group = (int *)malloc(sizeof(int) * rank);
    SSA_Opt ssa;
    memset(&ssa, 0, sizeof(ssa));
    int rc = -1;
    if (!old_v || !new_v || !rec || !group) goto out;
    for (int i = 0; i < rank; i++) group[i] = i;
    for (int i = 0, g = 0; i < N; i++) if (gap[i]) old_v[g++] = x[i];
    double diff = 1e10;
    int converged = 0, it;
    for (it = 0; it < max_iter; it++) {
        /* a fresh decomposition of the current signal (:3591-3603) */
        if (it == 0 ? ssa_opt_init(&ssa, x, N, L) != 0 : ssa_opt_update_signal(&ssa, x) != 0) goto out;
        ssa.calls = 0;
        ssa_opt_prepare(&ssa, rank, 8);
        if (ssa_opt_decompose_randomized(&ssa, rank, 8) != 0 || ssa_opt_reconstruct(&ssa, group, rank, rec) != 0)
            goto out;
        for (int i = 0, g = 0; i < N; i++)
            if (gap[i]) {
                new_v[g++] = rec[i];
                x[i] = rec[i];
            }
        double sd = 0.0, so = 0.0;
        for (int g = 0; g < n_gaps; g++) {
            double d = new_v[g] - old_v[g];
            sd += d * d;
            so += old_v[g] * old_v[g];
        }
        diff = (so > 1e-12) ? sqrt(sd / so) : sqrt(sd);
        if (diff < tol) {
            converged = 1;
            it++;
            break;
        }
        memcpy(old_v, new_v, sizeof(double) * n_gaps);
    }
    if (result) {
        result->iterations = it;
        result->final_diff = diff;
        result->converged = converged;
        result->n_gaps = n_gaps;
    }
    rc = 0;
out:
    ssa_opt_free(&ssa);
    free(gap); free(old_v); free(new_v); free(rec); free(group);
    return rc;
}

/* forecast of h values continuing seg[0:len) from its rank-`rank` decomposition
 * (a fresh context, randomized, oversampling 8); 0 on success */
static int segment_forecast(const double *seg, int len, int L, int rank, const int *group, int h, double *out)
{
    SSA_Opt s;
    memset(&s, 0, sizeof(s));
    int rc = ssa_opt_init(&s, seg, len, L);
    if (rc == 0) {
        ssa_opt_prepare(&s, rank, 8);
        rc = ssa_opt_decompose_randomized(&s, rank, 8);
    }
    if (rc == 0) rc = ssa_opt_forecast(&s, group, rank, h, out);
    ssa_opt_free(&s);
    return rc;
}

/* MKL/ssa_opt.h:3685-3877.  Each gap in turn (left to right, earlier fills
 * included in later left segments): a forecast from the data before it when
 * that is at least L + rank long, a backcast (the forecast of the reversed
 * tail of the series) when the data after it is, blended with weights falling
 * linearly away from each side; neither: the mean of the two neighbours.  A
 * side whose forecast fails (verticality) counts as absent. */
int ssa_opt_gapfill_simple(ssa_real *x, int N, int L, int rank, SSA_GapFillResult *result)
{
    if (!x || N < 4 || L < 2 || L > N - 1 || rank < 1) return -1;
    unsigned char *gap = (unsigned char *)calloc(N, 1);
    int *group = (int *)malloc(sizeof(int) * rank);
    double *fl = (double *)malloc(sizeof(double) * N), *fr = (double *)malloc(sizeof(double) * N);
    double *rev = (double *)malloc(sizeof(double) * N), *tmp = (double *)malloc(sizeof(double) * N);
    int rc = -1, n_gaps = 0;
    if (!gap || !group || !fl || !fr || !rev || !tmp) goto out;
    for (int i = 0; i < rank; i++) group[i] = i;
    for (int i = 0; i < N; i++)
        if (isnan(x[i])) {
            gap[i] = 1;
            n_gaps++;
        }
    for (int i = 0; i < N;) {
        if (!gap[i]) {
            i++;
            continue;
        }
        int a = i;
        while (i < N && gap[i]) i++;
        int b = i, len = b - a;
        int left_len = a, right_len = N - b;
        int have_l = 0, have_r = 0;
        if (left_len >= L + rank) have_l = segment_forecast(x, left_len, L, rank, group, len, fl) == 0;
        if (right_len >= L + rank) {
            for (int j = 0; j < right_len; j++) rev[j] = x[N - 1 - j];
            have_r = segment_forecast(rev, right_len, L, rank, group, len, tmp) == 0;
            if (have_r) for (int j = 0; j < len; j++) fr[j] = tmp[len - 1 - j];
        }
        for (int j = 0; j < len; j++) {
            double v = 0.0, w = 0.0;
            if (have_l) {
                double wl = (double)(len - j) / len;
                v += wl * fl[j];
                w += wl;
            }
            if (have_r) {
                double wr = (double)(j + 1) / len;
                v += wr * fr[j];
                w += wr;
            }
            if (w > 0) x[a + j] = v / w;
            else {
                double lv = (a > 0) ? x[a - 1] : 0.0;
                double rv = (b < N) ? x[b] : lv;
                x[a + j] = (lv + rv) / 2.0;
            }
        }
    }
    if (result) {
        result->iterations = 1;
        result->final_diff = 0.0;
        result->converged = 1;
        result->n_gaps = n_gaps;
    }
    rc = 0;
out:
    free(gap); free(group); free(fl); free(fr); free(rev); free(tmp);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* MKL configuration hooks (MKL/py/ssa_wrapper.c:47-124)                      */
/* ------------------------------------------------------------------------ */
/* The reference's library exports these for its benches, which call them
 * inside try/except.  The device path has no MKL: there is nothing to
 * configure, the calls succeed and do nothing, and the CPU query reports the
 * host's online cores as performance cores. */
int ssa_mkl_init(int verbose) { (void)verbose; return 0; }
int ssa_mkl_init_14900kf(int verbose) { (void)verbose; return 0; }
int ssa_mkl_init_generic(int verbose) { (void)verbose; return 0; }
void ssa_mkl_set_threads(int n) { (void)n; }
int ssa_mkl_get_threads(void) { return 1; }
void ssa_mkl_get_cpu_info(int *p_cores, int *e_cores, int *is_hybrid, int *has_avx512)
{
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    if (p_cores) *p_cores = n > 0 ? (int)n : 1;
    if (e_cores) *e_cores = 0;
    if (is_hybrid) *is_hybrid = 0;
    if (has_avx512) *has_avx512 = 0;
}
void ssa_mkl_disable_affinity(void) {}

/* ------------------------------------------------------------------------ */
/* extensions                                                                */
/* ------------------------------------------------------------------------ */

/* The host U/V mirrors are refreshed by every decomposition unless
 * mirroring is off; then they are copied when host code needs them (LRF,
 * ESPRIT) or the caller asks. */
int ssa_amd_set_mirror(SSA_Opt *ssa, int on)
{
    if (!ssa || !ssa->initialized) return -1;
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    c->no_mirror = !on;
    return on ? ssa_amd_sync_mirrors(ssa) : 0;
}

int ssa_amd_sync_mirrors(const SSA_Opt *ssa)
{
    if (!ssa || !ssa->initialized) return -1;
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    if (!c->uv_stale) return 0;
    if (!ssa->decomposed) return -1;
    if (ssa_engine_results(c->eng, 0, NULL, NULL, NULL, ssa->U, ssa->V, NULL) != ssa->n_components) return -1;
    c->uv_stale = 0;
    return 0;
}

int ssa_amd_set_options(SSA_Opt *ssa, int q, int precision, unsigned long long seed)
{
    if (!ssa || !ssa->initialized || q < 0 || (precision != 0 && precision != 1)) return -1;
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    ssa->q = q;
    ssa->seed = seed;
    ssa->calls = 0;
    if (precision != ssa->precision) {
        ssa_engine_destroy(c->eng);
        c->eng = NULL;
        ssa->precision = precision;
        free_results(ssa);
        ssa->prepared = false;
        ssa->prepared_kp = 0;
        ssa->decomposed = false;
        ssa->fft_cached = false;   /* the cache went with the engine */
        ssa->n_components = 0;
        if (build_engine(ssa, c)) return -1;
    }
    return 0;
}

int ssa_amd_inject_omega(SSA_Opt *ssa, const ssa_real *Omega, int K, int b)
{
    if (!ssa || !ssa->prepared || !Omega || K != ssa->K || b < 1 || b > ssa->prepared_kp) return -1;
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    /* the engine's current width must be b for the copy */
    if (ssa_engine_set_bk(c->eng, 0, imax(b, 2), imin(ssa->n_components, imax(b, 2)))) return -1;
    if (b < 2) return -1;
    if (ssa_engine_inject_omega(c->eng, 0, Omega)) return -1;
    c->injected = b;
    return 0;
}

int ssa_amd_plan_info(const SSA_Opt *ssa, int *out6)
{
    if (!ssa || !ssa->initialized || !out6) return -1;
    ssa_ctx *c = (ssa_ctx *)ssa->dev;
    const ssa_desc *d = ssa_engine_desc(c->eng, 0);
    out6[0] = d->n; out6[1] = d->CL; out6[2] = d->CK; out6[3] = d->AL; out6[4] = d->SK;
    double sw = 0;
    out6[5] = 0;
    if (ssa->decomposed && ssa_engine_results(c->eng, 0, NULL, NULL, NULL, NULL, NULL, &sw) >= 0) out6[5] = (int)sw;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* MSSA (MKL/ssa_opt.h:3881-4283)                                            */
/* ------------------------------------------------------------------------ */

static int mssa_build(MSSA_Opt *m, ssa_ctx *c)
{
    int err = 0;
    c->eng = ssa_engine_create(m->precision, 1, &m->N, &m->L, &m->M, NULL, NULL, m->q, &err);
    if (!c->eng) return -1;
    if (ssa_engine_upload_signal(c->eng, 0, c->x) || ssa_engine_spectra(c->eng)) return -1;
    const ssa_desc *d = ssa_engine_desc(c->eng, 0);
    m->fft_len = d->n;
    m->r2c_len = d->n / 2 + 1;
    return 0;
}

/* MKL/ssa_opt.h:3907-4057 */
int mssa_opt_init(MSSA_Opt *mssa, const ssa_real *X, int M, int N, int L)
{
    if (!mssa || !X || M < 1 || N < 4 || L < 2 || L > N - 1) return -1;
    memset(mssa, 0, sizeof(MSSA_Opt));
    mssa->M = M; mssa->N = N; mssa->L = L; mssa->K = N - L + 1;
    mssa->q = env_int("SSA_AMD_Q", 2);
    if (mssa->q < 0) mssa->q = 0;
    mssa->precision = env_precision();
    mssa->seed = 42;
    ssa_ctx *c = (ssa_ctx *)calloc(1, sizeof(ssa_ctx));
    if (!c) return -1;
    mssa->dev = c;
    c->x = (double *)malloc(sizeof(double) * (size_t)M * N);
    mssa->inv_diag_count = (double *)malloc(sizeof(double) * N);
    if (!c->x || !mssa->inv_diag_count) { mssa_opt_free(mssa); return -1; }
    memcpy(c->x, X, sizeof(double) * (size_t)M * N);
    weights(mssa->inv_diag_count, N, L, mssa->K);
    if (mssa_build(mssa, c)) { mssa_opt_free(mssa); return -1; }
    mssa->initialized = true;
    return 0;
}

void mssa_opt_free(MSSA_Opt *mssa)   /* MKL/ssa_opt.h:3881-3905 */
{
    if (!mssa) return;
    ssa_ctx *c = (ssa_ctx *)mssa->dev;
    if (c) {
        if (c->eng) ssa_engine_destroy(c->eng);
        free(c->x);
        free(c);
    }
    ssa_engine_host_free(mssa->U); ssa_engine_host_free(mssa->V); free(mssa->sigma); free(mssa->eigenvalues);
    free(mssa->inv_diag_count);
    memset(mssa, 0, sizeof(MSSA_Opt));
}

int mssa_amd_set_options(MSSA_Opt *mssa, int q, int precision, unsigned long long seed)
{
    if (!mssa || !mssa->initialized || q < 0 || (precision != 0 && precision != 1)) return -1;
    ssa_ctx *c = (ssa_ctx *)mssa->dev;
    mssa->q = q;
    mssa->seed = seed;
    mssa->calls = 0;
    if (precision != mssa->precision) {
        ssa_engine_destroy(c->eng);
        c->eng = NULL;
        mssa->precision = precision;
        mssa->decomposed = false;
        if (mssa_build(mssa, c)) return -1;
    }
    return 0;
}

/* MKL/ssa_opt.h:4116-4210: kp = min(k + p, M L, K), p <= 0 -> 8 */
int mssa_opt_decompose(MSSA_Opt *mssa, int k, int oversampling)
{
    if (!mssa || !mssa->initialized || k < 1) return -1;
    ssa_ctx *c = (ssa_ctx *)mssa->dev;
    int ML = mssa->M * mssa->L;
    int p = oversampling <= 0 ? 8 : oversampling;
    int kp = imin(k + p, imin(ML, mssa->K));
    k = imin(k, kp);
    int b = imax(kp, 2);
    if (b > SSA_BMAX) return -1;
    if (ssa_engine_set_caps(c->eng, 0, b, k) || ssa_engine_prepare(c->eng, 1, 0)) return -1;
    if (ssa_engine_set_bk(c->eng, 0, b, k) || ssa_engine_set_rng(c->eng, 0, mssa->seed, mssa->calls)) return -1;
    int rc = ssa_engine_decompose(c->eng, mssa->q, 0);
    mssa->calls++;
    if (rc) return -1;
    if (!(mssa->U && mssa->V && mssa->sigma && mssa->eigenvalues && mssa->n_components == k)) {
        /* page-locked mirrors, kept while k is unchanged */
        ssa_engine_host_free(mssa->U); ssa_engine_host_free(mssa->V); free(mssa->sigma); free(mssa->eigenvalues);
        mssa->n_components = 0;
        mssa->U = ssa_engine_host_alloc((size_t)ML * k);
        mssa->V = ssa_engine_host_alloc((size_t)mssa->K * k);
        mssa->sigma = (double *)malloc(sizeof(double) * k);
        mssa->eigenvalues = (double *)malloc(sizeof(double) * k);
    }
    if (!mssa->U || !mssa->V || !mssa->sigma || !mssa->eigenvalues) return -1;
    double tv = 0;
    if (ssa_engine_results(c->eng, 0, mssa->sigma, mssa->eigenvalues, &tv, mssa->U, mssa->V, NULL) != k) return -1;
    mssa->total_variance = tv;
    mssa->n_components = k;
    mssa->decomposed = true;
    return 0;
}

/* MKL/ssa_opt.h:4212-4242 */
int mssa_opt_reconstruct(const MSSA_Opt *mssa, int series_idx, const int *group, int n_group, ssa_real *output)
{
    if (!mssa || !mssa->decomposed || !group || !output || n_group < 1 || series_idx < 0 || series_idx >= mssa->M)
        return -1;
    ssa_ctx *c = (ssa_ctx *)mssa->dev;
    return ssa_engine_reconstruct(c->eng, 0, series_idx, group, &n_group, 1, 1, 0, output);
}

/* MKL/ssa_opt.h:4244-4254 */
int mssa_opt_reconstruct_all(const MSSA_Opt *mssa, const int *group, int n_group, ssa_real *output)
{
    if (!mssa || !mssa->decomposed || !group || !output || n_group < 1) return -1;
    for (int m = 0; m < mssa->M; m++)
        if (mssa_opt_reconstruct(mssa, m, group, n_group, output + (size_t)m * mssa->N) != 0) return -1;
    return 0;
}

/* MKL/ssa_opt.h:4256-4270 */
int mssa_opt_series_contributions(const MSSA_Opt *mssa, ssa_real *contributions)
{
    if (!mssa || !mssa->decomposed || !contributions) return -1;
    int ML = mssa->M * mssa->L, k = mssa->n_components;
    for (int i = 0; i < k; i++)
        for (int m = 0; m < mssa->M; m++) {
            const double *u = mssa->U + (size_t)i * ML + (size_t)m * mssa->L;
            double s = 0;
            for (int t = 0; t < mssa->L; t++) s += u[t] * u[t];
            contributions[(size_t)m * k + i] = s;
        }
    return 0;
}

/* MKL/ssa_opt.h:4272-4283 */
ssa_real mssa_opt_variance_explained(const MSSA_Opt *mssa, int start, int end)
{
    if (!mssa || !mssa->decomposed || start < 0 || mssa->total_variance <= 0) return 0.0;
    if (end < 0 || end >= mssa->n_components) end = mssa->n_components - 1;
    double s = 0;
    for (int i = start; i <= end; i++) s += mssa->eigenvalues[i];
    return s / mssa->total_variance;
}

/* ------------------------------------------------------------------------ */
/* batched path                                                              */
/* ------------------------------------------------------------------------ */

/* A batch is G engines over disjoint sets of series, stepped as parallel
 * branches of one graph (G = 1: a single engine).  Series are dealt to the
 * branches round-robin in order of decreasing cost, so every branch holds a
 * similar mix of sizes; the Philox key of series i depends on i only. */
#define BATCH_MAX_BRANCHES 4

struct ssa_amd_batch {
    ssa_engine *eng[BATCH_MAX_BRANCHES];
    ssa_multi *multi;
    int Gmax, G, count, cap, q, p, h, precision;
    int wide_min;              /* ssa_amd_batch_set_wide_min, 0: the engines' default */
    unsigned long long seed;
    int *gid, *lid;            /* series i -> branch, index within that branch */
};

static int batch_branches(void)
{
    const char *v = getenv("SSA_AMD_BRANCHES");
    int G = (v && *v) ? atoi(v) : 2;
    if (G < 1) G = 1;
    if (G > BATCH_MAX_BRANCHES) G = BATCH_MAX_BRANCHES;
    return G;
}

static double *g_sort_cost;
static int by_cost_desc(const void *a, const void *b)
{
    const double ca = g_sort_cost[*(const int *)a], cb = g_sort_cost[*(const int *)b];
    if (ca != cb) return ca < cb ? 1 : -1;
    return *(const int *)a - *(const int *)b;
}

/* deal the series to the branches (round-robin in order of decreasing cost)
 * and (re)build each branch's engine; engines and their device buffers are
 * reused across calls */
static int batch_bind(ssa_amd_batch *b, int count, const int *N, const int *L, const int *k, const int *g)
{
    if (count < 1 || !N || !L || !k) return -1;
    const int p = b->p;
    int rc = -1;
    int *bm = (int *)malloc(sizeof(int) * count), *km = (int *)malloc(sizeof(int) * count);
    int *ord = (int *)malloc(sizeof(int) * count);
    double *cost = (double *)malloc(sizeof(double) * count);
    int *sN = (int *)malloc(sizeof(int) * count), *sL = (int *)malloc(sizeof(int) * count);
    int *sb = (int *)malloc(sizeof(int) * count), *sk = (int *)malloc(sizeof(int) * count);
    int *sg = (int *)malloc(sizeof(int) * count);
    if (!bm || !km || !ord || !cost || !sN || !sL || !sb || !sk || !sg) goto out;
    if (count > b->cap) {
        free(b->gid); free(b->lid);
        b->gid = (int *)malloc(sizeof(int) * count);
        b->lid = (int *)malloc(sizeof(int) * count);
        b->cap = (b->gid && b->lid) ? count : 0;
        if (!b->cap) goto out;
    }
    for (int i = 0; i < count; i++) {
        if (N[i] < 4 || L[i] < 2 || L[i] > N[i] - 1 || k[i] < 1) goto out;
        int r = imin(L[i], N[i] - L[i] + 1);
        int kp = imin(k[i] + p, r);
        bm[i] = imax(kp, 2);
        km[i] = imin(k[i], kp);
        if (bm[i] > SSA_BMAX) goto out;
        ord[i] = i;
        cost[i] = (double)N[i] * bm[i];
    }
    g_sort_cost = cost;
    qsort(ord, count, sizeof(int), by_cost_desc);
    b->count = count;
    b->G = imin(b->Gmax, count);
    for (int j = 0; j < count; j++) b->gid[ord[j]] = j % b->G;
    rc = 0;
    for (int br = 0; rc == 0 && br < b->G; br++) {
        /* Engine order of the branch's series: in decreasing cost, dealt in a
         * snake over N_XCD_BINS bins which are then concatenated.  Launches
         * give each XCD a contiguous range of work items, so every range holds
         * a similar mix of sizes, largest first (the launch tails are short
         * items). */
        enum { N_XCD_BINS = 8 };
        int nbr = 0;
        for (int j = 0; j < count; j++) if (b->gid[ord[j]] == br) nbr++;
        int n = 0;
        for (int bin = 0; bin < N_XCD_BINS; bin++) {
            int pos = 0;   /* position of the series in the branch's cost order */
            for (int j = 0; j < count; j++) {
                const int i = ord[j];
                if (b->gid[i] != br) continue;
                const int rnd = pos / N_XCD_BINS, slot = pos % N_XCD_BINS;
                const int bi = (rnd & 1) ? N_XCD_BINS - 1 - slot : slot;
                pos++;
                if (bi != bin) continue;
                b->lid[i] = n;
                sN[n] = N[i]; sL[n] = L[i]; sb[n] = bm[i]; sk[n] = km[i];
                sg[n] = g ? imin(imax(g[i], 1), km[i]) : 0;
                n++;
            }
        }
        if (n != nbr) { rc = -1; break; }
        if (!b->eng[br]) {
            int err = 0;
            b->eng[br] = ssa_engine_create(b->precision, n, sN, sL, NULL, sb, sk, b->q, &err);
            if (!b->eng[br]) { rc = -1; break; }
        } else if (ssa_engine_rebind(b->eng[br], n, sN, sL, NULL, sb, sk)) {
            rc = -1;
            break;
        }
        if (b->wide_min > 0) ssa_engine_set_wide_min(b->eng[br], b->wide_min);
        rc = ssa_engine_prepare(b->eng[br], 1, b->h);
        for (int i = 0; rc == 0 && i < count; i++) {
            if (b->gid[i] != br) continue;
            rc = ssa_engine_set_rng(b->eng[br], b->lid[i],
                                    b->seed ^ (0x9E3779B97F4A7C15ULL * (unsigned long long)(i + 1)), 0);
            if (rc == 0 && g) rc = ssa_engine_set_step(b->eng[br], b->lid[i], sg[b->lid[i]]);
        }
        if (rc == 0 && b->h > 0) rc = ssa_engine_set_step_h(b->eng[br], b->h);
    }
    ssa_multi_destroy(b->multi);
    b->multi = NULL;
    if (rc == 0 && b->G > 1) {
        b->multi = ssa_multi_create(b->eng, b->G);
        if (!b->multi) rc = -1;
    }
out:
    free(bm); free(km); free(ord); free(cost);
    free(sN); free(sL); free(sb); free(sk); free(sg);
    return rc;
}

ssa_amd_batch *ssa_amd_batch_create_g(int count, const int *N, const int *L, const int *k, const int *g, int h,
                                      int p, int q, int precision, unsigned long long seed, int branches)
{
    if (count < 1 || !N || !L || !k || q < 0 || h < 0 || (precision != 0 && precision != 1)) return NULL;
    if (branches < 0 || branches > BATCH_MAX_BRANCHES) return NULL;
    ssa_amd_batch *b = (ssa_amd_batch *)calloc(1, sizeof(ssa_amd_batch));
    if (!b) return NULL;
    b->p = p <= 0 ? 8 : p;
    b->q = q;
    b->h = h;
    b->precision = precision;
    b->seed = seed;
    b->Gmax = branches ? branches : batch_branches();
    if (batch_bind(b, count, N, L, k, g)) { ssa_amd_batch_free(b); return NULL; }
    return b;
}

ssa_amd_batch *ssa_amd_batch_create_ex(int count, const int *N, const int *L, const int *k, const int *g, int h,
                                       int p, int q, int precision, unsigned long long seed)
{
    return ssa_amd_batch_create_g(count, N, L, k, g, h, p, q, precision, seed, 0);
}

int ssa_amd_batch_rebind(ssa_amd_batch *b, int count, const int *N, const int *L, const int *k, const int *g)
{
    if (!b) return -1;
    if (b->multi) ssa_multi_sync(b->multi);
    for (int br = 0; br < b->G; br++) ssa_engine_sync(b->eng[br]);
    return batch_bind(b, count, N, L, k, g);
}

ssa_amd_batch *ssa_amd_batch_create(int count, const int *N, const int *L, const int *k, int p, int q,
                                    int precision, unsigned long long seed)
{
    return ssa_amd_batch_create_ex(count, N, L, k, NULL, 0, p, q, precision, seed);
}

int ssa_amd_batch_query(ssa_amd_batch *b)
{
    if (!b) return -1;
    return b->multi ? ssa_multi_query(b->multi) : ssa_engine_query(b->eng[0]);
}

int ssa_amd_batch_status(ssa_amd_batch *b, int i)
{
    if (!b || i < 0 || i >= b->count) return -1;
    return ssa_engine_status(b->eng[b->gid[i]], b->lid[i]);
}

int ssa_amd_batch_set_signals_f32(ssa_amd_batch *b, const float *cat, const long long *offsets)
{
    if (!b || !cat || !offsets) return -1;
    long long *off = (long long *)malloc(sizeof(long long) * b->count);
    if (!off) return -1;
    int rc = 0;
    for (int br = 0; rc == 0 && br < b->G; br++) {
        for (int i = 0; i < b->count; i++)
            if (b->gid[i] == br) off[b->lid[i]] = offsets[i];
        rc = ssa_engine_upload_signals_f32(b->eng[br], cat, off);
    }
    free(off);
    return rc;
}

int ssa_amd_autotune(int precision, double *ns8)
{
    if (precision != 0 && precision != 1) return -1;
    return ssa_engine_autotune(precision == 1 ? SSA_PREC_F32 : SSA_PREC_F64, ns8);
}

int ssa_amd_batch_set_wide_min(ssa_amd_batch *b, int n)
{
    if (!b || n < 1) return -1;
    b->wide_min = n;
    for (int br = 0; br < b->G; br++) ssa_engine_set_wide_min(b->eng[br], n);
    if (b->multi) {   /* a captured graph holds the old launches */
        ssa_multi_destroy(b->multi);
        b->multi = ssa_multi_create(b->eng, b->G);
        if (!b->multi) return -1;
    }
    return 0;
}

int ssa_amd_batch_set_key(ssa_amd_batch *b, int i, unsigned long long key)
{
    if (!b || i < 0 || i >= b->count) return -1;
    return ssa_engine_set_rng(b->eng[b->gid[i]], b->lid[i], key, 0);
}

int ssa_amd_batch_run(ssa_amd_batch *b, int use_graph)
{
    if (!b) return -1;
    return b->multi ? ssa_multi_step(b->multi, b->q, use_graph) : ssa_engine_step(b->eng[0], b->q, use_graph);
}

int ssa_amd_batch_sync(ssa_amd_batch *b)
{
    if (!b) return -1;
    return b->multi ? ssa_multi_sync(b->multi) : ssa_engine_sync(b->eng[0]);
}

int ssa_amd_batch_sigma(ssa_amd_batch *b, int i, ssa_real *sigma)
{
    if (!b || i < 0 || i >= b->count || !sigma) return -1;
    return ssa_engine_results(b->eng[b->gid[i]], b->lid[i], sigma, NULL, NULL, NULL, NULL, NULL);
}

int ssa_amd_batch_output(ssa_amd_batch *b, int i, ssa_real *out)
{
    if (!b || i < 0 || i >= b->count || !out) return -1;
    return ssa_engine_step_outputs(b->eng[b->gid[i]], b->lid[i], out);
}

int ssa_amd_batch_outputs(ssa_amd_batch *b, ssa_real *out, const long long *offsets, int *status)
{
    if (!b || !out || !offsets) return -1;
    long long *off = (long long *)malloc(sizeof(long long) * b->count);
    int *st = status ? (int *)malloc(sizeof(int) * b->count) : NULL;
    if (!off || (status && !st)) { free(off); free(st); return -1; }
    int rc = 0;
    for (int br = 0; rc == 0 && br < b->G; br++) {
        for (int i = 0; i < b->count; i++)
            if (b->gid[i] == br) off[b->lid[i]] = offsets[i];
        rc = ssa_engine_step_outputs_all(b->eng[br], out, off, st);
        if (rc == 0 && status)
            for (int i = 0; i < b->count; i++)
                if (b->gid[i] == br) status[i] = st[b->lid[i]];
    }
    free(off);
    free(st);
    return rc;
}

long long ssa_amd_batch_bytes(const ssa_amd_batch *b)
{
    if (!b) return 0;
    long long t = 0;
    for (int br = 0; br < b->G; br++) t += ssa_engine_bytes(b->eng[br]);
    return t;
}

/* per-launch timings of one eager step, branch after branch */
int ssa_amd_batch_profile(ssa_amd_batch *b, int max_rec, int *phase, int *lgn, float *ms)
{
    if (!b || !phase || !lgn || !ms) return -1;
    int n = 0;
    for (int br = 0; br < b->G; br++) {
        int r = ssa_engine_profile_step(b->eng[br], b->q, max_rec - n, phase + n, lgn + n, ms + n);
        if (r < 0) return -1;
        n += r;
    }
    return n;
}

/* No counterpart in the reference: device bytes one series (N, L, k) takes
 * in a batch of oversampling p, q power iterations and horizon h (the layout
 * the batch's engines use; a batch's ssa_amd_batch_bytes is the sum over its
 * series).  Host arithmetic only, for admission against a device budget. */
long long ssa_amd_series_bytes(int N, int L, int k, int p, int q, int precision, int h)
{
    if (N < 4 || L < 2 || L > N - 1 || k < 1 || q < 0 || h < 0) return -1;
    if (p <= 0) p = 8;
    const int r = imin(L, N - L + 1);
    const int kp = imin(k + p, r), bm = imax(kp, 2), km = imin(k, kp);
    if (bm > SSA_BMAX) return -1;
    return ssa_engine_series_bytes(precision ? SSA_PREC_F32 : SSA_PREC_F64, N, L, bm, km, 1, h, q);
}

int ssa_amd_batch_plan(const ssa_amd_batch *b, int i, int *out8)
{
    if (!b || i < 0 || i >= b->count || !out8) return -1;
    const ssa_desc *d = ssa_engine_desc(b->eng[b->gid[i]], b->lid[i]);
    out8[0] = d->n; out8[1] = d->CL; out8[2] = d->CK; out8[3] = d->AL; out8[4] = d->SK;
    out8[5] = d->b; out8[6] = d->k; out8[7] = d->lgn;
    return 0;
}

void ssa_amd_batch_free(ssa_amd_batch *b)
{
    if (!b) return;
    ssa_multi_destroy(b->multi);
    for (int br = 0; br < BATCH_MAX_BRANCHES; br++) ssa_engine_destroy(b->eng[br]);
    free(b->gid);
    free(b->lid);
    free(b);
}
