/*
 * ssa_engine.h -- host C engine that keeps a batch of SSA / MSSA series
 * resident in HBM and drives the HIP kernels of ssa_kernels.hip.
 *
 * One engine = one HIP stream + one bulk arena (signals, segment spectra,
 * panels, results; FP64 or FP32) + one FP64 arena (Gram partials and the
 * b x b factors).  A single SSA_Opt context is an engine with one series;
 * the batched and serving paths put many series in one engine.
 */
#ifndef SSA_ENGINE_H
#define SSA_ENGINE_H

#include "ssa_device.h"
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ssa_engine ssa_engine;

/* chunking chosen for a series: n, CL, CK (see ssa_device.h) */
typedef struct { int n, lgn, CL, CK, AL, SK; double cost; } ssa_plan;

/* planner: correlation length and chunking for (N, L, M) at precision prec,
 * for block width b, k components and q power iterations.  When lgn_force
 * is in [6, 13] only that length is considered. */
int ssa_plan_series(int prec, int N, int L, int M, int b, int k, int q, int lgn_force, ssa_plan *out);
/* device bytes of one series in an engine (signal, panel and small arenas) */
long long ssa_engine_series_bytes(int prec, int N, int L, int b, int k, int max_groups, int max_h, int q);

/* measure the device cost of one transform at every length 2^6..2^13 for
 * precision prec and plan every later engine of that precision with it;
 * ns_out (may be NULL) receives the 8 costs in ns */
int ssa_engine_autotune(int prec, double *ns_out);

/* create: count series with lengths N[i], windows L[i], channels M[i] (NULL -> 1) */
ssa_engine *ssa_engine_create(int prec, int count, const int *N, const int *L, const int *M,
                              const int *bmax, const int *kmax, int q_hint, int *err);
void ssa_engine_destroy(ssa_engine *e);
/* re-target an idle engine to a new set of series (as create); device
 * buffers are kept when large enough, so a serving loop allocates nothing
 * per wave once its engines have grown to the largest wave.  Call
 * ssa_engine_prepare (and the per-series setters) again afterwards. */
int ssa_engine_rebind(ssa_engine *e, int count, const int *N, const int *L, const int *M, const int *bmax,
                      const int *kmax);
int ssa_engine_count(const ssa_engine *e);
int ssa_engine_precision(const ssa_engine *e);
const ssa_desc *ssa_engine_desc(const ssa_engine *e, int s);

/* signals: x has M*N values (channel-major) */
int ssa_engine_upload_signal(ssa_engine *e, int s, const double *x);
/* all series at once from one float array, offsets[s] = first value of series s */
int ssa_engine_upload_signals_f32(ssa_engine *e, const float *cat, const long long *offsets);
int ssa_engine_spectra(ssa_engine *e);

/* capacities (block width, components) of series s; takes effect at the next prepare */
int ssa_engine_set_caps(ssa_engine *e, int s, int bmax, int kmax);
/* first ncomplex values of the spectrum of segment (0, 0) of series s */
int ssa_engine_read_spectrum(ssa_engine *e, int s, double *out, int ncomplex);

/* prepare panels and staging: groups = max reconstruction groups per call,
 * max_h = max forecast horizon; may be called again to enlarge */
int ssa_engine_prepare(ssa_engine *e, int max_groups, int max_h);
/* block width and components of the next decomposition of series s
 * (b <= bmax, k <= kmax of create); seed and call number of its test matrix */
int ssa_engine_set_bk(ssa_engine *e, int s, int b, int k);
int ssa_engine_set_rng(ssa_engine *e, int s, unsigned long long seed, unsigned long long call);
/* block-matvec launches of at least n items take the wide transform geometry
 * (default MV_WIDE_MIN_ITEMS or SSA_AMD_MV_WIDE_MIN; 1: every launch); drops a
 * captured step graph */
void ssa_engine_set_wide_min(ssa_engine *e, int n);
/* copy a K x b column-major test matrix into the Z panel; the next
 * decomposition of the engine uses injected matrices for every series */
int ssa_engine_inject_omega(ssa_engine *e, int s, const double *Omega);

/* decomposition of every series (q power iterations) */
int ssa_engine_decompose(ssa_engine *e, int q, int injected);
/* results of series s (any pointer may be NULL); returns k.  U and V, when
 * given, must come from ssa_engine_host_alloc: an FP64 engine copies into
 * them directly, a mixed one through its staging buffer (FP32 to FP64) */
int ssa_engine_results(ssa_engine *e, int s, double *sigma, double *lambda, double *totvar,
                       double *U, double *V, double *sweeps);
/* zeroed page-locked host memory for n doubles (NULL on failure), and its release */
double *ssa_engine_host_alloc(size_t n);
void ssa_engine_host_free(double *p);

/* reconstruction of n_groups groups of series s (channel ch for MSSA):
 * glist holds the concatenated groups, gcount their sizes; out receives
 * n_groups * N values (weight 0: diagonal averaging).  scale_sigma 1: sigma-weighted */
int ssa_engine_reconstruct(ssa_engine *e, int s, int ch, const int *glist, const int *gcount, int n_groups,
                           int scale_sigma, int weight, double *out);
/* as ssa_engine_reconstruct; use_cache 0 forms every component's transforms on the fly
 * even when the component spectra cache is valid (ssa_engine_reconstruct uses it then) */
int ssa_engine_reconstruct_ex(ssa_engine *e, int s, int ch, const int *glist, const int *gcount, int n_groups,
                              int scale_sigma, int weight, int use_cache, double *out);
/* W-correlation matrix (k x k row-major) of series s, from the component spectra
 * cache when use_cache is set and the cache is valid */
int ssa_engine_wcorr(ssa_engine *e, int s, int use_cache, double *W);
/* component spectra cache: the transforms of every chunk of every component's U
 * and V columns of every series, kept in HBM until the next decomposition (or
 * prepare, rebind, ssa_engine_cache_free); reconstructions read them instead of
 * transforming the factors again */
int ssa_engine_cache_build(ssa_engine *e);
void ssa_engine_cache_free(ssa_engine *e);
int ssa_engine_cache_valid(const ssa_engine *e);
/* reconstruction of group + LRF forecast of h steps.  out: N + h values
 * (full) or the h forecasts; returns -1 when the verticality is >= 1 */
/* R-forecast (vmode 0) or V-forecast (vmode 1) of group from the group's
 * reconstruction, or (base != NULL) from the last L-1 values of base */
int ssa_engine_forecast(ssa_engine *e, int s, const int *group, int n_group, int h, double *out, int full,
                        int vmode, const double *base, int base_len);

/* single-vector Hankel products of series s (dir 0: y = H v, 1: z = H^T u) */
int ssa_engine_matvec(ssa_engine *e, int s, int dir, const double *in, double *out);

/* batched step of every series: spectra, decomposition, reconstruction of
 * the group {0..k-1} into the series' output slot; captured into a HIP
 * graph on first use when use_graph is set */
int ssa_engine_step(ssa_engine *e, int q, int use_graph);
int ssa_engine_step_outputs(ssa_engine *e, int s, double *out);   /* N (+ step h) values */
/* all series: outputs at out + offsets[s], device status words (may be NULL), one sync */
int ssa_engine_step_outputs_all(ssa_engine *e, double *out, const long long *offsets, int *status);
/* group {0..gsize-1} reconstructed (and forecast) by the batched step for series s */
int ssa_engine_set_step(ssa_engine *e, int s, int gsize);
/* forecast horizon of the batched step (0: none; <= max_h of prepare) */
int ssa_engine_set_step_h(ssa_engine *e, int h);
/* 1 when all work queued on the engine stream has finished, 0 if not, -1 on error */
int ssa_engine_query(ssa_engine *e);
/* device status word of series s (2: forecast verticality >= 1) */
int ssa_engine_status(ssa_engine *e, int s);
/* one eager step with a HIP event pair around every kernel launch; returns
 * the number of records (phase id, FFT-length log2 or panel, milliseconds).
 * Phases: 0 spectra, 1 matvec H, 2 matvec H^T, 3 Gram, 4 Cholesky,
 * 5 apply R^-1, 6 small SVD, 7 U/V products, 8 reconstruction, 9 overlap-add,
 * 10 forecast */
int ssa_engine_profile_step(ssa_engine *e, int q, int max_rec, int *phase, int *lgn, float *ms);
int ssa_engine_sync(ssa_engine *e);

/* Several engines (disjoint sets of series) stepped as parallel branches of
 * one launch: forked from and joined into the first engine's stream, captured
 * as one HIP graph when use_graph.  The kernels of one branch fill the tails
 * of the other branches' launches. */
typedef struct ssa_multi ssa_multi;
ssa_multi *ssa_multi_create(ssa_engine *const *es, int G);
void ssa_multi_destroy(ssa_multi *m);
int ssa_multi_step(ssa_multi *m, int q, int use_graph);
int ssa_multi_sync(ssa_multi *m);
int ssa_multi_query(ssa_multi *m);
void *ssa_engine_stream(ssa_engine *e);
/* bytes of device memory held */
long long ssa_engine_bytes(const ssa_engine *e);

#ifdef __cplusplus
}
#endif
#endif
