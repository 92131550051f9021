/*
 * ssa_oracle.h -- FP64 CPU restatement of SSA-Opt's decomposition hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline).  The product library (libssa.so) never links,
 * loads or calls anything here.
 *
 * It restates, in plain C with its own FFT / QR / Jacobi SVD:
 *   - the Hankel matvec and adjoint      MKL/ssa_opt.h:871-903 (conj-multiply correlation)
 *   - init (FFT length, diagonal weights) MKL/ssa_opt.h:1026-1284
 *   - decompose_randomized               MKL/ssa_opt.h:1685-1763, extended with q
 *                                        subspace (power) iterations, Halko et al. 2011
 *                                        Alg. 4.4, which the reference does not have
 *   - reconstruct                        MKL/ssa_opt.h:2325-2400
 *   - wcorr_matrix                       MKL/ssa_opt.h:2407-2477
 *   - compute_lrf / forecast(_full)      MKL/ssa_opt.h:2766-2885
 *   - MSSA init/decompose/reconstruct    MKL/ssa_opt.h:3907-4283
 *
 * Unlike the reference, the Gaussian test matrix is never drawn from an
 * internal stream: it is passed in (Omega, K x b column-major), so the same
 * matrix can be fed to this oracle, to the reference's own dependency-free
 * build (oracle/_ref) and to the device path.  orc_philox_gaussian() draws it
 * from a counter-based Philox4x32-10 generator (Salmon et al., SC'11) with
 * Box-Muller, the same definition the device path uses.
 *
 * Parity pins (tests/test_oracle_*.py): brute-force double sums for the
 * matvecs, numpy FFT, dense LAPACK SVD of the explicit Hankel matrix, explicit
 * anti-diagonal averaging, the reference's ssa_opt_ref.h compiled from source
 * (q = 0, same Omega), the sinusoid recurrence x[n] = 2cos(w)x[n-1] - x[n-2]
 * for the LRF, and the Philox4x32-10 known-answer vectors of Random123.
 */
#ifndef SSA_ORACLE_H
#define SSA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int N, L, K;
    int n;              /* FFT length: next_pow2(N + K - 1), MKL/ssa_opt.h:1047-1049 */
    double *X;          /* forward FFT of zero-padded x, n complex (interleaved) */
    double *w;          /* inv_diag_count[t] = 1/min(t+1, L, K, N-t), MKL/ssa_opt.h:1270-1279 */
    double *x;          /* copy of the signal */
    int k, b;           /* components kept, block width k+p (clamped) */
    double *U;          /* L x k column-major */
    double *V;          /* K x k column-major */
    double *sigma, *lambda;
    double total_variance;
    int decomposed;
} orc_ssa;

/* ---- building blocks (exported so tests can pin each one) ---- */
void orc_fft(double *z, int n, int sign);                 /* in place, interleaved, unscaled */
int  orc_next_pow2(int v);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_philox_gaussian(uint64_t seed, uint64_t call, int rows, int cols, double *out);
int  orc_qr_q(double *A, int m, int n, double *R);        /* Householder; A <- Q (m x n), R n x n (may be NULL) */
int  orc_jacobi_svd(double *A, int m, int n, double *S, double *Vt_colmajor);  /* one-sided */

/* ---- SSA ---- */
int  orc_init(orc_ssa *s, const double *x, int N, int L);
void orc_free(orc_ssa *s);
int  orc_update_signal(orc_ssa *s, const double *x);
void orc_matvec(const orc_ssa *s, const double *v, double *y);     /* y = H v,  v: K, y: L */
void orc_matvec_T(const orc_ssa *s, const double *u, double *z);   /* z = H^T u, u: L, z: K */
void orc_matvec_block(const orc_ssa *s, const double *V, double *Y, int b, int transpose);
int  orc_decompose(orc_ssa *s, int k, int p, int q, const double *Omega);
int  orc_reconstruct(const orc_ssa *s, const int *group, int n_group, double *out);
int  orc_wcorr(const orc_ssa *s, double *W);
int  orc_lrf(const orc_ssa *s, const int *group, int n_group, double *R, double *nu2);
int  orc_forecast(const orc_ssa *s, const int *group, int n_group, int h, double *out);
int  orc_forecast_full(const orc_ssa *s, const int *group, int n_group, int h, double *out);
int  orc_vforecast(const orc_ssa *s, const int *group, int n_group, const double *base, int base_len, int h,
                   double *out);
double orc_variance_explained(const orc_ssa *s, int start, int end);
/* Cadzow / gap filling (MKL/ssa_opt.h:3147-3878); out3 = iterations, final_diff, converged;
 * out4 adds n_gaps */
int  orc_cadzow(const double *x, int N, int L, int rank, int max_iter, double tol, double *output, double *out3);
int  orc_cadzow_inplace(orc_ssa *s, int rank, int max_iter, double tol, uint64_t counter0, double *out3);
int  orc_gapfill(double *x, int N, int L, int rank, int max_iter, double tol, double *out4);
int  orc_gapfill_simple(double *x, int N, int L, int rank, double *out4);

/* ---- accessors for ctypes ---- */
int  orc_dims(const orc_ssa *s, int *out4);   /* N, L, K, n */
int  orc_get(const orc_ssa *s, double *U, double *V, double *sigma); /* copies; returns k */
orc_ssa *orc_new(void);
void orc_delete(orc_ssa *s);

/* ---- MSSA ---- */
typedef struct {
    int M, N, L, K, n;
    double *X;          /* M spectra, n complex each */
    double *w;
    int k, b;
    double *U;          /* (M L) x k */
    double *V;          /* K x k */
    double *sigma, *lambda;
    double total_variance;
    int decomposed;
} orc_mssa;

orc_mssa *orc_mssa_new(void);
void orc_mssa_delete(orc_mssa *m);
int  orc_mssa_init(orc_mssa *m, const double *X, int M, int N, int L);
void orc_mssa_free(orc_mssa *m);
void orc_mssa_matvec(const orc_mssa *m, const double *v, double *y);    /* y: M*L */
void orc_mssa_matvec_T(const orc_mssa *m, const double *u, double *z);  /* z: K */
int  orc_mssa_decompose(orc_mssa *m, int k, int p, int q, const double *Omega);
int  orc_mssa_reconstruct(const orc_mssa *m, int series, const int *group, int n_group, double *out);
int  orc_mssa_get(const orc_mssa *m, double *U, double *V, double *sigma);
int  orc_mssa_contributions(const orc_mssa *m, double *c);

#ifdef __cplusplus
}
#endif
#endif
