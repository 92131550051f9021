/*
 * ssa_device.h -- the thin C ABI between the host C code (ssa_opt.c,
 * ssa_batch.c) and the HIP kernels (ssa_kernels.hip).  Plain C types only.
 *
 * A "series" is one SSA (M = 1) or MSSA (M > 1) problem resident in HBM.
 * Its Hankel operator H (rows: M stacked windows of length L; columns: K) is
 * applied chunk by chunk: rows in chunks of CL (AL chunks per channel),
 * columns in chunks of CK (SK chunks), so that every piece is a correlation
 * of length n >= CL + CK - 1 that fits one workgroup's LDS.  The spectra of
 * the signal segments x_m[a CL + s CK : ... + CL + CK - 1] are cached in HBM.
 */
#ifndef SSA_DEVICE_H
#define SSA_DEVICE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSA_BMAX 80           /* max block width k + p */
#define SSA_GRAM_ROWS 2048    /* panel rows per Gram / apply work item */
/* element multiple of the Y and Z panels' leading dimensions (128 bytes of FP32) */
#define SSA_PANEL_ALIGN 32
#define SSA_EDGE 64           /* samples at each end of a reconstruction summed directly (power of two) */
#define SSA_OLAP_TILE 256     /* reconstruction samples per overlap-add work item of a short launch */

enum { SSA_PREC_F64 = 0, SSA_PREC_F32 = 1 };

/* per-series descriptor (device resident, one per series of a batch) */
typedef struct ssa_desc {
    int N, L, K, M;           /* series length, window, K = N-L+1, channels */
    int ML;                   /* M*L rows of the stacked trajectory */
    int b, k;                 /* block width (k+p clamped), components */
    int n, lgn;               /* correlation length (power of two) */
    int CL, CK, AL, SK;       /* chunk sizes and counts (AL per channel) */
    int tilesY, tilesZ;       /* Gram row tiles of the ML x b and K x b panels */
    int ldY, ldZ;             /* leading dimensions of the Y and Z panels: ML and K rounded up to
                               * SSA_PANEL_ALIGN elements, so every column starts on a 128-byte line */
    int status;               /* set non-zero by device code on failure */
    int gsize;                /* batched step: reconstruct / forecast the group {0..gsize-1} */
    /* element offsets into the bulk arena (T = double or float) */
    long long off_x;          /* M*N signal values */
    long long off_spec;       /* M*AL*SK spectra, n complex each (2n T) */
    long long off_Y;          /* ML x b, ld ldY */
    long long off_Z;          /* K x b, ld ldZ */
    long long off_U;          /* ML x k, ld ML */
    long long off_V;          /* K x k, ld K */
    long long off_stage;      /* reconstruction staging */
    long long off_out;        /* reconstruction outputs */
    long long off_w;          /* N diagonal weights 1/min(t+1, L, K, N-t) */
    /* double offsets into the small (FP64) arena */
    long long off_part;       /* Gram partials: tiles x (b*b + b) */
    long long off_small;      /* small-matrix slots, see SSA_SLOT_* */
    unsigned long long seed;  /* Philox key of this series' test matrix */
    unsigned long long call;  /* decomposition counter (Philox counter word) */
    long long stage_cap;      /* staging capacity in T elements */
    long long out_cap;        /* output capacity in T elements */
    /* component spectra cache (element offset into the bulk arena): for
     * component g and chunk c (c < M*AL: row chunk c of U, else column chunk
     * c - M*AL of V), the n/2 + 1 complex values of T-(chunk) at
     * off_cache + 2 ((g (M AL + SK) + c) (n/2 + 1)) */
    long long off_cache;
    /* the batched step's output (N + h values) gathered here, every series'
     * slot after the previous one, so that one copy moves all of them */
    long long off_gather;
    /* staged block-matvec scratch (series chunked in both dimensions): the
     * input-chunk spectra of every column pair, 2 n T elements each */
    long long off_mvs;
} ssa_desc;

/* FP64 small-matrix slots, each SSA_BMAX*SSA_BMAX doubles, relative to off_small.
 * CholeskyQR pass 1 writes R1, R1INV; pass 2 writes R2, R2INV; the products
 * R = R2 R1 (R factor of the last CholeskyQR2 of Z, i.e. of B) and
 * RINV = R1INV R2INV (of the last CholeskyQR2 of Y, applied to form Q) are
 * formed only where they are used. */
enum {
    SSA_SLOT_R1 = 0, SSA_SLOT_R1INV, SSA_SLOT_R2, SSA_SLOT_R2INV, SSA_SLOT_R, SSA_SLOT_RINV,
    SSA_SLOT_W, SSA_SLOT_WK, SSA_SLOT_WKS, SSA_SLOT_MISC, SSA_NSLOTS
};
/* SSA_SLOT_MISC layout (doubles): [0, BMAX) colsum of raw Y, [BMAX, 2BMAX) colsum of Q,
 * [2BMAX, 3BMAX) sigma, [3BMAX, 4BMAX) lambda, 4BMAX: total variance,
 * 4BMAX+1: Jacobi sweeps used, 4BMAX+2: dead columns seen */
#define SSA_MISC_COLSUMY 0
#define SSA_MISC_COLSUMQ (SSA_BMAX)
#define SSA_MISC_SIGMA (2 * SSA_BMAX)
#define SSA_MISC_LAMBDA (3 * SSA_BMAX)
#define SSA_MISC_TOTVAR (4 * SSA_BMAX)
#define SSA_MISC_SWEEPS (4 * SSA_BMAX + 1)
#define SSA_MISC_DEAD (4 * SSA_BMAX + 2)

typedef struct { int s, x, y, z; } ssa_item;

/* panel selectors */
/* SSA_PANEL_OUT: the N x k weighted component panel at off_out (W-correlation) */
enum { SSA_PANEL_Y = 0, SSA_PANEL_Z = 1, SSA_PANEL_U = 2, SSA_PANEL_V = 3, SSA_PANEL_OUT = 4 };
/* column blocks of 16 of a panel of width c (Gram / apply instantiations) */
#define SSA_NB(c) (((c) + 15) / 16)

/* ---- launchers (return 0 or a hipError_t value) ---- */
/* twiddle tables of the precision on the current device (static device arrays,
 * filled on the first call per device and precision; no device allocation) */
int ssa_dev_twiddles(int prec, void **table_out, void *stream);
int ssa_dev_spectra(int prec, int lgn, const ssa_desc *D, const ssa_item *items, int n_items,
                    void *bulk, const void *tw, void *stream);
/* dir 0: out = H * in (in: K-dim panel, out: L-dim panel); dir 1: out = H^T * in.
 * wide 1: the wide transform geometry (k_matvec<T, lgn, true>: E = 32 values per
 * thread in FP32 at n = 8192, E = 16 in FP64 at 1024 <= n <= 4096, half the
 * waves); it pays on launches that fill the device.  Other lengths compile to
 * the same code either way. */
int ssa_dev_matvec(int prec, int lgn, int dir, int in_panel, int out_panel, int wide,
                   const ssa_desc *D, const ssa_item *items, int n_items,
                   void *bulk, const void *tw, void *stream);
/* staged block matvec for series chunked in both dimensions: items (s, pair,
 * input chunk) transform each input chunk once into the scratch at off_mvs;
 * then items (s, pair, output chunk) sum the stored spectra and transform back */
int ssa_dev_mv_spec(int prec, int lgn, int dir, int in_panel, int wide, const ssa_desc *D, const ssa_item *items,
                    int n_items, void *bulk, const void *tw, void *stream);
int ssa_dev_mv_acc(int prec, int lgn, int dir, int out_panel, int wide, const ssa_desc *D, const ssa_item *items,
                   int n_items, void *bulk, const void *tw, void *stream);
/* Philox Gaussian test matrix into panel Z; items (s, row tile) of the K-dim panel.
 * call >= 0 replaces every series' descriptor counter (call) for this launch */
int ssa_dev_omega(int prec, const ssa_desc *D, const ssa_item *items, int n_items, void *bulk, long long call,
                  void *stream);
/* Gram / apply: every panel of the launch has at most nb*16 columns (1 <= nb <= 5) */
int ssa_dev_gram(int prec, int panel, int use_m, int nb, const ssa_desc *D, const ssa_item *items, int n_items,
                 void *bulk, double *small, void *stream);
/* Cholesky of the summed Gram; a pivot at or below tol times its column's
 * original diagonal marks the column dependent (zero row of R, zero column
 * of R^-1) */
int ssa_dev_chol(int panel, int pass, int bmax, double tol, const ssa_desc *D, const ssa_item *items, int n_items,
                 double *small, void *stream);
int ssa_dev_apply(int prec, int src, int dst, int slot, int ncols_k, int nb, const ssa_desc *D,
                  const ssa_item *items, int n_items, void *bulk, double *small, void *stream);
/* C = A B of the upper-triangular b x b slot matrices (FP64 MFMA), one workgroup per series */
int ssa_dev_trmm(int nb, int slot_a, int slot_b, int slot_c, const ssa_desc *D, const ssa_item *items, int n_items,
                 double *small, void *stream);
/* SVD of the R factor of B in slot r_slot; colsum(Q) = colsum(Y) * (slot rinv_slot), the R^-1 of the final Q.
 * Every series of the launch has b <= bmax; instantiated per column-block count SSA_NB(bmax). */
int ssa_dev_svd(int bmax, int r_slot, int rinv_slot, double early, const ssa_desc *D, const ssa_item *items,
                int n_items, double *small, void *stream);
/* reconstruction: items (s, gid * S + split, row chunk o, col chunk sk); groups: gdesc[gid] =
 * {start, count}, glist = component indices; scale_sigma 1: sum_g sigma_g u_g (*) v_g, 0: u (*) v.
 * S = min(nsplit, staging slots of the series): split j sums its share of the group's components
 * into its own staged piece; ssa_dev_olap with the same nsplit sums the pieces. */
int ssa_dev_recon(int prec, int lgn, int scale_sigma, int limit_gsize, int nsplit, const ssa_desc *D,
                  const ssa_item *items, int n_items, const int *gdesc, const int *glist, void *bulk,
                  const double *small, const void *tw, void *stream);
/* overlap-add of the staging pieces: items (s, gid, t tile of `tile` samples, a multiple
 * of 256); weight 0: w_t, 1: sqrt(w_t); channel index taken from gdesc[gid].z.  Output of
 * gid goes to out + gid*N. */
int ssa_dev_olap(int prec, int weight, int nsplit, int tile, const ssa_desc *D, const ssa_item *items,
                 int n_items, const int *gdesc, void *bulk, void *stream);
/* component spectra cache: items (s, component pair p, chunk c); components 2p and
 * 2p + 1 (< k) of U's row chunk c (c < M AL) or V's column chunk c - M AL, transformed
 * together and separated; the n/2 + 1 values of each land at off_cache (ssa_desc) */
int ssa_dev_cache_build(int prec, int lgn, const ssa_desc *D, const ssa_item *items, int n_items, void *bulk,
                        const void *tw, void *stream);
/* reconstruction pieces from the cache: pieces (gid, a, sk) numbered (gid AL + a) SK + sk
 * for n_pieces = groups AL SK; item (s, first piece) takes that piece and the next
 * (two real pieces share one inverse transform); staged as ssa_dev_recon stages
 * them with one split, for ssa_dev_olap with nsplit 1 */
int ssa_dev_recon_cached(int prec, int lgn, int scale_sigma, const ssa_desc *D, const ssa_item *items, int n_items,
                         int n_pieces, const int *gdesc, const int *glist, void *bulk, const double *small,
                         const void *tw, void *stream);
/* misc_n FP64 values from misc, then un panel elements from U and vn from V,
 * packed in that order into out (8 misc_n + (un + vn) element bytes; out may
 * be the device address of page-locked host memory) */
int ssa_dev_pack_results(int prec, const double *misc, int misc_n, const void *U, long long un, const void *V,
                         long long vn, void *out, void *stream);
/* copy every series' step output (N + h values at off_out) to its off_gather slot; items (s) */
int ssa_dev_gather(int prec, int h, const ssa_desc *D, const ssa_item *items, int n_items, void *bulk, void *stream);
/* the first and last SSA_EDGE samples of every group's reconstruction by direct
 * anti-diagonal sums (FP64), overwriting the overlap-add's values there; items
 * (s, gid); scale_sigma, weight, limit_gsize and groups as in ssa_dev_recon / ssa_dev_olap */
int ssa_dev_edge(int prec, int scale_sigma, int weight, int limit_gsize, const ssa_desc *D, const ssa_item *items,
                 int n_items, const int *gdesc, const int *glist, void *bulk, const double *small, void *stream);
/* W-correlation: Gram of the sqrt(w)-weighted component panel (stored at off_out, N x k)
 * then normalisation; W written row-major k x k to small + off_small + SLOT_W */
int ssa_dev_wcorr(int prec, int nb, const ssa_desc *D, const ssa_item *tile_items, int n_tiles,
                  const ssa_item *series_items, int n_series, void *bulk, double *small, void *stream);
/* LRF + forecast (vmode 0) or V-forecast (vmode 1): per series, group gdesc[0]; reads the
 * last L-1 values before off_out + N, writes h forecasts after them (off_out + N) */
int ssa_dev_forecast(int prec, int h, int limit_gsize, int vmode, const ssa_desc *D, const ssa_item *items,
                     int n_items, const int *gdesc, const int *glist, void *bulk, double *small, void *stream);

#ifdef __cplusplus
}
#endif
#endif
