/* mkl_baseline.c -- TEST INFRASTRUCTURE ONLY: the CPU baseline timed beside
 * the device path (bench.py's cpu_baseline leg) and its check (tests/).
 *
 * The reference's MKL call sequence for one series, restated in C against
 * MKL's public C API, which is resolved at run time from libmkl_rt (the MKL
 * runtime ships without its headers here, so the few entry points used are
 * declared below with their documented signatures):
 *   init            x zero-padded to n = next_pow2(N + K - 1), real-to-complex
 *                   DFTI transform (MKL/ssa_opt.h:1026-1200)
 *   block matvec    batches of columns: real-to-complex transform, product
 *                   with conj(X) ... as X conj(V), complex-to-real transform,
 *                   first L (or K) values (MKL/ssa_opt.h:910-1004)
 *   decompose       Gaussian test matrix, Y = H Omega, dgeqrf + dorgqr,
 *                   B = H^T Q, dgesdd('S'), U = Q Vb^T by dgemm, V = Ub,
 *                   sign rule sum(U[:, i]) >= 0 (MKL/ssa_opt.h:1685-1763);
 *                   q > 0 adds power iterations Z = H^T Q, Q_Z = qr(Z),
 *                   Y = H Q_Z, Q = qr(Y) with the same dgeqrf + dorgqr
 *   reconstruct     per component: transforms of sigma u and v, their product
 *                   accumulated, one inverse transform, times the diagonal
 *                   averaging weights (MKL/ssa_opt.h:2325-2400)
 *   wcorr           per component h = C2R(FFT(u) FFT(v)), row h[t] / sqrt(c_t)
 *                   normalised, W = G G^T by dsyrk, mirrored
 *                   (MKL/ssa_opt.h:2407-2490)
 * The test matrix comes from the caller (the same Philox matrix the oracle
 * and the device use), so the outputs can be checked against the oracle.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* DFTI configuration parameters and values (MKL's documented enumerations) */
enum {
    DFTI_BACKWARD_SCALE_ = 5,
    DFTI_NUMBER_OF_TRANSFORMS_ = 7,
    DFTI_CONJUGATE_EVEN_STORAGE_ = 10,
    DFTI_PLACEMENT_ = 11,
    DFTI_INPUT_DISTANCE_ = 14,
    DFTI_OUTPUT_DISTANCE_ = 15,
    DFTI_REAL_ = 33,
    DFTI_DOUBLE_ = 36,
    DFTI_COMPLEX_COMPLEX_ = 39,
    DFTI_NOT_INPLACE_ = 44,
};
enum {
    LAPACK_COL_MAJOR_ = 102,
    CBLAS_ROW_MAJOR_ = 101,
    CBLAS_COL_MAJOR_ = 102,
    CBLAS_NO_TRANS_ = 111,
    CBLAS_TRANS_ = 112,
    CBLAS_UPPER_ = 121,
};

typedef void *dfti_handle;
static long (*p_create)(dfti_handle *, int, int, long, ...);
static long (*p_set)(dfti_handle, int, ...);
static long (*p_commit)(dfti_handle);
static long (*p_fwd)(dfti_handle, void *, ...);
static long (*p_bwd)(dfti_handle, void *, ...);
static long (*p_free)(dfti_handle *);
static char *(*p_errmsg)(long);
static int (*p_geqrf)(int, int, int, double *, int, double *);
static int (*p_orgqr)(int, int, int, int, double *, int, const double *);
static int (*p_gesdd)(int, char, int, int, double *, int, double *, double *, int, double *, int);
static void (*p_dgemm)(int, int, int, int, int, int, double, const double *, int, const double *, int, double,
                       double *, int);
static void (*p_dsyrk)(int, int, int, int, int, double, const double *, int, double, double *, int);
static void (*p_set_threads)(int);
static void (*p_version)(char *, int);
static void *g_mkl;

#define SYM(var, name)                                          \
    do {                                                        \
        *(void **)(&var) = dlsym(g_mkl, name);                  \
        if (!var) {                                             \
            fprintf(stderr, "mkl_baseline: %s missing\n", name); \
            return -1;                                          \
        }                                                       \
    } while (0)

/* 0 when the MKL runtime at path (NULL: libmkl_rt.so.1 on the loader path)
 * is loaded and every entry point resolved */
int mklb_load(const char *path)
{
    if (g_mkl) return 0;
    /* RTLD_DEEPBIND: MKL's own libraries resolve their symbols among
     * themselves first.  A process that has already loaded a library exporting
     * BLAS/LAPACK names (PyTorch's CPU library does) would otherwise bind
     * MKL's internal calls to those copies. */
    g_mkl = dlopen(path ? path : "libmkl_rt.so.1", RTLD_NOW | RTLD_LOCAL | RTLD_DEEPBIND);
    if (!g_mkl) return -1;
    SYM(p_create, "DftiCreateDescriptor");
    SYM(p_set, "DftiSetValue");
    SYM(p_commit, "DftiCommitDescriptor");
    SYM(p_fwd, "DftiComputeForward");
    SYM(p_bwd, "DftiComputeBackward");
    SYM(p_free, "DftiFreeDescriptor");
    SYM(p_errmsg, "DftiErrorMessage");
    SYM(p_geqrf, "LAPACKE_dgeqrf");
    SYM(p_orgqr, "LAPACKE_dorgqr");
    SYM(p_gesdd, "LAPACKE_dgesdd");
    SYM(p_dgemm, "cblas_dgemm");
    SYM(p_dsyrk, "cblas_dsyrk");
    SYM(p_set_threads, "MKL_Set_Num_Threads");
    SYM(p_version, "MKL_Get_Version_String");
    return 0;
}

void mklb_set_threads(int n)
{
    if (p_set_threads) p_set_threads(n);
}

int mklb_version(char *buf, int len)
{
    if (!p_version || len < 2) return -1;
    p_version(buf, len);
    return 0;
}

typedef struct {
    int N, L, K, n, r2c, batch;
    double *X;                     /* spectrum of x: r2c complex values */
    double *wr, *wc;               /* batch workspaces: batch x n real, batch x r2c complex */
    double *w;                     /* diagonal averaging weights */
    dfti_handle f1, b1, fb, bb;    /* single and batched transforms */
} mklb_ctx;

static int dfti_ok(long st, const char *what)
{
    if (st == 0) return 1;
    fprintf(stderr, "mkl_baseline: %s: %s\n", what, p_errmsg ? p_errmsg(st) : "DFTI error");
    return 0;
}

static int make_desc(dfti_handle *h, int n, int count, int r2c, double bscale)
{
    if (!dfti_ok(p_create(h, DFTI_DOUBLE_, DFTI_REAL_, 1L, (long)n), "create")) return -1;
    if (!dfti_ok(p_set(*h, DFTI_PLACEMENT_, DFTI_NOT_INPLACE_), "placement")) return -1;
    if (!dfti_ok(p_set(*h, DFTI_CONJUGATE_EVEN_STORAGE_, DFTI_COMPLEX_COMPLEX_), "storage")) return -1;
    if (bscale != 0.0 && !dfti_ok(p_set(*h, DFTI_BACKWARD_SCALE_, bscale), "scale")) return -1;
    if (count > 1) {
        if (!dfti_ok(p_set(*h, DFTI_NUMBER_OF_TRANSFORMS_, (long)count), "count")) return -1;
        /* forward: real rows of n in, complex rows of r2c out; backward the reverse */
        const long din = bscale != 0.0 ? r2c : n, dout = bscale != 0.0 ? n : r2c;
        if (!dfti_ok(p_set(*h, DFTI_INPUT_DISTANCE_, din), "in distance")) return -1;
        if (!dfti_ok(p_set(*h, DFTI_OUTPUT_DISTANCE_, dout), "out distance")) return -1;
    }
    return dfti_ok(p_commit(*h), "commit") ? 0 : -1;
}

void mklb_free(mklb_ctx *c);

/* the reference's batch size: about 8 MB of real transform input, 4..64, a power of two */
static int batch_size(int n)
{
    int t = (8 * 1024 * 1024) / (n * (int)sizeof(double));
    if (t < 4) t = 4;
    if (t > 64) t = 64;
    int b = 1;
    while (b * 2 <= t) b *= 2;
    return b;
}

mklb_ctx *mklb_create(const double *x, int N, int L)
{
    if (!g_mkl || N < 4 || L < 2 || L > N - 1) return NULL;
    mklb_ctx *c = (mklb_ctx *)calloc(1, sizeof(mklb_ctx));
    if (!c) return NULL;
    c->N = N; c->L = L; c->K = N - L + 1;
    int n = 1;
    while (n < N + c->K - 1) n *= 2;
    c->n = n;
    c->r2c = n / 2 + 1;
    c->batch = batch_size(n);
    c->X = (double *)malloc(sizeof(double) * 2 * c->r2c);
    c->wr = (double *)calloc((size_t)c->batch * n, sizeof(double));
    c->wc = (double *)malloc(sizeof(double) * 2 * (size_t)c->batch * c->r2c);
    c->w = (double *)malloc(sizeof(double) * N);
    if (!c->X || !c->wr || !c->wc || !c->w) { mklb_free(c); return NULL; }
    if (make_desc(&c->f1, n, 1, c->r2c, 0.0) || make_desc(&c->b1, n, 1, c->r2c, 1.0 / n) ||
        make_desc(&c->fb, n, c->batch, c->r2c, 0.0) || make_desc(&c->bb, n, c->batch, c->r2c, 1.0 / n)) {
        mklb_free(c);
        return NULL;
    }
    memcpy(c->wr, x, sizeof(double) * N);
    memset(c->wr + N, 0, sizeof(double) * (n - N));
    p_fwd(c->f1, c->wr, c->X);
    for (int t = 0; t < N; t++) {
        int m = t + 1;
        if (L < m) m = L;
        if (c->K < m) m = c->K;
        if (N - t < m) m = N - t;
        c->w[t] = 1.0 / m;
    }
    return c;
}

void mklb_free(mklb_ctx *c)
{
    if (!c) return;
    if (c->f1) p_free(&c->f1);
    if (c->b1) p_free(&c->b1);
    if (c->fb) p_free(&c->fb);
    if (c->bb) p_free(&c->bb);
    free(c->X); free(c->wr); free(c->wc); free(c->w);
    free(c);
}

/* out[:, j] = first n_out values of IFFT(X conj(FFT(in[:, j]))), in columns of
 * n_in values (column-major, ld n_in / n_out): H v (n_in = K, n_out = L) and
 * H^T u (n_in = L, n_out = K) */
static void block_corr(mklb_ctx *c, const double *in, int n_in, double *out, int n_out, int b)
{
    const int n = c->n, r = c->r2c;
    for (int col = 0; col < b; col += c->batch) {
        const int cnt = b - col < c->batch ? b - col : c->batch;
        for (int i = 0; i < cnt; i++) {
            double *dst = c->wr + (size_t)i * n;
            memcpy(dst, in + (size_t)(col + i) * n_in, sizeof(double) * n_in);
            memset(dst + n_in, 0, sizeof(double) * (n - n_in));
        }
        if (cnt == c->batch) p_fwd(c->fb, c->wr, c->wc);
        else for (int i = 0; i < cnt; i++) p_fwd(c->f1, c->wr + (size_t)i * n, c->wc + (size_t)2 * i * r);
        for (int i = 0; i < cnt; i++) {
            double *z = c->wc + (size_t)2 * i * r;
            for (int k = 0; k < r; k++) {
                const double xr = c->X[2 * k], xi = c->X[2 * k + 1], vr = z[2 * k], vi = z[2 * k + 1];
                z[2 * k] = xr * vr + xi * vi;
                z[2 * k + 1] = xi * vr - xr * vi;
            }
        }
        if (cnt == c->batch) p_bwd(c->bb, c->wc, c->wr);
        else for (int i = 0; i < cnt; i++) p_bwd(c->b1, c->wc + (size_t)2 * i * r, c->wr + (size_t)i * n);
        for (int i = 0; i < cnt; i++) memcpy(out + (size_t)(col + i) * n_out, c->wr + (size_t)i * n, sizeof(double) * n_out);
    }
}

static int qr_q(double *A, int m, int b, double *tau)
{
    if (p_geqrf(LAPACK_COL_MAJOR_, m, b, A, m, tau) != 0) return -1;
    return p_orgqr(LAPACK_COL_MAJOR_, m, b, b, A, m, tau) != 0 ? -1 : 0;
}

/* Randomized decomposition of the series with the caller's K x b test matrix
 * omega (b = k + p, column-major), q power iterations.  Outputs sigma (k),
 * U (L x k) and V (K x k), column-major.  0 on success. */
int mklb_decompose(mklb_ctx *c, const double *omega, int k, int b, int q, double *sigma, double *U, double *V)
{
    const int L = c->L, K = c->K;
    if (k < 1 || b < k || b > L || b > K) return -1;
    double *Y = (double *)malloc(sizeof(double) * (size_t)L * b);
    double *Z = (double *)malloc(sizeof(double) * (size_t)K * b);
    double *tau = (double *)malloc(sizeof(double) * b);
    double *S = (double *)malloc(sizeof(double) * b);
    double *Ub = (double *)malloc(sizeof(double) * (size_t)K * b);
    double *Vt = (double *)malloc(sizeof(double) * (size_t)b * b);
    int rc = -1;
    if (!Y || !Z || !tau || !S || !Ub || !Vt) goto out;
    block_corr(c, omega, K, Y, L, b);
    if (qr_q(Y, L, b, tau)) goto out;
    for (int it = 0; it < q; it++) {
        block_corr(c, Y, L, Z, K, b);
        if (qr_q(Z, K, b, tau)) goto out;
        block_corr(c, Z, K, Y, L, b);
        if (qr_q(Y, L, b, tau)) goto out;
    }
    block_corr(c, Y, L, Z, K, b);   /* B = H^T Q, K x b */
    if (p_gesdd(LAPACK_COL_MAJOR_, 'S', K, b, Z, K, S, Ub, K, Vt, b) != 0) goto out;
    p_dgemm(CBLAS_COL_MAJOR_, CBLAS_NO_TRANS_, CBLAS_TRANS_, L, k, b, 1.0, Y, L, Vt, b, 0.0, U, L);
    memcpy(V, Ub, sizeof(double) * (size_t)K * k);
    for (int i = 0; i < k; i++) {
        sigma[i] = S[i];
        double s = 0.0;
        for (int t = 0; t < L; t++) s += U[(size_t)i * L + t];
        if (s < 0) {
            for (int t = 0; t < L; t++) U[(size_t)i * L + t] = -U[(size_t)i * L + t];
            for (int t = 0; t < K; t++) V[(size_t)i * K + t] = -V[(size_t)i * K + t];
        }
    }
    rc = 0;
out:
    free(Y); free(Z); free(tau); free(S); free(Ub); free(Vt);
    return rc;
}

/* out (N) = diagonal average of sum over the group of sigma_i u_i v_i^T, by
 * transforms of sigma_i u_i and v_i */
int mklb_reconstruct(mklb_ctx *c, const double *sigma, const double *U, const double *V, const int *group,
                     int n_group, double *out)
{
    const int n = c->n, r = c->r2c, L = c->L, K = c->K, N = c->N;
    if (n_group < 1 || c->batch < 2) return -1;
    double *acc = (double *)calloc((size_t)2 * r, sizeof(double));
    if (!acc) return -1;
    double *a = c->wr, *bv = c->wr + n, *fa = c->wc, *fb = c->wc + 2 * (size_t)r;
    for (int g = 0; g < n_group; g++) {
        const int i = group[g];
        for (int t = 0; t < L; t++) a[t] = sigma[i] * U[(size_t)i * L + t];
        memset(a + L, 0, sizeof(double) * (n - L));
        memcpy(bv, V + (size_t)i * K, sizeof(double) * K);
        memset(bv + K, 0, sizeof(double) * (n - K));
        p_fwd(c->f1, a, fa);
        p_fwd(c->f1, bv, fb);
        for (int k = 0; k < r; k++) {
            const double ar = fa[2 * k], ai = fa[2 * k + 1], br = fb[2 * k], bi = fb[2 * k + 1];
            acc[2 * k] += ar * br - ai * bi;
            acc[2 * k + 1] += ar * bi + ai * br;
        }
    }
    p_bwd(c->b1, acc, c->wr);
    for (int t = 0; t < N; t++) out[t] = c->wr[t] * c->w[t];
    free(acc);
    return 0;
}

/* W (k x k, row-major) = the W-correlation matrix of the k components: row i
 * of G is h_i[t] / sqrt(c_t) over its weighted norm, h_i = u_i (*) v_i (sigma
 * cancels), W = G G^T */
int mklb_wcorr(mklb_ctx *c, const double *U, const double *V, int k, double *W)
{
    const int n = c->n, r = c->r2c, L = c->L, K = c->K, N = c->N;
    if (k < 1) return -1;
    double *G = (double *)malloc(sizeof(double) * (size_t)k * N);
    if (!G) return -1;
    double *a = c->wr, *bv = c->wr + n, *fa = c->wc, *fb = c->wc + 2 * (size_t)r;
    for (int i = 0; i < k; i++) {
        memcpy(a, U + (size_t)i * L, sizeof(double) * L);
        memset(a + L, 0, sizeof(double) * (n - L));
        memcpy(bv, V + (size_t)i * K, sizeof(double) * K);
        memset(bv + K, 0, sizeof(double) * (n - K));
        p_fwd(c->f1, a, fa);
        p_fwd(c->f1, bv, fb);
        for (int j = 0; j < r; j++) {
            const double ar = fa[2 * j], ai = fa[2 * j + 1], br = fb[2 * j], bi = fb[2 * j + 1];
            fa[2 * j] = ar * br - ai * bi;
            fa[2 * j + 1] = ar * bi + ai * br;
        }
        p_bwd(c->b1, fa, a);
        double *g = G + (size_t)i * N, nn = 0.0;
        for (int t = 0; t < N; t++) {
            g[t] = a[t] * sqrt(c->w[t]);
            nn += g[t] * g[t];
        }
        const double inv = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
        for (int t = 0; t < N; t++) g[t] *= inv;
    }
    p_dsyrk(CBLAS_ROW_MAJOR_, CBLAS_UPPER_, CBLAS_NO_TRANS_, k, N, 1.0, G, N, 0.0, W, k);
    for (int i = 0; i < k; i++)
        for (int j = 0; j < i; j++) W[(size_t)i * k + j] = W[(size_t)j * k + i];
    free(G);
    return 0;
}
