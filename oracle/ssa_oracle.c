/*
 * ssa_oracle.c -- FP64 CPU restatement of SSA-Opt's decomposition hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see ssa_oracle.h).  Every function cites the
 * reference file:line whose behaviour it restates.
 */
#include "ssa_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_TWO_PI_L 6.283185307179586476925286766559005768L

/* ------------------------------------------------------------------------ */
/* small helpers                                                            */
/* ------------------------------------------------------------------------ */

int orc_next_pow2(int v)   /* MKL/ssa_opt.h:434-440 */
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

static int imin(int a, int b) { return a < b ? a : b; }

static void *xcalloc(size_t n, size_t sz)
{
    if (n == 0) n = 1;
    return calloc(n, sz);
}

/* ------------------------------------------------------------------------ */
/* FFT: iterative radix-2 Cooley-Tukey on interleaved complex data.          */
/* Twiddles exp(-2 pi i j / n) are evaluated in long double and rounded once, */
/* so every table entry is within ~0.5 ulp of the exact root of unity.      */
/* Stands in for MKL DFTI (MKL/ssa_opt.h:1119-1196).                        */
/* ------------------------------------------------------------------------ */

#define ORC_MAX_LOG2 22
static double *g_tw[ORC_MAX_LOG2 + 1];   /* tw[j] = exp(-2 pi i j/n), j < n/2 */

static const double *twiddles(int n, int lg)
{
    if (!g_tw[lg]) {
        int h = n / 2 > 0 ? n / 2 : 1;
        double *t = (double *)malloc(sizeof(double) * 2 * (size_t)h);
        for (int j = 0; j < h; j++) {
            long double a = ORC_TWO_PI_L * (long double)j / (long double)n;
            t[2 * j] = (double)cosl(a);
            t[2 * j + 1] = (double)(-sinl(a));
        }
        g_tw[lg] = t;
    }
    return g_tw[lg];
}

void orc_fft(double *z, int n, int sign)
{
    int lg = 0;
    while ((1 << lg) < n) lg++;
    if (n < 2) return;
    const double *tw = twiddles(n, lg);
    /* bit reversal */
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = z[2 * i], ti = z[2 * i + 1];
            z[2 * i] = z[2 * j]; z[2 * i + 1] = z[2 * j + 1];
            z[2 * j] = tr; z[2 * j + 1] = ti;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int j = 0; j < half; j++) {
                double wr = tw[2 * j * step], wi = tw[2 * j * step + 1];
                if (sign > 0) wi = -wi;
                double *a = z + 2 * (i + j), *b = z + 2 * (i + j + half);
                double tr = wr * b[0] - wi * b[1];
                double ti = wr * b[1] + wi * b[0];
                b[0] = a[0] - tr; b[1] = a[1] - ti;
                a[0] += tr; a[1] += ti;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as  */
/* easy as 1, 2, 3", SC'11) and Box-Muller.  Replaces the reference's VSL   */
/* MT2203 stream (MKL/ssa_opt.h:1245, :1719), which cannot be reproduced    */
/* without MKL; the test matrix is therefore an explicit input everywhere.  */
/* ------------------------------------------------------------------------ */

void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* out[c*rows + r] ~ N(0,1), column-major rows x cols.  Row r of the column
 * quad (4j .. 4j+3) comes from one Philox call with counter (r, j, call_lo,
 * call_hi) under key (seed_lo, seed_hi); its four 32-bit outputs o give two
 * Box-Muller pairs: u1 = o[2h] 2^-32, u2 = o[2h+1] 2^-32,
 * rad = sqrt(-2 ln(1 - u1)), rad*cos(2 pi u2) -> column 4j+2h,
 * rad*sin(2 pi u2) -> column 4j+2h+1.  The device path draws the same values. */
void orc_philox_gaussian(uint64_t seed, uint64_t call, int rows, int cols, double *out)
{
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) };
    for (int j = 0; 4 * j < cols; j++) {
        for (int r = 0; r < rows; r++) {
            uint32_t ctr[4] = { (uint32_t)r, (uint32_t)j, (uint32_t)call, (uint32_t)(call >> 32) };
            uint32_t o[4];
            orc_philox4x32_10(ctr, key, o);
            for (int h = 0; h < 2; h++) {
                double u1 = (double)o[2 * h] * 0x1p-32, u2 = (double)o[2 * h + 1] * 0x1p-32;
                double rad = sqrt(-2.0 * log(1.0 - u1));
                double ang = 6.283185307179586476925286766559 * u2;
                int c = 4 * j + 2 * h;
                if (c < cols) out[(size_t)c * rows + r] = rad * cos(ang);
                if (c + 1 < cols) out[(size_t)(c + 1) * rows + r] = rad * sin(ang);
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Dense FP64 kernels standing in for LAPACKE_dgeqrf/dorgqr and dgesdd      */
/* (MKL/ssa_opt.h:1723-1731).                                               */
/* ------------------------------------------------------------------------ */

/* Householder QR of A (m x n, col-major, m >= n).  On return A holds the
 * explicit thin Q; R (n x n col-major, upper) if non-NULL. */
int orc_qr_q(double *A, int m, int n, double *R)
{
    if (m < n) return -1;
    double *Hv = (double *)xcalloc((size_t)m * n, sizeof(double));  /* reflectors */
    double *tau = (double *)xcalloc(n, sizeof(double));
    if (!Hv || !tau) { free(Hv); free(tau); return -1; }
    for (int j = 0; j < n; j++) {
        double *a = A + (size_t)j * m;
        double nrm = 0.0;
        for (int i = j; i < m; i++) nrm += a[i] * a[i];
        nrm = sqrt(nrm);
        double alpha = a[j];
        double beta = (alpha >= 0.0) ? -nrm : nrm;
        double *v = Hv + (size_t)j * m;
        if (nrm == 0.0) { tau[j] = 0.0; continue; }
        double v0 = alpha - beta;
        v[j] = 1.0;
        for (int i = j + 1; i < m; i++) v[i] = a[i] / v0;
        tau[j] = (beta - alpha) / beta;
        /* apply (I - tau v v^T) to A[:, j:] */
        for (int c = j; c < n; c++) {
            double *ac = A + (size_t)c * m;
            double s = 0.0;
            for (int i = j; i < m; i++) s += v[i] * ac[i];
            s *= tau[j];
            for (int i = j; i < m; i++) ac[i] -= s * v[i];
        }
    }
    if (R) {
        for (int c = 0; c < n; c++)
            for (int r = 0; r < n; r++)
                R[(size_t)c * n + r] = (r <= c) ? A[(size_t)c * m + r] : 0.0;
    }
    /* form Q = H_0 H_1 ... H_{n-1} [I; 0] backwards */
    for (int c = 0; c < n; c++)
        for (int i = 0; i < m; i++) A[(size_t)c * m + i] = (i == c) ? 1.0 : 0.0;
    for (int j = n - 1; j >= 0; j--) {
        if (tau[j] == 0.0) continue;
        double *v = Hv + (size_t)j * m;
        for (int c = j; c < n; c++) {
            double *ac = A + (size_t)c * m;
            double s = 0.0;
            for (int i = j; i < m; i++) s += v[i] * ac[i];
            s *= tau[j];
            for (int i = j; i < m; i++) ac[i] -= s * v[i];
        }
    }
    free(Hv); free(tau);
    return 0;
}

/* One-sided (Hestenes) Jacobi SVD of A (m x n, col-major, m >= n).
 * On return A holds U*diag(S) rescaled to U (unit columns), S descending,
 * W (n x n col-major) the right singular vectors: A_in = U diag(S) W^T. */
int orc_jacobi_svd(double *A, int m, int n, double *S, double *W)
{
    for (int c = 0; c < n; c++)
        for (int r = 0; r < n; r++) W[(size_t)c * n + r] = (r == c) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 80; sweep++) {
        int rotated = 0;
        for (int p = 0; p < n - 1; p++) {
            for (int q = p + 1; q < n; q++) {
                double *ap = A + (size_t)p * m, *aq = A + (size_t)q * m;
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int i = 0; i < m; i++) { al += ap[i] * ap[i]; be += aq[i] * aq[i]; ga += ap[i] * aq[i]; }
                if (ga == 0.0 || fabs(ga) <= 1e-15 * sqrt(al * be)) continue;
                rotated = 1;
                double zeta = (be - al) / (2.0 * ga);
                double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
                double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
                for (int i = 0; i < m; i++) {
                    double x = ap[i], y = aq[i];
                    ap[i] = c * x - s * y; aq[i] = s * x + c * y;
                }
                double *wp = W + (size_t)p * n, *wq = W + (size_t)q * n;
                for (int i = 0; i < n; i++) {
                    double x = wp[i], y = wq[i];
                    wp[i] = c * x - s * y; wq[i] = s * x + c * y;
                }
            }
        }
        if (!rotated) break;
    }
    for (int c = 0; c < n; c++) {
        double s2 = 0.0;
        double *ac = A + (size_t)c * m;
        for (int i = 0; i < m; i++) s2 += ac[i] * ac[i];
        S[c] = sqrt(s2);
        if (S[c] > 0.0)
            for (int i = 0; i < m; i++) ac[i] /= S[c];
    }
    /* selection sort, descending */
    double *tmp = (double *)xcalloc(m > n ? m : n, sizeof(double));
    for (int i = 0; i < n - 1; i++) {
        int best = i;
        for (int j = i + 1; j < n; j++) if (S[j] > S[best]) best = j;
        if (best != i) {
            double t = S[i]; S[i] = S[best]; S[best] = t;
            memcpy(tmp, A + (size_t)i * m, sizeof(double) * m);
            memcpy(A + (size_t)i * m, A + (size_t)best * m, sizeof(double) * m);
            memcpy(A + (size_t)best * m, tmp, sizeof(double) * m);
            memcpy(tmp, W + (size_t)i * n, sizeof(double) * n);
            memcpy(W + (size_t)i * n, W + (size_t)best * n, sizeof(double) * n);
            memcpy(W + (size_t)best * n, tmp, sizeof(double) * n);
        }
    }
    free(tmp);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* SSA context                                                              */
/* ------------------------------------------------------------------------ */

static void spectrum(const double *x, int N, int n, double *X)
{
    memset(X, 0, sizeof(double) * 2 * (size_t)n);
    for (int t = 0; t < N; t++) X[2 * t] = x[t];
    orc_fft(X, n, -1);
}

/* MKL/ssa_opt.h:1026-1284 (validation, dimensions, FFT(x), weights) */
int orc_init(orc_ssa *s, const double *x, int N, int L)
{
    if (!s || !x || N < 4 || L < 2 || L > N - 1) return -1;
    memset(s, 0, sizeof(*s));
    s->N = N; s->L = L; s->K = N - L + 1;
    s->n = orc_next_pow2(N + s->K - 1);
    s->X = (double *)xcalloc(2 * (size_t)s->n, sizeof(double));
    s->w = (double *)xcalloc(N, sizeof(double));
    s->x = (double *)xcalloc(N, sizeof(double));
    if (!s->X || !s->w || !s->x) { orc_free(s); return -1; }
    memcpy(s->x, x, sizeof(double) * N);
    spectrum(x, N, s->n, s->X);
    for (int t = 0; t < N; t++) {
        int c = imin(imin(t + 1, L), imin(s->K, N - t));
        s->w[t] = c > 0 ? 1.0 / c : 0.0;
    }
    return 0;
}

void orc_free(orc_ssa *s)   /* MKL/ssa_opt.h:1286-1350 */
{
    if (!s) return;
    free(s->X); free(s->w); free(s->x);
    free(s->U); free(s->V); free(s->sigma); free(s->lambda);
    memset(s, 0, sizeof(*s));
}

orc_ssa *orc_new(void) { return (orc_ssa *)calloc(1, sizeof(orc_ssa)); }
void orc_delete(orc_ssa *s) { if (s) { orc_free(s); free(s); } }

int orc_update_signal(orc_ssa *s, const double *x)   /* MKL/ssa_opt.h:1545-1558 */
{
    if (!s || !s->X || !x) return -1;
    memcpy(s->x, x, sizeof(double) * s->N);
    spectrum(x, s->N, s->n, s->X);
    s->decomposed = 0;
    return 0;
}

/* Correlation of x with two real vectors packed as a + i b:
 * FFT(corr)[k] = X[k] * Z[(n-k) mod n] with Z = FFT(a + i b)
 * (the conj-multiply of MKL/ssa_opt.h:871-903, two vectors per transform).
 * Writes Re -> ya[0:len_out), Im -> yb[0:len_out) (yb may be NULL). */
static void corr2(const double *X, int n, const double *a, const double *b, int len_in,
                  double *ya, double *yb, int len_out, double *ws)
{
    memset(ws, 0, sizeof(double) * 2 * (size_t)n);
    for (int i = 0; i < len_in; i++) { ws[2 * i] = a[i]; ws[2 * i + 1] = b ? b[i] : 0.0; }
    orc_fft(ws, n, -1);
    double *P = ws + 2 * (size_t)n;
    for (int kk = 0; kk < n; kk++) {
        int m = (n - kk) & (n - 1);
        double xr = X[2 * kk], xi = X[2 * kk + 1], zr = ws[2 * m], zi = ws[2 * m + 1];
        P[2 * kk] = xr * zr - xi * zi;
        P[2 * kk + 1] = xr * zi + xi * zr;
    }
    orc_fft(P, n, +1);
    double sc = 1.0 / n;
    for (int i = 0; i < len_out; i++) {
        ya[i] = P[2 * i] * sc;
        if (yb) yb[i] = P[2 * i + 1] * sc;
    }
}

void orc_matvec(const orc_ssa *s, const double *v, double *y)      /* MKL/ssa_opt.h:871-883 */
{
    double *ws = (double *)xcalloc(4 * (size_t)s->n, sizeof(double));
    corr2(s->X, s->n, v, NULL, s->K, y, NULL, s->L, ws);
    free(ws);
}

void orc_matvec_T(const orc_ssa *s, const double *u, double *z)    /* MKL/ssa_opt.h:891-903 */
{
    double *ws = (double *)xcalloc(4 * (size_t)s->n, sizeof(double));
    corr2(s->X, s->n, u, NULL, s->L, z, NULL, s->K, ws);
    free(ws);
}

/* MKL/ssa_opt.h:910-1004 -- block form, two columns per transform */
void orc_matvec_block(const orc_ssa *s, const double *Vin, double *Yout, int b, int transpose)
{
    int lin = transpose ? s->L : s->K, lout = transpose ? s->K : s->L;
    double *ws = (double *)xcalloc(4 * (size_t)s->n, sizeof(double));
    for (int c = 0; c < b; c += 2) {
        const double *a = Vin + (size_t)c * lin;
        const double *bb = (c + 1 < b) ? Vin + (size_t)(c + 1) * lin : NULL;
        corr2(s->X, s->n, a, bb, lin, Yout + (size_t)c * lout,
              bb ? Yout + (size_t)(c + 1) * lout : NULL, lout, ws);
    }
    free(ws);
}

/* Randomized SVD: MKL/ssa_opt.h:1685-1763 with q subspace iterations
 * (Halko, Martinsson, Tropp 2011, Alg. 4.4) inserted between the range
 * finder and the projection.  Omega: K x b column-major, b = min(k+p, L, K). */
int orc_decompose(orc_ssa *s, int k, int p, int q, const double *Omega)
{
    if (!s || !s->X || k < 1 || !Omega || q < 0) return -1;
    int L = s->L, K = s->K;
    if (p <= 0) p = 8;                          /* MKL/ssa_opt.h:1692 */
    int b = imin(k + p, imin(L, K));            /* :1694 */
    k = imin(k, b);                             /* :1695 */
    free(s->U); free(s->V); free(s->sigma); free(s->lambda);
    s->U = (double *)xcalloc((size_t)L * k, sizeof(double));
    s->V = (double *)xcalloc((size_t)K * k, sizeof(double));
    s->sigma = (double *)xcalloc(k, sizeof(double));
    s->lambda = (double *)xcalloc(k, sizeof(double));
    double *Y = (double *)xcalloc((size_t)L * b, sizeof(double));
    double *Z = (double *)xcalloc((size_t)K * b, sizeof(double));
    double *R = (double *)xcalloc((size_t)b * b, sizeof(double));
    double *Wr = (double *)xcalloc((size_t)b * b, sizeof(double));
    double *S = (double *)xcalloc(b, sizeof(double));
    double *Bl = (double *)xcalloc((size_t)K * b, sizeof(double));
    if (!s->U || !s->V || !s->sigma || !s->lambda || !Y || !Z || !R || !Wr || !S || !Bl) {
        free(Y); free(Z); free(R); free(Wr); free(S); free(Bl);
        return -1;
    }
    orc_matvec_block(s, Omega, Y, b, 0);        /* Y = H Omega (:1720) */
    orc_qr_q(Y, L, b, NULL);                    /* Q = orth(Y) (:1723-1725) */
    for (int it = 0; it < q; it++) {
        orc_matvec_block(s, Y, Z, b, 1);
        orc_qr_q(Z, K, b, NULL);
        orc_matvec_block(s, Z, Y, b, 0);
        orc_qr_q(Y, L, b, NULL);
    }
    orc_matvec_block(s, Y, Z, b, 1);            /* B = H^T Q (:1728) */
    /* SVD of B (K x b) (:1731): QR then Jacobi on R */
    memcpy(Bl, Z, sizeof(double) * (size_t)K * b);
    orc_qr_q(Bl, K, b, R);
    orc_jacobi_svd(R, b, b, S, Wr);             /* R = Ur S Wr^T, R now holds Ur */
    /* U = Q Wr[:, :k] (:1736); V = (Q_B Ur)[:, :k] (:1737-1738) */
    for (int j = 0; j < k; j++) {
        for (int i = 0; i < L; i++) {
            double acc = 0.0;
            for (int c = 0; c < b; c++) acc += Y[(size_t)c * L + i] * Wr[(size_t)j * b + c];
            s->U[(size_t)j * L + i] = acc;
        }
        for (int i = 0; i < K; i++) {
            double acc = 0.0;
            for (int c = 0; c < b; c++) acc += Bl[(size_t)c * K + i] * R[(size_t)j * b + c];
            s->V[(size_t)j * K + i] = acc;
        }
    }
    s->total_variance = 0.0;
    for (int i = 0; i < k; i++) {               /* :1741-1746 */
        s->sigma[i] = S[i];
        s->lambda[i] = S[i] * S[i];
        s->total_variance += s->lambda[i];
    }
    for (int i = 0; i < k; i++) {               /* sign fix :1749-1759 */
        double sum = 0.0;
        for (int t = 0; t < L; t++) sum += s->U[(size_t)i * L + t];
        if (sum < 0) {
            for (int t = 0; t < L; t++) s->U[(size_t)i * L + t] = -s->U[(size_t)i * L + t];
            for (int t = 0; t < K; t++) s->V[(size_t)i * K + t] = -s->V[(size_t)i * K + t];
        }
    }
    s->k = k; s->b = b; s->decomposed = 1;
    free(Y); free(Z); free(R); free(Wr); free(S); free(Bl);
    return 0;
}

/* Linear convolution of a (len la) and c (len lc) through FFT(a)*FFT(c),
 * accumulated into acc (n complex).  Both factors are scaled by sqrt(scale)
 * so the packed transform stays balanced. */
static void conv_acc(const double *a, int la, const double *c, int lc, double scale,
                     int n, double *acc, double *ws)
{
    double r = sqrt(fabs(scale)), sg = scale < 0 ? -1.0 : 1.0;
    memset(ws, 0, sizeof(double) * 2 * (size_t)n);
    for (int i = 0; i < la; i++) ws[2 * i] = r * a[i];
    for (int i = 0; i < lc; i++) ws[2 * i + 1] = r * c[i];
    orc_fft(ws, n, -1);
    for (int kk = 0; kk < n; kk++) {
        int m = (n - kk) & (n - 1);
        double zr = ws[2 * kk], zi = ws[2 * kk + 1], mr = ws[2 * m], mi = -ws[2 * m + 1];
        /* A = (Z[k] + conj Z[-k])/2 ; C = (Z[k] - conj Z[-k])/(2i) */
        double Ar = 0.5 * (zr + mr), Ai = 0.5 * (zi + mi);
        double Cr = 0.5 * (zi - mi), Ci = -0.5 * (zr - mr);
        acc[2 * kk] += sg * (Ar * Cr - Ai * Ci);
        acc[2 * kk + 1] += sg * (Ar * Ci + Ai * Cr);
    }
}

/* MKL/ssa_opt.h:2325-2400: frequency-domain diagonal averaging;
 * out-of-range indices are skipped, n_group < 1 fails. */
int orc_reconstruct(const orc_ssa *s, const int *group, int n_group, double *out)
{
    if (!s || !s->decomposed || !group || !out || n_group < 1) return -1;
    int n = s->n, L = s->L, K = s->K, N = s->N;
    double *acc = (double *)xcalloc(2 * (size_t)n, sizeof(double));
    double *ws = (double *)xcalloc(2 * (size_t)n, sizeof(double));
    for (int g = 0; g < n_group; g++) {
        int idx = group[g];
        if (idx < 0 || idx >= s->k) continue;
        conv_acc(s->U + (size_t)idx * L, L, s->V + (size_t)idx * K, K, s->sigma[idx], n, acc, ws);
    }
    orc_fft(acc, n, +1);
    for (int t = 0; t < N; t++) out[t] = acc[2 * t] / n * s->w[t];
    free(acc); free(ws);
    return 0;
}

/* MKL/ssa_opt.h:2407-2477: G[i,t] = h_i[t] sqrt(w_t) / ||h_i||_w,
 * h_i = u_i (*) v_i, W = G G^T (rows with sigma*||h||_w <= 1e-12 are zero). */
int orc_wcorr(const orc_ssa *s, double *W)
{
    if (!s || !s->decomposed || !W) return -1;
    int n = s->n, L = s->L, K = s->K, N = s->N, k = s->k;
    double *G = (double *)xcalloc((size_t)k * N, sizeof(double));
    double *acc = (double *)xcalloc(2 * (size_t)n, sizeof(double));
    double *ws = (double *)xcalloc(2 * (size_t)n, sizeof(double));
    for (int i = 0; i < k; i++) {
        memset(acc, 0, sizeof(double) * 2 * (size_t)n);
        conv_acc(s->U + (size_t)i * L, L, s->V + (size_t)i * K, K, 1.0, n, acc, ws);
        orc_fft(acc, n, +1);
        double nsq = 0.0;
        for (int t = 0; t < N; t++) { double h = acc[2 * t] / n; nsq += h * h * s->w[t]; }
        double sg = s->sigma[i];
        double norm = sqrt(nsq * sg * sg);
        double scale = norm > 1e-12 ? sg / norm : 0.0;
        for (int t = 0; t < N; t++) G[(size_t)i * N + t] = scale * (acc[2 * t] / n) * sqrt(s->w[t]);
    }
    for (int i = 0; i < k; i++)
        for (int j = i; j < k; j++) {
            double a = 0.0;
            for (int t = 0; t < N; t++) a += G[(size_t)i * N + t] * G[(size_t)j * N + t];
            W[(size_t)i * k + j] = a;
            W[(size_t)j * k + i] = a;
        }
    free(G); free(acc); free(ws);
    return 0;
}

/* MKL/ssa_opt.h:2766-2816 */
int orc_lrf(const orc_ssa *s, const int *group, int n_group, double *R, double *nu2)
{
    if (!s || !s->decomposed || !group || n_group < 1) return -1;
    int L = s->L;
    double nu = 0.0;
    for (int g = 0; g < n_group; g++) {
        int idx = group[g];
        if (idx < 0 || idx >= s->k) return -1;
        double pi = s->U[(size_t)idx * L + L - 1];
        nu += pi * pi;
    }
    if (nu2) *nu2 = nu;
    if (nu >= 1.0 - 1e-10) return -1;
    double sc = 1.0 / (1.0 - nu);
    for (int j = 0; j < L - 1; j++) {
        double acc = 0.0;
        for (int g = 0; g < n_group; g++) {
            int idx = group[g];
            acc += s->U[(size_t)idx * L + L - 1] * s->U[(size_t)idx * L + j];
        }
        R[j] = acc * sc;
    }
    return 0;
}

/* MKL/ssa_opt.h:2818-2841 recurrence on the last L-1 values */
static int lrf_run(const double *R, int L, const double *base, int base_len, int h, double *out)
{
    int wsz = L - 1;
    if (base_len < wsz || h < 1) return -1;
    double *buf = (double *)xcalloc((size_t)wsz + h, sizeof(double));
    for (int i = 0; i < wsz; i++) buf[i] = base[base_len - wsz + i];
    for (int t = 0; t < h; t++) {
        double f = 0.0;
        for (int j = 0; j < wsz; j++) f += R[j] * buf[t + j];
        buf[wsz + t] = f;
        out[t] = f;
    }
    free(buf);
    return 0;
}

int orc_forecast(const orc_ssa *s, const int *group, int n_group, int h, double *out)  /* :2843-2867 */
{
    if (!s || !s->decomposed || !group || !out || n_group < 1 || h < 1) return -1;
    double *R = (double *)xcalloc(s->L, sizeof(double));
    double *rec = (double *)xcalloc(s->N, sizeof(double));
    int rc = orc_lrf(s, group, n_group, R, NULL);
    if (rc == 0) rc = orc_reconstruct(s, group, n_group, rec);
    if (rc == 0) rc = lrf_run(R, s->L, rec, s->N, h, out);
    free(R); free(rec);
    return rc;
}

int orc_forecast_full(const orc_ssa *s, const int *group, int n_group, int h, double *out)  /* :2869-2885 */
{
    if (!s || !s->decomposed || !group || !out || n_group < 1 || h < 1) return -1;
    double *R = (double *)xcalloc(s->L, sizeof(double));
    int rc = orc_lrf(s, group, n_group, R, NULL);
    if (rc == 0) rc = orc_reconstruct(s, group, n_group, out);
    if (rc == 0) rc = lrf_run(R, s->L, out, s->N, h, out + s->N);
    free(R);
    return rc;
}

/* V-forecast, MKL/ssa_opt.h:2906-2965 (plain), :3016-3078 (fast: base given):
 * each step projects the last L-1 values on every eigenvector of the group,
 * x = (sum_g pi_g (U_g[0:L-1] . z)) / (1 - nu^2), in the reference's order
 * (one sequential dot per component, then the weighted sum). */
int orc_vforecast(const orc_ssa *s, const int *group, int n_group, const double *base, int base_len, int h,
                  double *out)
{
    if (!s || !s->decomposed || !group || !out || n_group < 1 || h < 1) return -1;
    int L = s->L, w = L - 1;
    double nu = 0.0;
    for (int g = 0; g < n_group; g++) {
        int idx = group[g];
        if (idx < 0 || idx >= s->k) return -1;
        double pi = s->U[(size_t)idx * L + L - 1];
        nu += pi * pi;
    }
    if (nu >= 1.0 - 1e-10) return -1;
    double sc = 1.0 / (1.0 - nu);
    double *rec = NULL;
    if (!base) {
        rec = (double *)xcalloc(s->N, sizeof(double));
        if (orc_reconstruct(s, group, n_group, rec) != 0) { free(rec); return -1; }
        base = rec;
        base_len = s->N;
    }
    if (base_len < w) { free(rec); return -1; }
    double *buf = (double *)xcalloc((size_t)w + h, sizeof(double));
    for (int i = 0; i < w; i++) buf[i] = base[base_len - w + i];
    for (int t = 0; t < h; t++) {
        double x = 0.0;
        for (int g = 0; g < n_group; g++) {
            const double *u = s->U + (size_t)group[g] * L;
            double dot = 0.0;
            for (int j = 0; j < w; j++) dot += u[j] * buf[t + j];
            x += u[L - 1] * dot;
        }
        x *= sc;
        buf[w + t] = x;
        out[t] = x;
    }
    free(buf);
    free(rec);
    return 0;
}

/* ---- Cadzow iterations and gap filling (MKL/ssa_opt.h:3147-3878) ----
 * The reference builds a fresh context (test-matrix stream restarted) for each
 * iteration; every rank-r decomposition here is the one the device path
 * makes: b = min(r + 8, L, K) columns, q = 2, the Philox matrix of seed 42 at
 * the given counter. */
static int cz_decompose(orc_ssa *s, int rank, uint64_t counter)
{
    int b = imin(rank + 8, imin(s->L, s->K));
    double *om = (double *)xcalloc((size_t)s->K * b, sizeof(double));
    orc_philox_gaussian(42, counter, s->K, b, om);
    int rc = orc_decompose(s, rank, 8, 2, om);
    free(om);
    return rc;
}

static double cz_norm(const double *x, int n)
{
    double a = 0.0;
    for (int i = 0; i < n; i++) a += x[i] * x[i];
    return sqrt(a);
}

/* :3165-3292; out3 = {iterations, final_diff, converged}; returns iterations or -1 */
int orc_cadzow(const double *x, int N, int L, int rank, int max_iter, double tol, double *output, double *out3)
{
    if (!x || !output || N < 4 || L < 2 || L > N - 1 || rank < 1 || max_iter < 1) return -1;
    double *y = (double *)xcalloc(N, sizeof(double)), *yn = (double *)xcalloc(N, sizeof(double));
    int *group = (int *)xcalloc(rank, sizeof(int));
    for (int i = 0; i < rank; i++) group[i] = i;
    memcpy(y, x, sizeof(double) * N);
    double ynorm = cz_norm(y, N);
    if (ynorm < 1e-15) ynorm = 1.0;
    double diff = 0.0;
    int conv = 0, it, rc = 0;
    for (it = 0; it < max_iter; it++) {
        orc_ssa s;
        memset(&s, 0, sizeof(s));
        rc = orc_init(&s, y, N, L);
        if (rc == 0) rc = cz_decompose(&s, rank, 0);
        if (rc == 0) rc = orc_reconstruct(&s, group, rank, yn);
        orc_free(&s);
        if (rc) break;
        double d2 = 0.0;
        for (int i = 0; i < N; i++) d2 += (yn[i] - y[i]) * (yn[i] - y[i]);
        diff = sqrt(d2) / ynorm;
        double *t = y; y = yn; yn = t;
        if (diff < tol) { conv = 1; it++; break; }
    }
    if (rc == 0) {
        memcpy(output, y, sizeof(double) * N);
        if (out3) { out3[0] = it; out3[1] = diff; out3[2] = conv; }
    }
    free(y); free(yn); free(group);
    return rc ? -1 : it;
}

/* :3319-3440 on a context whose test-matrix counter starts at `counter0`
 * (one increment per decomposition) */
int orc_cadzow_inplace(orc_ssa *s, int rank, int max_iter, double tol, uint64_t counter0, double *out3)
{
    if (!s || !s->X || rank < 1 || max_iter < 1) return -1;
    int N = s->N;
    double *yn = (double *)xcalloc(N, sizeof(double)), *yc = (double *)xcalloc(N, sizeof(double));
    int *group = (int *)xcalloc(rank, sizeof(int));
    for (int i = 0; i < rank; i++) group[i] = i;
    double ynorm = 1.0, diff = 0.0;
    int conv = 0, it, rc = 0;
    uint64_t ctr = counter0;
    for (it = 0; it < max_iter; it++) {
        rc = cz_decompose(s, rank, ctr++);
        if (rc) break;
        if (it > 0) memcpy(yc, yn, sizeof(double) * N);
        rc = orc_reconstruct(s, group, rank, yn);
        if (rc) break;
        if (it == 0) {
            ynorm = cz_norm(yn, N);
            if (ynorm < 1e-15) ynorm = 1.0;
        } else {
            double d2 = 0.0;
            for (int i = 0; i < N; i++) d2 += (yn[i] - yc[i]) * (yn[i] - yc[i]);
            diff = sqrt(d2) / ynorm;
            if (diff < tol) { conv = 1; it++; break; }
        }
        if (it < max_iter - 1) orc_update_signal(s, yn);
    }
    if (rc == 0 && out3) { out3[0] = it; out3[1] = diff; out3[2] = conv; }
    free(yn); free(yc); free(group);
    return rc ? -1 : it;
}

/* :3455-3496 */
static void cz_interp(double *x, const unsigned char *gap, int N)
{
    for (int i = 0; i < N;) {
        if (!gap[i]) { i++; continue; }
        int a = i;
        while (i < N && gap[i]) i++;
        int b = i, len = b - a;
        double lv = a > 0 ? x[a - 1] : (b < N ? x[b] : 0.0);
        double rv = b < N ? x[b] : lv;
        for (int j = 0; j < len; j++) {
            double t = len > 1 ? (double)(j + 1) / (len + 1) : 0.5;
            x[a + j] = lv + t * (rv - lv);
        }
    }
}

/* :3498-3683; out4 = {iterations, final_diff, converged, n_gaps} */
int orc_gapfill(double *x, int N, int L, int rank, int max_iter, double tol, double *out4)
{
    if (!x || N < 4 || L < 2 || L > N - 1 || rank < 1) return -1;
    unsigned char *gap = (unsigned char *)xcalloc(N, 1);
    int ng = 0;
    for (int i = 0; i < N; i++) if (isnan(x[i])) { gap[i] = 1; ng++; }
    if (ng == 0) {
        free(gap);
        if (out4) { out4[0] = 0; out4[1] = 0; out4[2] = 1; out4[3] = 0; }
        return 0;
    }
    if (N - ng < L + rank) { free(gap); return -1; }
    cz_interp(x, gap, N);
    double *ov = (double *)xcalloc(ng, sizeof(double)), *nv = (double *)xcalloc(ng, sizeof(double));
    double *rec = (double *)xcalloc(N, sizeof(double));
    int *group = (int *)xcalloc(rank, sizeof(int));
    for (int i = 0; i < rank; i++) group[i] = i;
    for (int i = 0, g = 0; i < N; i++) if (gap[i]) ov[g++] = x[i];
    double diff = 1e10;
    int conv = 0, it, rc = 0;
    for (it = 0; it < max_iter; it++) {
        orc_ssa s;
        memset(&s, 0, sizeof(s));
        rc = orc_init(&s, x, N, L);
        if (rc == 0) rc = cz_decompose(&s, rank, 0);
        if (rc == 0) rc = orc_reconstruct(&s, group, rank, rec);
        orc_free(&s);
        if (rc) break;
        for (int i = 0, g = 0; i < N; i++) if (gap[i]) { nv[g++] = rec[i]; x[i] = rec[i]; }
        double sd = 0.0, so = 0.0;
        for (int g = 0; g < ng; g++) { sd += (nv[g] - ov[g]) * (nv[g] - ov[g]); so += ov[g] * ov[g]; }
        diff = so > 1e-12 ? sqrt(sd / so) : sqrt(sd);
        if (diff < tol) { conv = 1; it++; break; }
        memcpy(ov, nv, sizeof(double) * ng);
    }
    if (rc == 0 && out4) { out4[0] = it; out4[1] = diff; out4[2] = conv; out4[3] = ng; }
    free(gap); free(ov); free(nv); free(rec); free(group);
    return rc ? -1 : 0;
}

static int cz_seg_forecast(const double *seg, int len, int L, int rank, const int *group, int h, double *out)
{
    orc_ssa s;
    memset(&s, 0, sizeof(s));
    int rc = orc_init(&s, seg, len, L);
    if (rc == 0) rc = cz_decompose(&s, rank, 0);
    if (rc == 0) rc = orc_forecast(&s, group, rank, h, out);
    orc_free(&s);
    return rc;
}

/* :3685-3877 (a side whose forecast fails counts as absent) */
int orc_gapfill_simple(double *x, int N, int L, int rank, double *out4)
{
    if (!x || N < 4 || L < 2 || L > N - 1 || rank < 1) return -1;
    unsigned char *gap = (unsigned char *)xcalloc(N, 1);
    int ng = 0;
    for (int i = 0; i < N; i++) if (isnan(x[i])) { gap[i] = 1; ng++; }
    int *group = (int *)xcalloc(rank, sizeof(int));
    for (int i = 0; i < rank; i++) group[i] = i;
    double *fl = (double *)xcalloc(N, sizeof(double)), *fr = (double *)xcalloc(N, sizeof(double));
    double *rev = (double *)xcalloc(N, sizeof(double)), *tmp = (double *)xcalloc(N, sizeof(double));
    for (int i = 0; i < N;) {
        if (!gap[i]) { i++; continue; }
        int a = i;
        while (i < N && gap[i]) i++;
        int b = i, len = b - a, hl = 0, hr = 0;
        if (a >= L + rank) hl = cz_seg_forecast(x, a, L, rank, group, len, fl) == 0;
        if (N - b >= L + rank) {
            for (int j = 0; j < N - b; j++) rev[j] = x[N - 1 - j];
            hr = cz_seg_forecast(rev, N - b, L, rank, group, len, tmp) == 0;
            if (hr) for (int j = 0; j < len; j++) fr[j] = tmp[len - 1 - j];
        }
        for (int j = 0; j < len; j++) {
            double v = 0.0, w = 0.0;
            if (hl) { double wl = (double)(len - j) / len; v += wl * fl[j]; w += wl; }
            if (hr) { double wr = (double)(j + 1) / len; v += wr * fr[j]; w += wr; }
            if (w > 0) x[a + j] = v / w;
            else {
                double lv = a > 0 ? x[a - 1] : 0.0, rv = b < N ? x[b] : lv;
                x[a + j] = (lv + rv) / 2.0;
            }
        }
    }
    if (out4) { out4[0] = 1; out4[1] = 0; out4[2] = 1; out4[3] = ng; }
    free(gap); free(group); free(fl); free(fr); free(rev); free(tmp);
    return 0;
}

double orc_variance_explained(const orc_ssa *s, int start, int end)   /* :3118-3128 */
{
    if (!s || !s->decomposed || start < 0 || s->total_variance <= 0) return 0.0;
    if (end < 0 || end >= s->k) end = s->k - 1;
    double a = 0.0;
    for (int i = start; i <= end; i++) a += s->lambda[i];
    return a / s->total_variance;
}

int orc_dims(const orc_ssa *s, int *o)
{
    o[0] = s->N; o[1] = s->L; o[2] = s->K; o[3] = s->n;
    return s->decomposed ? s->k : 0;
}

int orc_get(const orc_ssa *s, double *U, double *V, double *sigma)
{
    if (!s->decomposed) return 0;
    if (U) memcpy(U, s->U, sizeof(double) * (size_t)s->L * s->k);
    if (V) memcpy(V, s->V, sizeof(double) * (size_t)s->K * s->k);
    if (sigma) memcpy(sigma, s->sigma, sizeof(double) * s->k);
    return s->k;
}

/* ------------------------------------------------------------------------ */
/* MSSA: MKL/ssa_opt.h:3907-4283.  Stacked (M L) x K trajectory; the        */
/* reference uses the reversed-convolution form (:4069-4101), which equals  */
/* the correlation used here in exact arithmetic.                           */
/* ------------------------------------------------------------------------ */

orc_mssa *orc_mssa_new(void) { return (orc_mssa *)calloc(1, sizeof(orc_mssa)); }
void orc_mssa_delete(orc_mssa *m) { if (m) { orc_mssa_free(m); free(m); } }

void orc_mssa_free(orc_mssa *m)
{
    if (!m) return;
    free(m->X); free(m->w); free(m->U); free(m->V); free(m->sigma); free(m->lambda);
    memset(m, 0, sizeof(*m));
}

int orc_mssa_init(orc_mssa *m, const double *X, int M, int N, int L)   /* :3907-4057 */
{
    if (!m || !X || M < 1 || N < 4 || L < 2 || L > N - 1) return -1;
    memset(m, 0, sizeof(*m));
    m->M = M; m->N = N; m->L = L; m->K = N - L + 1;
    m->n = orc_next_pow2(N + m->K - 1);
    m->X = (double *)xcalloc((size_t)M * 2 * m->n, sizeof(double));
    m->w = (double *)xcalloc(N, sizeof(double));
    if (!m->X || !m->w) { orc_mssa_free(m); return -1; }
    for (int c = 0; c < M; c++) spectrum(X + (size_t)c * N, N, m->n, m->X + (size_t)c * 2 * m->n);
    for (int t = 0; t < N; t++) {
        int c = imin(imin(t + 1, L), imin(m->K, N - t));
        m->w[t] = c > 0 ? 1.0 / c : 0.0;
    }
    return 0;
}

void orc_mssa_matvec(const orc_mssa *m, const double *v, double *y)   /* :4063-4077 */
{
    double *ws = (double *)xcalloc(4 * (size_t)m->n, sizeof(double));
    for (int c = 0; c < m->M; c++)
        corr2(m->X + (size_t)c * 2 * m->n, m->n, v, NULL, m->K, y + (size_t)c * m->L, NULL, m->L, ws);
    free(ws);
}

void orc_mssa_matvec_T(const orc_mssa *m, const double *u, double *z)  /* :4079-4101 */
{
    double *ws = (double *)xcalloc(4 * (size_t)m->n, sizeof(double));
    double *tmp = (double *)xcalloc(m->K, sizeof(double));
    memset(z, 0, sizeof(double) * m->K);
    for (int c = 0; c < m->M; c++) {
        corr2(m->X + (size_t)c * 2 * m->n, m->n, u + (size_t)c * m->L, NULL, m->L, tmp, NULL, m->K, ws);
        for (int j = 0; j < m->K; j++) z[j] += tmp[j];
    }
    free(ws); free(tmp);
}

int orc_mssa_decompose(orc_mssa *m, int k, int p, int q, const double *Omega)  /* :4116-4210 */
{
    if (!m || !m->X || k < 1 || !Omega || q < 0) return -1;
    int ML = m->M * m->L, K = m->K;
    if (p <= 0) p = 8;
    int b = imin(k + p, imin(ML, K));
    k = imin(k, b);
    free(m->U); free(m->V); free(m->sigma); free(m->lambda);
    m->U = (double *)xcalloc((size_t)ML * k, sizeof(double));
    m->V = (double *)xcalloc((size_t)K * k, sizeof(double));
    m->sigma = (double *)xcalloc(k, sizeof(double));
    m->lambda = (double *)xcalloc(k, sizeof(double));
    double *Y = (double *)xcalloc((size_t)ML * b, sizeof(double));
    double *Z = (double *)xcalloc((size_t)K * b, sizeof(double));
    double *R = (double *)xcalloc((size_t)b * b, sizeof(double));
    double *Wr = (double *)xcalloc((size_t)b * b, sizeof(double));
    double *S = (double *)xcalloc(b, sizeof(double));
    for (int c = 0; c < b; c++) orc_mssa_matvec(m, Omega + (size_t)c * K, Y + (size_t)c * ML);
    orc_qr_q(Y, ML, b, NULL);
    for (int it = 0; it < q; it++) {
        for (int c = 0; c < b; c++) orc_mssa_matvec_T(m, Y + (size_t)c * ML, Z + (size_t)c * K);
        orc_qr_q(Z, K, b, NULL);
        for (int c = 0; c < b; c++) orc_mssa_matvec(m, Z + (size_t)c * K, Y + (size_t)c * ML);
        orc_qr_q(Y, ML, b, NULL);
    }
    for (int c = 0; c < b; c++) orc_mssa_matvec_T(m, Y + (size_t)c * ML, Z + (size_t)c * K);
    orc_qr_q(Z, K, b, R);                        /* Z <- Q_B */
    orc_jacobi_svd(R, b, b, S, Wr);
    for (int j = 0; j < k; j++) {
        for (int i = 0; i < ML; i++) {
            double acc = 0.0;
            for (int c = 0; c < b; c++) acc += Y[(size_t)c * ML + i] * Wr[(size_t)j * b + c];
            m->U[(size_t)j * ML + i] = acc;
        }
        for (int i = 0; i < K; i++) {
            double acc = 0.0;
            for (int c = 0; c < b; c++) acc += Z[(size_t)c * K + i] * R[(size_t)j * b + c];
            m->V[(size_t)j * K + i] = acc;
        }
    }
    m->total_variance = 0.0;
    for (int i = 0; i < k; i++) {
        m->sigma[i] = S[i]; m->lambda[i] = S[i] * S[i]; m->total_variance += m->lambda[i];
    }
    for (int i = 0; i < k; i++) {               /* sign fix over all M L rows (:4188-4199) */
        double sum = 0.0;
        for (int t = 0; t < ML; t++) sum += m->U[(size_t)i * ML + t];
        if (sum < 0) {
            for (int t = 0; t < ML; t++) m->U[(size_t)i * ML + t] = -m->U[(size_t)i * ML + t];
            for (int t = 0; t < K; t++) m->V[(size_t)i * K + t] = -m->V[(size_t)i * K + t];
        }
    }
    m->k = k; m->b = b; m->decomposed = 1;
    free(Y); free(Z); free(R); free(Wr); free(S);
    return 0;
}

int orc_mssa_reconstruct(const orc_mssa *m, int series, const int *group, int n_group, double *out)  /* :4212-4242 */
{
    if (!m || !m->decomposed || !group || !out || n_group < 1 || series < 0 || series >= m->M) return -1;
    int n = m->n, L = m->L, K = m->K, N = m->N, ML = m->M * m->L;
    double *acc = (double *)xcalloc(2 * (size_t)n, sizeof(double));
    double *ws = (double *)xcalloc(2 * (size_t)n, sizeof(double));
    for (int g = 0; g < n_group; g++) {
        int idx = group[g];
        if (idx < 0 || idx >= m->k) continue;
        conv_acc(m->U + (size_t)idx * ML + (size_t)series * L, L, m->V + (size_t)idx * K, K,
                 m->sigma[idx], n, acc, ws);
    }
    orc_fft(acc, n, +1);
    for (int t = 0; t < N; t++) out[t] = acc[2 * t] / n * m->w[t];
    free(acc); free(ws);
    return 0;
}

int orc_mssa_get(const orc_mssa *m, double *U, double *V, double *sigma)
{
    if (!m->decomposed) return 0;
    int ML = m->M * m->L;
    if (U) memcpy(U, m->U, sizeof(double) * (size_t)ML * m->k);
    if (V) memcpy(V, m->V, sizeof(double) * (size_t)m->K * m->k);
    if (sigma) memcpy(sigma, m->sigma, sizeof(double) * m->k);
    return m->k;
}

int orc_mssa_contributions(const orc_mssa *m, double *c)   /* :4256-4270 */
{
    if (!m || !m->decomposed || !c) return -1;
    int ML = m->M * m->L, k = m->k;
    for (int i = 0; i < k; i++)
        for (int s = 0; s < m->M; s++) {
            double a = 0.0;
            const double *u = m->U + (size_t)i * ML + (size_t)s * m->L;
            for (int t = 0; t < m->L; t++) a += u[t] * u[t];
            c[(size_t)s * k + i] = a;
        }
    return 0;
}
