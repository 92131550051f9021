/* ssa_esprit.c -- ESPRIT signal-pole estimation from the eigenvectors of a
 * decomposed context (MKL/ssa_opt.h:4289-4558): small FP64 host arithmetic on
 * the host mirror of U, k <= SSA_BMAX.
 *
 *   U_up = U[0:L-1, g], U_down = U[1:L, g]     (the group's eigenvectors)
 *   Z    = argmin || U_up Z - U_down ||_F      (Householder QR least squares;
 *                                              the reference calls dgels)
 *   poles = eig(Z)                             (Hessenberg reduction and
 *                                              Francis double-shift QR; the
 *                                              reference calls dgeev)
 *
 * From each pole: modulus, rate log(modulus) (-30 below 1e-12), frequency
 * arg/(2 pi) in cycles per sample, period 1/|frequency| (infinite below
 * 1e-12).  The order of the poles is this solver's deflation order, not
 * dgeev's. */
#include "../../include/ssa_opt.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* min ||A Z - B||_F for A m x k, B m x k (column-major, ld m), m >= k, by
 * Householder QR of A applied to B; A and B are overwritten, Z (k x k,
 * column-major) returned.  -1 when A is exactly rank deficient. */
static int lstsq_qr(int m, int k, double *A, double *B, double *Z)
{
    for (int j = 0; j < k; j++) {
        double *a = A + (size_t)j * m;
        double nx = 0.0;
        for (int i = j; i < m; i++) nx += a[i] * a[i];
        nx = sqrt(nx);
        if (nx == 0.0) return -1;
        const double alpha = a[j] > 0 ? -nx : nx;
        /* v = x - alpha e1 stored in a[j..m), v^T v = 2 nx (nx + |x_j|) */
        a[j] -= alpha;
        const double vv = 2.0 * nx * (nx + fabs(a[j] + alpha));
        for (int c = j + 1; c < k; c++) {
            double *col = A + (size_t)c * m, s = 0.0;
            for (int i = j; i < m; i++) s += a[i] * col[i];
            s = 2.0 * s / vv;
            for (int i = j; i < m; i++) col[i] -= s * a[i];
        }
        for (int c = 0; c < k; c++) {
            double *col = B + (size_t)c * m, s = 0.0;
            for (int i = j; i < m; i++) s += a[i] * col[i];
            s = 2.0 * s / vv;
            for (int i = j; i < m; i++) col[i] -= s * a[i];
        }
        a[j] = alpha;   /* R[j][j] */
    }
    /* back substitution R Z = (Q^T B)[0:k) */
    for (int c = 0; c < k; c++) {
        const double *b = B + (size_t)c * m;
        double *z = Z + (size_t)c * k;
        for (int i = k - 1; i >= 0; i--) {
            double s = b[i];
            for (int l = i + 1; l < k; l++) s -= A[(size_t)l * m + i] * z[l];
            const double rii = A[(size_t)i * m + i];
            if (rii == 0.0) return -1;
            z[i] = s / rii;
        }
    }
    return 0;
}

#define H(i, j) h[(size_t)(i) * n + (j)]

/* reduce the row-major n x n matrix h to upper Hessenberg form by Householder
 * similarity transforms (the eigenvalues are kept) */
static void hessenberg(int n, double *h, double *v)
{
    for (int c = 0; c + 2 < n; c++) {
        const int len = n - c - 1;
        double nx = 0.0;
        for (int i = 0; i < len; i++) {
            v[i] = H(c + 1 + i, c);
            nx += v[i] * v[i];
        }
        nx = sqrt(nx);
        if (nx == 0.0) continue;
        const double alpha = v[0] > 0 ? -nx : nx;
        v[0] -= alpha;
        double vv = 0.0;
        for (int i = 0; i < len; i++) vv += v[i] * v[i];
        if (vv == 0.0) continue;
        for (int j = 0; j < n; j++) {          /* rows c+1.. from the left */
            double s = 0.0;
            for (int i = 0; i < len; i++) s += v[i] * H(c + 1 + i, j);
            s = 2.0 * s / vv;
            for (int i = 0; i < len; i++) H(c + 1 + i, j) -= s * v[i];
        }
        for (int i = 0; i < n; i++) {          /* columns c+1.. from the right */
            double s = 0.0;
            for (int j = 0; j < len; j++) s += H(i, c + 1 + j) * v[j];
            s = 2.0 * s / vv;
            for (int j = 0; j < len; j++) H(i, c + 1 + j) -= s * v[j];
        }
        for (int i = 1; i < len; i++) H(c + 1 + i, c) = 0.0;
    }
}

/* eigenvalues of the upper Hessenberg h by the Francis double-shift QR
 * iteration with deflation (the scheme of EISPACK's hqr, Wilkinson and
 * Reinsch, Handbook for Automatic Computation II, 1971): wr + i wi.
 * -1 if a window needs more than 60 iterations. */
static int francis_eigvals(int n, double *h, double *wr, double *wi)
{
    double anorm = 0.0;
    for (int i = 0; i < n; i++)
        for (int j = i > 0 ? i - 1 : 0; j < n; j++) anorm += fabs(H(i, j));
    int nn = n - 1, its = 0;
    double t = 0.0;   /* accumulated exceptional shifts */
    while (nn >= 0) {
        int l;
        for (l = nn; l >= 1; l--) {           /* small subdiagonal: split */
            double s = fabs(H(l - 1, l - 1)) + fabs(H(l, l));
            if (s == 0.0) s = anorm;
            if (fabs(H(l, l - 1)) <= DBL_EPSILON * s) {
                H(l, l - 1) = 0.0;
                break;
            }
        }
        double x = H(nn, nn);
        if (l == nn) {                        /* one real root */
            wr[nn] = x + t;
            wi[nn] = 0.0;
            nn--;
            its = 0;
            continue;
        }
        double y = H(nn - 1, nn - 1), w = H(nn, nn - 1) * H(nn - 1, nn);
        if (l == nn - 1) {                    /* 2 x 2 block: two roots */
            const double p = 0.5 * (y - x), q = p * p + w;
            double z = sqrt(fabs(q));
            x += t;
            if (q >= 0.0) {
                z = p + (p >= 0 ? z : -z);
                wr[nn - 1] = wr[nn] = x + z;
                if (z != 0.0) wr[nn] = x - w / z;
                wi[nn - 1] = wi[nn] = 0.0;
            } else {
                wr[nn - 1] = wr[nn] = x + p;
                wi[nn - 1] = -z;
                wi[nn] = z;
            }
            nn -= 2;
            its = 0;
            continue;
        }
        if (its == 60) return -1;
        if (its == 10 || its == 20) {         /* exceptional shift */
            t += x;
            for (int i = 0; i <= nn; i++) H(i, i) -= x;
            const double s = fabs(H(nn, nn - 1)) + fabs(H(nn - 1, nn - 2));
            x = y = 0.75 * s;
            w = -0.4375 * s * s;
        }
        its++;
        /* shifts from the trailing 2 x 2; start the bulge where two
         * consecutive subdiagonals are small enough */
        int m;
        double p = 0.0, q = 0.0, r = 0.0, z;
        for (m = nn - 2; m >= l; m--) {
            z = H(m, m);
            r = x - z;
            double s = y - z;
            p = (r * s - w) / H(m + 1, m) + H(m, m + 1);
            q = H(m + 1, m + 1) - z - r - s;
            r = H(m + 2, m + 1);
            s = fabs(p) + fabs(q) + fabs(r);
            p /= s;
            q /= s;
            r /= s;
            if (m == l) break;
            const double u = fabs(H(m, m - 1)) * (fabs(q) + fabs(r));
            const double v = fabs(p) * (fabs(H(m - 1, m - 1)) + fabs(z) + fabs(H(m + 1, m + 1)));
            if (u <= DBL_EPSILON * v) break;
        }
        for (int i = m + 2; i <= nn; i++) {
            H(i, i - 2) = 0.0;
            if (i != m + 2) H(i, i - 3) = 0.0;
        }
        /* chase the bulge with 3 x 3 (last: 2 x 2) reflectors */
        for (int k = m; k <= nn - 1; k++) {
            if (k != m) {
                p = H(k, k - 1);
                q = H(k + 1, k - 1);
                r = (k != nn - 1) ? H(k + 2, k - 1) : 0.0;
                x = fabs(p) + fabs(q) + fabs(r);
                if (x != 0.0) {
                    p /= x;
                    q /= x;
                    r /= x;
                }
            }
            double s = sqrt(p * p + q * q + r * r);
            if (p < 0) s = -s;
            if (s == 0.0) continue;
            if (k == m) {
                if (l != m) H(k, k - 1) = -H(k, k - 1);
            } else {
                H(k, k - 1) = -s * x;
            }
            p += s;
            x = p / s;
            y = q / s;
            z = r / s;
            q /= p;
            r /= p;
            for (int j = k; j <= nn; j++) {
                p = H(k, j) + q * H(k + 1, j);
                if (k != nn - 1) {
                    p += r * H(k + 2, j);
                    H(k + 2, j) -= p * z;
                }
                H(k + 1, j) -= p * y;
                H(k, j) -= p * x;
            }
            const int imax = nn < k + 3 ? nn : k + 3;
            for (int i = l; i <= imax; i++) {
                p = x * H(i, k) + y * H(i, k + 1);
                if (k != nn - 1) {
                    p += z * H(i, k + 2);
                    H(i, k + 2) -= p * r;
                }
                H(i, k + 1) -= p * q;
                H(i, k) -= p;
            }
        }
    }
    return 0;
}

#undef H

/* Poles of the shift-invariance operator of the k columns U (L x k,
 * column-major): wr + i wi, k each.  0, or -1 (L < 3, rank-deficient U_up,
 * no convergence).  No counterpart in the reference; the ESPRIT core of
 * ssa_opt_parestimate, exported so that it can be checked on the host. */
int ssa_amd_esprit_poles(const double *U, int L, int k, double *wr, double *wi)
{
    if (!U || !wr || !wi || L < 3 || k < 1 || k > L - 1) return -1;
    const int m = L - 1;
    double *up = (double *)malloc(sizeof(double) * (size_t)m * k);
    double *dn = (double *)malloc(sizeof(double) * (size_t)m * k);
    double *Z = (double *)malloc(sizeof(double) * (size_t)k * k);
    double *h = (double *)malloc(sizeof(double) * (size_t)k * k);
    double *v = (double *)malloc(sizeof(double) * (size_t)k);
    int rc = -1;
    if (!up || !dn || !Z || !h || !v) goto out;
    for (int j = 0; j < k; j++) {
        memcpy(up + (size_t)j * m, U + (size_t)j * L, sizeof(double) * m);
        memcpy(dn + (size_t)j * m, U + (size_t)j * L + 1, sizeof(double) * m);
    }
    if (lstsq_qr(m, k, up, dn, Z) != 0) goto out;
    for (int i = 0; i < k; i++)
        for (int j = 0; j < k; j++) h[(size_t)i * k + j] = Z[(size_t)j * k + i];   /* row-major */
    hessenberg(k, h, v);
    rc = francis_eigvals(k, h, wr, wi);
out:
    free(up); free(dn); free(Z); free(h); free(v);
    return rc;
}

/* MKL/ssa_opt.h:4289-4302 */
void ssa_opt_parestimate_free(SSA_ParEstimate *result)
{
    if (!result) return;
    free(result->periods);
    free(result->frequencies);
    free(result->moduli);
    free(result->rates);
    memset(result, 0, sizeof(*result));
}

/* MKL/ssa_opt.h:4304-4558: group NULL or n_group <= 0 takes every computed
 * component */
int ssa_opt_parestimate(const SSA_Opt *ssa, const int *group, int n_group, SSA_ParEstimate *result)
{
    if (!ssa || !ssa->decomposed || !result || !ssa->U) return -1;
    if (ssa_amd_sync_mirrors(ssa)) return -1;
    const int k = (n_group > 0 && group) ? n_group : ssa->n_components;
    const int L = ssa->L;
    if (k < 1 || k > ssa->n_components || L < 3) return -1;
    double *Us = (double *)malloc(sizeof(double) * (size_t)L * k);
    double *wr = (double *)malloc(sizeof(double) * k), *wi = (double *)malloc(sizeof(double) * k);
    int rc = -1;
    if (!Us || !wr || !wi) goto out;
    for (int j = 0; j < k; j++) {
        const int c = (n_group > 0 && group) ? group[j] : j;
        if (c < 0 || c >= ssa->n_components) goto out;
        memcpy(Us + (size_t)j * L, ssa->U + (size_t)c * L, sizeof(double) * L);
    }
    if (ssa_amd_esprit_poles(Us, L, k, wr, wi) != 0) goto out;
    memset(result, 0, sizeof(*result));
    result->periods = (double *)malloc(sizeof(double) * k);
    result->frequencies = (double *)malloc(sizeof(double) * k);
    result->moduli = (double *)malloc(sizeof(double) * k);
    result->rates = (double *)malloc(sizeof(double) * k);
    result->n_components = k;
    if (!result->periods || !result->frequencies || !result->moduli || !result->rates) {
        ssa_opt_parestimate_free(result);
        goto out;
    }
    for (int i = 0; i < k; i++) {
        const double mod = sqrt(wr[i] * wr[i] + wi[i] * wi[i]);
        const double f = atan2(wi[i], wr[i]) / 6.283185307179586476925;
        result->moduli[i] = mod;
        result->rates[i] = mod > 1e-12 ? log(mod) : -30.0;
        result->frequencies[i] = f;
        result->periods[i] = fabs(f) > 1e-12 ? 1.0 / fabs(f) : INFINITY;
    }
    rc = 0;
out:
    free(Us); free(wr); free(wi);
    return rc;
}
