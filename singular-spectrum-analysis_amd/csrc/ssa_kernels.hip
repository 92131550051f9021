// ssa_kernels.hip -- HIP kernels of the SSA decomposition hot path (gfx950).
//
// Every kernel works on a batch of series described by ssa_desc records and
// a work list of ssa_item; a single SSA_Opt context is a batch of one.
//
//   k_spectra   cached spectra of the signal segments (init / update_signal,
//               MKL/ssa_opt.h:1255-1257, :1545-1558)
//   k_matvec    fused block Hankel matvec H*X / H^T*X: two columns per complex
//               transform, spectrum multiply, inverse, extract
//               (MKL/ssa_opt.h:910-1004); optionally draws the Gaussian test
//               matrix on the fly (MKL/ssa_opt.h:1719)
//   k_gram      partial Gram P^T P (optionally of P*M) over row tiles: the
//               CholeskyQR2 replacement of dgeqrf/dorgqr (MKL/ssa_opt.h:1723-1725)
//   k_chol      FP64 Cholesky of the summed Gram, triangular inverse, products
//   k_apply     panel * small matrix (Q = Y R^-1, U = Q W, V = B W S^-1;
//               k_apply_f32: the FP32-MFMA form for FP32 panels;
//               MKL/ssa_opt.h:1736-1738)
//   k_svd       FP64 one-sided Jacobi SVD of the b x b R factor of B = H^T Q,
//               sign convention sum(U[:,i]) >= 0 (MKL/ssa_opt.h:1731, :1749-1759)
//   k_recon     frequency-domain diagonal averaging (MKL/ssa_opt.h:2325-2400)
//   k_olap      overlap-add of the chunked convolutions and 1/count weights
//   k_wnorm     W-correlation normalisation (MKL/ssa_opt.h:2407-2477)
//   k_forecast  LRF coefficients and recurrence (MKL/ssa_opt.h:2766-2885), or the
//               V-forecast recurrence (MKL/ssa_opt.h:2906-3084)
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>

#include <mutex>

#include "ssa_device.h"
#include "ssa_fft.hpp"

using namespace ssa;

namespace {

constexpr int LDP = SSA_BMAX + 1;       // padded LDS row length of small matrices

__device__ __forceinline__ double *slot(double *small, const ssa_desc &d, int s)
{
    return small + d.off_small + (long long)s * SSA_BMAX * SSA_BMAX;
}

template <typename T> struct Panel {
    T *p;
    int rows, ld, cols;
};

template <typename T> __device__ __forceinline__ Panel<T> panel_of(T *bulk, const ssa_desc &d, int which)
{
    switch (which) {
    case SSA_PANEL_Y: return {bulk + d.off_Y, d.ML, d.ldY, d.b};
    case SSA_PANEL_Z: return {bulk + d.off_Z, d.K, d.ldZ, d.b};
    case SSA_PANEL_U: return {bulk + d.off_U, d.ML, d.ML, d.k};
    case SSA_PANEL_V: return {bulk + d.off_V, d.K, d.K, d.k};
    default: return {bulk + d.off_out, d.N, d.N, d.k};   // weighted component panel (W-corr)
    }
}

// ---------------------------------------------------------------------------
// Philox4x32-10 + Box-Muller: identical definition to the oracle's
// orc_philox_gaussian.  Row r of column quad j (columns 4j .. 4j+3) comes from
// counter (r, j, call_lo, call_hi); the call's four 32-bit outputs o are the
// uniforms of two Box-Muller pairs: u1 = o[2h] 2^-32, u2 = o[2h+1] 2^-32,
// rad = sqrt(-2 ln(1 - u1)), rad cos(2 pi u2) -> column 4j+2h and
// rad sin(2 pi u2) -> column 4j+2h+1.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ void gauss_quad(uint64_t seed, uint64_t call, int row, int quad, double (&g)[4])
{
    uint32_t c[4] = {(uint32_t)row, (uint32_t)quad, (uint32_t)call, (uint32_t)(call >> 32)};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const double u1 = (double)c[2 * h] * 0x1p-32, u2 = (double)c[2 * h + 1] * 0x1p-32;
        const double rad = sqrt(-2.0 * log(1.0 - u1));
        double sn, cs;
        sincos(6.283185307179586476925286766559 * u2, &sn, &cs);
        g[2 * h] = rad * cs;
        g[2 * h + 1] = rad * sn;
    }
}

// The same draw with the transcendental steps in FP32 (the FP32 panels of
// the mixed form): the uniforms and 1 - u1 are formed exactly as above in
// FP64, then rounded, so the tail of the radius is kept.
__device__ __forceinline__ void gauss_quad_f(uint64_t seed, uint64_t call, int row, int quad, float (&g)[4])
{
    uint32_t c[4] = {(uint32_t)row, (uint32_t)quad, (uint32_t)call, (uint32_t)(call >> 32)};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const double u1 = (double)c[2 * h] * 0x1p-32, u2 = (double)c[2 * h + 1] * 0x1p-32;
        const float rad = sqrtf(-2.0f * logf((float)(1.0 - u1)));
        float sn, cs;
        sincosf(6.2831853071795865f * (float)u2, &sn, &cs);
        g[2 * h] = rad * cs;
        g[2 * h + 1] = rad * sn;
    }
}

// Workgroups are dealt to the 8 XCDs round-robin; launches over 8*ceil(n/8)
// workgroups map XCD x to the contiguous items [x*per, (x+1)*per), so the
// items of one series (which read the same spectra / U, V chunks) share that
// XCD's L2.  Returns -1 for the padding workgroups.
constexpr int N_XCD = 8;
__device__ __forceinline__ int xcd_item(int n_items)
{
    const int per = (n_items + N_XCD - 1) / N_XCD;
    const int i = (int)(blockIdx.x % N_XCD) * per + (int)(blockIdx.x / N_XCD);
    return i < n_items ? i : -1;
}
static inline unsigned xcd_grid(int n_items) { return (unsigned)(((n_items + N_XCD - 1) / N_XCD) * N_XCD); }

template <typename T> __device__ __forceinline__ void block_sum2(double &a, double &b, double *red)
{
    // 256-thread sum of two values; result broadcast to every thread
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
    }
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) { red[2 * w] = a; red[2 * w + 1] = b; }
    __syncthreads();
    a = 0; b = 0;
    const int nw = blockDim.x >> 6;
    for (int i = 0; i < nw; i++) { a += red[2 * i]; b += red[2 * i + 1]; }
}

// ---------------------------------------------------------------------------
// spectra of the signal segments: X[o][sk] = FFT(x_m[a CL + sk CK : +CL+CK-1]) / n
// ---------------------------------------------------------------------------
template <typename T, int LOGN>
__global__ __launch_bounds__((FftGeom<T, LOGN>::NT), 4) void k_spectra(const ssa_desc *__restrict__ D,
                                                                const ssa_item *__restrict__ items,
                                                                T *__restrict__ bulk, const C2<T> *__restrict__ tw)
{
    using G = FftGeom<T, LOGN>;
    constexpr bool SPLIT = G::SPLIT;
    __shared__ T lre[G::LDS];
    __shared__ T lim[SPLIT ? 1 : G::LDS];
    const ssa_item it = items[blockIdx.x];
    const ssa_desc d = D[it.s];
    const int o = it.x, sk = it.y;
    const int m = o / d.AL, a = o - m * d.AL;
    const int start = a * d.CL + sk * d.CK;
    int len = d.CL + d.CK - 1;
    if (start + len > d.N) len = d.N - start;
    const T *x = bulk + d.off_x + (long long)m * d.N + start;
    C2<T> c[G::E];
#pragma unroll
    for (int e = 0; e < G::E; e++) {
        const int p = threadIdx.x + e * G::NT;
        c[e].x = (p < len) ? x[p] : (T)0;
        c[e].y = 0;
    }
    fft<T, LOGN, -1, SPLIT>(c, lre, lim, tw);
    const T sc = (T)1 / (T)G::N;
    C2<T> *X = reinterpret_cast<C2<T> *>(bulk + d.off_spec) + ((long long)o * d.SK + sk) * G::N;
#pragma unroll
    for (int e = 0; e < G::E; e++) {
        const int p = threadIdx.x + e * G::NT;
        X[p] = C2<T>{c[e].x * sc, c[e].y * sc};
    }
}

// ---------------------------------------------------------------------------
// Gaussian test matrix into panel Z (K x b): columns 4j .. 4j+3 of row r are
// the Philox draw (row r, quad j) of the series' key and call counter (the
// descriptor's, or call_arg when it is >= 0); FP32 panels use the FP32
// Box-Muller of the same uniforms.  Item (s, tile of SSA_GRAM_ROWS rows).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_omega(const ssa_desc *__restrict__ D, const ssa_item *__restrict__ items,
                                               T *bulk, long long call_arg)
{
    const ssa_item it = items[blockIdx.x];
    const ssa_desc d = D[it.s];
    const unsigned long long call = call_arg >= 0 ? (unsigned long long)call_arg : d.call;
    const Panel<T> P = panel_of<T>(bulk, d, SSA_PANEL_Z);
    const int r0 = it.x * SSA_GRAM_ROWS, r1 = min(P.rows, r0 + SSA_GRAM_ROWS);
    const int quads = (d.b + 3) / 4;
    // column quads dealt over gridDim.y workgroups (more than one only when
    // the launch holds few row tiles)
    for (int j = blockIdx.y; j < quads; j += gridDim.y) {
        T *c0 = P.p + (long long)(4 * j) * P.ld;
        const int nc = min(4, d.b - 4 * j);
        for (int r = r0 + (int)threadIdx.x; r < r1; r += 256) {
            T g[4];
            if constexpr (sizeof(T) == 4) gauss_quad_f(d.seed, call, r, j, g);
            else gauss_quad(d.seed, call, r, j, g);
#pragma unroll
            for (int t = 0; t < 4; t++)
                if (t < nc) c0[(long long)t * P.ld + r] = g[t];
        }
    }
}

// ---------------------------------------------------------------------------
// fused block matvec.  Item (s, pair, o0, o1): columns 2*pair, 2*pair+1 and
// output chunks [o0, o1).  dir 0: out rows are the stacked L dimension
// (chunk o = m*AL + a), input chunks run over K; dir 1: the reverse.
// out[o][i] = sum_p sum_j seg_(o,p)[i + j] in_p[j]  via  T+( X_(o,p) . T+(in_p) ).
// ---------------------------------------------------------------------------
template <typename T, int LOGN, bool W>
__global__ __launch_bounds__((FftGeom<T, LOGN, W>::NT), (FftGeom<T, LOGN, W>::OCC)) void k_matvec(const ssa_desc *__restrict__ D,
                                                                     const ssa_item *__restrict__ items, T *bulk,
                                                                     const C2<T> *__restrict__ tw, int dir,
                                                                     int in_panel, int out_panel, int n_items)
{
    using G = FftGeom<T, LOGN, W>;
    constexpr int E = G::E, NT = G::NT;
    constexpr bool SPLIT = G::SPLIT;
    __shared__ T lre[G::LDS];
    __shared__ T lim[SPLIT ? 1 : G::LDS];
    const int item = xcd_item(n_items);
    if (item < 0) return;
    const ssa_item it = items[item];
    const ssa_desc d = D[it.s];
    const int c0 = 2 * it.x, c1 = c0 + 1;
    const bool two = c1 < d.b;
    const Panel<T> pin = panel_of<T>(bulk, d, in_panel);
    const Panel<T> pout = panel_of<T>(bulk, d, out_panel);
    const int n_in = dir == 0 ? d.SK : d.M * d.AL;
    const C2<T> *X = reinterpret_cast<const C2<T> *>(bulk + d.off_spec);

    C2<T> z[E], acc[E];
    for (int o = it.y; o < it.z; o++) {
#pragma unroll
        for (int e = 0; e < E; e++) acc[e] = C2<T>{0, 0};
        for (int p = 0; p < n_in; p++) {
            if (n_in > 1 || o == it.y) {
                // input chunk p: rows [row0, row0 + len) of the K or stacked L dimension
                int row0, len;
                if (dir == 0) {
                    row0 = p * d.CK;
                    len = min(d.CK, d.K - row0);
                } else {
                    const int m = p / d.AL, a = p - m * d.AL;
                    row0 = m * d.L + a * d.CL;
                    len = min(d.CL, d.L - a * d.CL);
                }
                {
                    const T *a0 = pin.p + (long long)c0 * pin.ld + row0;
                    const T *a1 = pin.p + (long long)(two ? c1 : c0) * pin.ld + row0;
#pragma unroll
                    for (int e = 0; e < E; e++) {
                        const int q = threadIdx.x + e * NT;
                        z[e].x = (q < len) ? a0[q] : (T)0;
                        z[e].y = (two && q < len) ? a1[q] : (T)0;
                    }
                }
                fft<T, LOGN, 1, SPLIT, W>(z, lre, lim, tw);
            }
            // spectra are indexed [L chunk][K chunk]
            const int lo = dir == 0 ? o : p, ko = dir == 0 ? p : o;
            const C2<T> *Xo = X + ((long long)lo * d.SK + ko) * G::N;
            // all spectrum loads first, then the products: interleaved, the
            // scheduler reused one register pair and waited on every load
            C2<T> xv[E];
#pragma unroll
            for (int e = 0; e < E; e++) xv[e] = Xo[threadIdx.x + e * NT];
#pragma unroll
            for (int e = 0; e < E; e++) {
                acc[e].x += xv[e].x * z[e].x - xv[e].y * z[e].y;
                acc[e].y += xv[e].x * z[e].y + xv[e].y * z[e].x;
            }
        }
        fft<T, LOGN, 1, SPLIT, W>(acc, lre, lim, tw);
        int row0, len;
        if (dir == 0) {
            const int m = o / d.AL, a = o - m * d.AL;
            row0 = m * d.L + a * d.CL;
            len = min(d.CL, d.L - a * d.CL);
        } else {
            row0 = o * d.CK;
            len = min(d.CK, d.K - row0);
        }
        T *b0 = pout.p + (long long)c0 * pout.ld + row0;
        T *b1 = pout.p + (long long)c1 * pout.ld + row0;
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int q = threadIdx.x + e * NT;
            if (q < len) {
                b0[q] = acc[e].x;
                if (two) b1[q] = acc[e].y;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// staged block matvec, for series chunked in both dimensions (more than one
// input and more than one output chunk): k_matvec would transform every input
// chunk again for every output chunk (n_in * n_out + n_out transforms per
// column pair); here each input chunk is transformed once into the series'
// scratch (k_mv_spec), and each output chunk sums the stored spectra
// (k_mv_acc): n_in + n_out transforms.  The stored spectrum is the transform's
// register values, so the sums are formed from the same values in the same
// order as in k_matvec.  Scratch layout: spectrum of input chunk p of column
// pair c at off_mvs + 2 N (c n_in + p), N complex values in natural order.
// ---------------------------------------------------------------------------
template <typename T, int LOGN, bool W>
__global__ __launch_bounds__((FftGeom<T, LOGN, W>::NT), (FftGeom<T, LOGN, W>::OCC)) void k_mv_spec(
    const ssa_desc *__restrict__ D, const ssa_item *__restrict__ items, T *bulk, const C2<T> *__restrict__ tw,
    int dir, int in_panel, int n_items)
{
    using G = FftGeom<T, LOGN, W>;
    constexpr int E = G::E, NT = G::NT;
    constexpr bool SPLIT = G::SPLIT;
    __shared__ T lre[G::LDS];
    __shared__ T lim[SPLIT ? 1 : G::LDS];
    const int item = xcd_item(n_items);
    if (item < 0) return;
    const ssa_item it = items[item];   // (s, pair, input chunk p)
    const ssa_desc d = D[it.s];
    const int c0 = 2 * it.x, c1 = c0 + 1, p = it.y;
    const bool two = c1 < d.b;
    const Panel<T> pin = panel_of<T>(bulk, d, in_panel);
    const int n_in = dir == 0 ? d.SK : d.M * d.AL;
    int row0, len;
    if (dir == 0) {
        row0 = p * d.CK;
        len = min(d.CK, d.K - row0);
    } else {
        const int m = p / d.AL, a = p - m * d.AL;
        row0 = m * d.L + a * d.CL;
        len = min(d.CL, d.L - a * d.CL);
    }
    const T *a0 = pin.p + (long long)c0 * pin.ld + row0;
    const T *a1 = pin.p + (long long)(two ? c1 : c0) * pin.ld + row0;
    C2<T> z[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int q = threadIdx.x + e * NT;
        z[e].x = (q < len) ? a0[q] : (T)0;
        z[e].y = (two && q < len) ? a1[q] : (T)0;
    }
    fft<T, LOGN, 1, SPLIT, W>(z, lre, lim, tw);
    C2<T> *S = reinterpret_cast<C2<T> *>(bulk + d.off_mvs) + ((long long)it.x * n_in + p) * G::N;
#pragma unroll
    for (int e = 0; e < E; e++) S[threadIdx.x + e * NT] = z[e];
}

template <typename T, int LOGN, bool W>
__global__ __launch_bounds__((FftGeom<T, LOGN, W>::NT), (FftGeom<T, LOGN, W>::OCC)) void k_mv_acc(
    const ssa_desc *__restrict__ D, const ssa_item *__restrict__ items, T *bulk, const C2<T> *__restrict__ tw,
    int dir, int out_panel, int n_items)
{
    using G = FftGeom<T, LOGN, W>;
    constexpr int E = G::E, NT = G::NT;
    constexpr bool SPLIT = G::SPLIT;
    __shared__ T lre[G::LDS];
    __shared__ T lim[SPLIT ? 1 : G::LDS];
    const int item = xcd_item(n_items);
    if (item < 0) return;
    const ssa_item it = items[item];   // (s, pair, output chunk o)
    const ssa_desc d = D[it.s];
    const int c0 = 2 * it.x, c1 = c0 + 1, o = it.y;
    const bool two = c1 < d.b;
    const Panel<T> pout = panel_of<T>(bulk, d, out_panel);
    const int n_in = dir == 0 ? d.SK : d.M * d.AL;
    const C2<T> *X = reinterpret_cast<const C2<T> *>(bulk + d.off_spec);
    const C2<T> *S = reinterpret_cast<const C2<T> *>(bulk + d.off_mvs) + (long long)it.x * n_in * G::N;
    C2<T> acc[E];
#pragma unroll
    for (int e = 0; e < E; e++) acc[e] = C2<T>{0, 0};
    for (int p = 0; p < n_in; p++) {
        // spectra are indexed [L chunk][K chunk]
        const int lo = dir == 0 ? o : p, ko = dir == 0 ? p : o;
        const C2<T> *Xo = X + ((long long)lo * d.SK + ko) * G::N;
        const C2<T> *Sp = S + (long long)p * G::N;
        C2<T> xv[E], zv[E];
#pragma unroll
        for (int e = 0; e < E; e++) {
            xv[e] = Xo[threadIdx.x + e * NT];
            zv[e] = Sp[threadIdx.x + e * NT];
        }
#pragma unroll
        for (int e = 0; e < E; e++) {
            acc[e].x += xv[e].x * zv[e].x - xv[e].y * zv[e].y;
            acc[e].y += xv[e].x * zv[e].y + xv[e].y * zv[e].x;
        }
    }
    fft<T, LOGN, 1, SPLIT, W>(acc, lre, lim, tw);
    int row0, len;
    if (dir == 0) {
        const int m = o / d.AL, a = o - m * d.AL;
        row0 = m * d.L + a * d.CL;
        len = min(d.CL, d.L - a * d.CL);
    } else {
        row0 = o * d.CK;
        len = min(d.CK, d.K - row0);
    }
    T *b0 = pout.p + (long long)c0 * pout.ld + row0;
    T *b1 = pout.p + (long long)c1 * pout.ld + row0;
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int q = threadIdx.x + e * NT;
        if (q < len) {
            b0[q] = acc[e].x;
            if (two) b1[q] = acc[e].y;
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 MFMA 16x16x4 (v_mfma_f64_16x16x4f64).  Operand layout: lane l holds
// A[l%16][l/16] of the 16x4 A and B[l/16][l%16] of the 4x16 B; the 4
// accumulator values of lane l are D[l/16 + 4v][l%16], v = 0..3.
// ---------------------------------------------------------------------------
typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4v mfma64(double a, double b, d4v c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Panels are read straight from HBM into MFMA operand registers; only the
// small b x b matrix of an apply (or of the fused second CholeskyQR pass)
// goes through LDS.  Two register layouts are used:
//  - columns in lanes (Gram of P): lane (c = l%16, q = l/16) holds 4
//    consecutive rows 4q..4q+3 of column c of a 16-row slab; MFMA step v
//    uses row 4q+v, which permutes the summation index of the Gram
//    identically for both operands.
//  - rows in lanes (P M, and (P M)^T): lane (r = l%16, q = l/16) holds VR
//    consecutive rows VR*r..VR*r+VR-1 of column 4kk+q (one 8-byte load);
//    row block rb takes physical row VR*r + rb for MFMA row r.
// ---------------------------------------------------------------------------
typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
typedef double d2u __attribute__((ext_vector_type(2), aligned(8)));

// 4 consecutive elements p[0..3], those at or beyond n read as zero
template <typename T> __device__ __forceinline__ void load4(const T *p, int n, T (&o)[4])
{
    if (n >= 4) {
        if constexpr (sizeof(T) == 4) {
            const f4u v = *(const f4u *)p;
            o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
        } else {
            const d2u v0 = *(const d2u *)p, v1 = *(const d2u *)(p + 2);
            o[0] = v0.x; o[1] = v0.y; o[2] = v1.x; o[3] = v1.y;
        }
    } else {
#pragma unroll
        for (int i = 0; i < 4; i++) o[i] = (i < n) ? p[i] : (T)0;
    }
}

// VR = 8 / sizeof(T) consecutive elements (one 8-byte load), zero at or beyond n
template <typename T> struct RowVec { static constexpr int VR = 8 / (int)sizeof(T); };

template <typename T> __device__ __forceinline__ void load_vr(const T *p, int n, T (&o)[RowVec<T>::VR])
{
    if constexpr (sizeof(T) == 4) {
        if (n >= 2) {
            const f2u v = *(const f2u *)p;
            o[0] = v.x; o[1] = v.y;
        } else {
            o[0] = (n >= 1) ? p[0] : 0.0f;
            o[1] = 0.0f;
        }
    } else {
        o[0] = (n >= 1) ? p[0] : 0.0;
    }
}

template <typename T> __device__ __forceinline__ void store_vr(T *p, int n, const T (&o)[RowVec<T>::VR])
{
    if constexpr (sizeof(T) == 4) {
        if (n >= 2) {
            *(f2u *)p = f2u{o[0], o[1]};
        } else if (n >= 1) {
            p[0] = o[0];
        }
    } else {
        if (n >= 1) p[0] = o[0];
    }
}

// LDS row stride (doubles) of a staged small matrix with NB column blocks:
// == 16 (mod 32) so the 4 row groups of one MFMA operand read hit disjoint banks
template <int NB> constexpr int small_ld() { return (NB & 1) ? NB * 16 : NB * 16 + 16; }

// Stage the b x nc column-major small matrix M (ld SSA_BMAX) into Ms[k*MP + c]
// (as type S), zero beyond b rows / nc columns, NB*16 rows and columns, with
// NTHREADS threads.  Every load is issued before the first store (addresses
// clamped into the slot, values masked afterwards), so staging costs one
// memory latency instead of one per pass.
template <int NB, int NTHREADS, typename S = double>
__device__ __forceinline__ void stage_small(const double *M, int b, int nc, S *Ms)
{
    constexpr int B16 = NB * 16, MP = small_ld<NB>(), PER = (B16 * B16 + NTHREADS - 1) / NTHREADS;
    double v[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int idx = (int)threadIdx.x + q * NTHREADS;
        const int c = idx / B16, k = idx - c * B16;
        const bool in = idx < B16 * B16 && k < b && c < nc;
        const double x = M[in ? (long long)c * SSA_BMAX + k : 0];
        v[q] = in ? x : 0.0;
    }
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int idx = (int)threadIdx.x + q * NTHREADS;
        const int c = idx / B16, k = idx - c * B16;
        if (idx < B16 * B16) Ms[k * MP + c] = (S)v[q];
    }
}

template <int NB> constexpr int n_upper_blocks() { return NB * (NB + 1) / 2; }

// ---------------------------------------------------------------------------
// partial Gram: item (s, tile) covers rows [tile*SSA_GRAM_ROWS, +SSA_GRAM_ROWS)
// of the panel, dealt to the 4 waves in slabs.  G_part = T^T T with T = P
// (USE_M = 0, columns-in-lanes layout) or T = P M, M = R1^-1 upper triangular
// (USE_M = 1, one row per lane: the accumulator layout of P M is the
// operand layout of the Gram, so T never leaves registers).  FP64 MFMA, FP64
// accumulation whatever the panel type; only the upper block triangle is
// multiplied, the partial is written in full.  USE_M = 0 also emits the column
// sums of P.  The 4 waves' sums are added in a fixed order (deterministic).
// NB = column blocks of 16 covering every panel of the launch.  SPLIT: the
// gridDim.y workgroups of an item share its upper blocks (block m to workgroup
// m % gridDim.y, which loads only the column blocks those use; the sums of
// column block I from the owner of block (I, I)) for launches of few items;
// each block is summed in the same order either way, so the partial is the
// same.
// ACC32 (FP32 panels, USE_M = 0 only): FP32 MFMA (v_mfma_f32_16x16x4f32) with
// FP32 accumulation inside the tile; the 4 waves' sums and the tiles are then
// added in FP64.  Only the W-correlation of the mixed form takes it: its
// entries are normalised inner products, bounded by 1, so the FP32 sum's
// relative error (about eps32 * sqrt(rows)) is far below what that output
// needs, unlike the orthonormalisation Grams, whose smallest eigenvalues the
// Cholesky needs to eps * kappa^2.
// PAD (Y and Z panels, USE_M = 0): the columns are padded to a multiple of 32
// rows, so a lane's 4 rows r .. r + 3 exist whenever r is inside the panel;
// every lane loads unconditionally (lanes outside the panel or its blocks read
// row r_begin of column 0 and discard it).  Without branches around them the
// next slab's loads stay in flight while the current slab's products run;
// with them the compiler waits for all loads before the first product.  The
// values, and so the sums, are the same either way.
// ---------------------------------------------------------------------------
template <typename T, int NB, int USE_M, bool SPLIT, bool ACC32 = false, bool PAD = false>
__global__ __launch_bounds__(256) void k_gram(const ssa_desc *__restrict__ D, const ssa_item *__restrict__ items,
                                              const T *__restrict__ bulk, double *small, int which)
{
    static_assert(!ACC32 || (sizeof(T) == 4 && !USE_M), "FP32 accumulation is for FP32 panels without M");
    static_assert(!PAD || !USE_M, "padded loads are for the P^T P form");
    extern __shared__ double sm[];
    constexpr int NBLK = n_upper_blocks<NB>(), MP = small_ld<NB>();
    const ssa_item it = items[blockIdx.x];
    const ssa_desc d = D[it.s];
    const Panel<T> P = panel_of<T>(const_cast<T *>(bulk), d, which);
    const int b = P.cols;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, cl = l & 15, kq = l >> 4;
    const int r_begin = it.x * SSA_GRAM_ROWS;
    const int r_end = min(P.rows, r_begin + SSA_GRAM_ROWS);
    d4v acc[ACC32 ? 1 : NBLK];
    f4v accf[ACC32 ? NBLK : 1];
#pragma unroll
    for (int m = 0; m < (ACC32 ? 1 : NBLK); m++) acc[m] = d4v{0, 0, 0, 0};
#pragma unroll
    for (int m = 0; m < (ACC32 ? NBLK : 1); m++) accf[m] = f4v{0, 0, 0, 0};
    double cs[NB];
#pragma unroll
    for (int I = 0; I < NB; I++) cs[I] = 0.0;

    // column blocks this workgroup's upper blocks use (all without SPLIT)
    unsigned need = ~0u;
    if constexpr (SPLIT) {
        need = 0;
        int m = 0;
#pragma unroll
        for (int I = 0; I < NB; I++)
#pragma unroll
            for (int J = I; J < NB; J++, m++)
                if (m % (int)gridDim.y == (int)blockIdx.y) need |= (1u << I) | (1u << J);
    }
    if constexpr (!USE_M) {
        // 16-row slabs, slab j of the item to wave j % 4
        T cur[NB][4], nxt[NB][4];
        auto load = [&](int r0, T (&x)[NB][4]) {
#pragma unroll
            for (int I = 0; I < NB; I++) {
                const int c = I * 16 + cl, r = r0 + 4 * kq;
                if constexpr (PAD) {
                    const bool okc = ((need >> I) & 1u) && c < b, okr = r < r_end;
                    T v4[4];
                    load4(P.p + (long long)(okc ? c : 0) * P.ld + (okr ? r : r_begin), 4, v4);
#pragma unroll
                    for (int v = 0; v < 4; v++) x[I][v] = (okc && r + v < r_end) ? v4[v] : (T)0;
                } else if (((need >> I) & 1u) && c < b && r0 < r_end) {
                    load4(P.p + (long long)c * P.ld + r, r_end - r, x[I]);
                } else {
#pragma unroll
                    for (int v = 0; v < 4; v++) x[I][v] = (T)0;
                }
            }
        };
        int r0 = r_begin + 16 * w;
        load(r0, cur);
#pragma unroll 1
        for (; r0 < r_end; r0 += 64) {
            load(r0 + 64, nxt);
#pragma unroll
            for (int v = 0; v < 4; v++) {
#pragma unroll
                for (int I = 0; I < NB; I++) cs[I] += (double)cur[I][v];
                int m = 0;
#pragma unroll
                for (int I = 0; I < NB; I++)
#pragma unroll
                    for (int J = I; J < NB; J++, m++)
                        if (!SPLIT || m % (int)gridDim.y == (int)blockIdx.y) {
                            if constexpr (ACC32)
                                accf[m] = __builtin_amdgcn_mfma_f32_16x16x4f32((float)cur[I][v], (float)cur[J][v],
                                                                               accf[m], 0, 0, 0);
                            else
                                acc[m] = mfma64((double)cur[I][v], (double)cur[J][v], acc[m]);
                        }
            }
#pragma unroll
            for (int I = 0; I < NB; I++)
#pragma unroll
                for (int v = 0; v < 4; v++) cur[I][v] = nxt[I][v];
        }
    } else {
        // one row per lane (16-row slabs): P M for a slab stays in registers
        constexpr int KK = NB * 4;
        double *Ms = sm;
        stage_small<NB, 256>(slot(small, d, SSA_SLOT_R1INV), b, b, Ms);
        __syncthreads();
        T cur[KK], nxt[KK];
        auto load = [&](int r0, T (&a)[KK]) {
            const int r = r0 + cl;
#pragma unroll
            for (int kk = 0; kk < KK; kk++) {
                const int c = 4 * kk + kq;
                a[kk] = (c < b && r < r_end) ? P.p[(long long)c * P.ld + r] : (T)0;
            }
        };
        int r0 = r_begin + 16 * w;
        load(r0, cur);
        for (; r0 < r_end; r0 += 64) {
            load(r0 + 64, nxt);
            d4v t[NB];
#pragma unroll
            for (int cb = 0; cb < NB; cb++) {
                t[cb] = d4v{0, 0, 0, 0};
                if (!((need >> cb) & 1u)) continue;   // column block used by none of this workgroup's blocks
                // M upper triangular: rows k < 16 (cb + 1) only
#pragma unroll
                for (int kk = 0; kk < 4 * (cb + 1); kk++)
                    t[cb] = mfma64((double)cur[kk], Ms[(4 * kk + kq) * MP + cb * 16 + cl], t[cb]);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                int m = 0;
#pragma unroll
                for (int I = 0; I < NB; I++)
#pragma unroll
                    for (int J = I; J < NB; J++, m++)
                        if (!SPLIT || m % (int)gridDim.y == (int)blockIdx.y) acc[m] = mfma64(t[I][v], t[J][v], acc[m]);
            }
#pragma unroll
            for (int kk = 0; kk < KK; kk++) cur[kk] = nxt[kk];
        }
    }
    // column sums: lanes of one column (kq = 0..3) first, then the waves
#pragma unroll
    for (int I = 0; I < NB; I++) {
        cs[I] += __shfl_xor(cs[I], 16);
        cs[I] += __shfl_xor(cs[I], 32);
    }
    __syncthreads();   // LDS (small matrix) free from here
    double *red = sm;  // NBLK x 4 x 64 accumulator values, then NB x 16 column sums
    double *rcs = sm + NBLK * 256;
    for (int ww = 0; ww < 4; ww++) {
        if (w == ww) {
#pragma unroll
            for (int m = 0; m < NBLK; m++)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    double *p = red + (m * 4 + v) * 64 + l;
                    double a;
                    if constexpr (ACC32) a = (double)accf[m][v];
                    else a = acc[m][v];
                    *p = (ww == 0 ? 0.0 : *p) + a;
                }
            if (kq == 0) {
#pragma unroll
                for (int I = 0; I < NB; I++) {
                    double *p = rcs + I * 16 + cl;
                    *p = (ww == 0 ? 0.0 : *p) + cs[I];
                }
            }
        }
        __syncthreads();
    }
    double *out = small + d.off_part + (long long)it.x * (b * b + b);
    for (int idx = threadIdx.x; idx < NBLK * 256; idx += 256) {
        const int m = idx >> 8, v = (idx >> 6) & 3, ll = idx & 63;
        if (SPLIT && m % (int)gridDim.y != (int)blockIdx.y) continue;
        int I = 0, mm = m;
        while (mm >= NB - I) { mm -= NB - I; I++; }
        const int J = I + mm;
        // accumulator value v of lane ll: D[ll/16 + 4v][ll%16] (FP64 MFMA),
        // D[4 (ll/16) + v][ll%16] (FP32 MFMA)
        const int i = I * 16 + (ACC32 ? 4 * (ll >> 4) + v : (ll >> 4) + 4 * v), j = J * 16 + (ll & 15);
        if (i < b && j < b) {
            const double g = red[idx];
            out[i * b + j] = g;
            if (I != J) out[j * b + i] = g;
        }
    }
    // the sums of column block I come from the workgroup of diagonal block (I, I)
    const int Ic = threadIdx.x >> 4, mdiag = Ic * NB - Ic * (Ic - 1) / 2;
    if (!USE_M && threadIdx.x < b && (!SPLIT || mdiag % (int)gridDim.y == (int)blockIdx.y))
        out[b * b + threadIdx.x] = rcs[threadIdx.x];
}

// ---------------------------------------------------------------------------
// Cholesky in one wave per series, the b x b matrix in registers
// (b <= NB*16 <= 64): lane j holds column j, slot i = A[i][j].  The matrix is
// padded to NB*16 with an identity block, so the inner loops run over the
// full width with static slot indices and no per-element branches; steps
// k >= b are skipped by one uniform branch each.  Each step writes its row of R into an LDS
// triangle (column j of R at Rl[tri_off(j) + i], i <= j); the trailing
// update and the inverse read R from there as uniform (broadcast) LDS reads.
// Same definition as k_chol pass 1: the Gram partials summed in tile order,
// pivots below tol * (original diagonal) mark dead columns (zero row of R,
// zero row and column of R^-1).  R^-1 is formed by rows (X R = I): lane i
// computes row i of X.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double rdlane(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

__host__ __device__ constexpr int tri_off(int j) { return j * (j + 1) / 2; }

// 1/x and 1/sqrt(x) for normal, finite x (x > 0 for the square root): the
// hardware estimate refined by two Newton steps, each of which squares the
// relative error.  They replace IEEE division and square root on the serial
// chains of the Cholesky and Jacobi steps; the results may differ from the
// correctly rounded ones in the last bit.
__device__ __forceinline__ double rcp_nr(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

__device__ __forceinline__ double rsq_nr(double x)
{
    const double hx = 0.5 * x;
    double y = __builtin_amdgcn_rsq(x);
    y = y * fma(-hx * y, y, 1.5);
    return y * fma(-hx * y, y, 1.5);
}

// v summed over each aligned group of 8 lanes, every lane of the group
// receiving the same bits: DPP quad permutations (lane ^ 1, lane ^ 2) then the
// half-row mirror (lane i of 8 with lane 7 - i).  All 8 lanes must be active.
template <int CTRL> __device__ __forceinline__ double dpp64(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double sum8(double v)
{
    v += dpp64<0xB1>(v);    // quad_perm [1, 0, 3, 2]
    v += dpp64<0x4E>(v);    // quad_perm [2, 3, 0, 1]
    return v + dpp64<0x141>(v);   // row_half_mirror
}

// Sum the Gram partials of a series in tile order: lane j < b gets column j
// (slots i < b, zero above), its diagonal dg and its column sum cy; lanes
// j >= b get column j of the identity (dg = 1).  Reads are clamped into the
// b x b partial and masked, so no slot index depends on b.
template <int NB>
__device__ __forceinline__ void gram_sum_regs(const double *part, int tiles, int b, int ln, double (&a)[NB * 16],
                                              double &dg, double &cy)
{
    constexpr int B16 = NB * 16;
    const int stride = b * b + b;
    const bool live = ln < b;
    const int lc = live ? ln : b - 1;
#pragma unroll
    for (int i = 0; i < B16; i++) a[i] = 0.0;
    dg = 0.0;
    cy = 0.0;
    // 16 slots per sweep over the tiles (bounded loads in flight)
#pragma unroll
    for (int h = 0; h < NB; h++) {
        if (16 * h >= b) break;
        for (int t = 0; t < tiles; t++) {
            const double *pt = part + (long long)t * stride;
            // the 16 loads first, then the sums (interleaved, the scheduler
            // reused one register pair and waited on every load)
            double v[16];
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const int ii = 16 * h + i;
                v[i] = pt[(ii < b ? ii : b - 1) * b + lc];
            }
#pragma unroll
            for (int i = 16 * h; i < 16 * h + 16; i++) a[i] += (i < b) ? v[i - 16 * h] : 0.0;
            if (h == 0) {
                dg += pt[lc * b + lc];
                cy += pt[b * b + lc];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < B16; i++)
        if (!live) a[i] = (i == ln) ? 1.0 : 0.0;
    if (!live) {
        dg = 1.0;
        cy = 0.0;
    }
}

// Cholesky A = R^T R in place (lane j: column j of R in slots 0..j), row k of
// R also stored to Rl.  Constant trip counts, static slot indices.
template <int NB>
__device__ __forceinline__ void chol_regs(double (&a)[NB * 16], double dg, int b, int ln, double tol, double *Rl,
                                          int &ndead, bool &mydead)
{
    constexpr int B16 = NB * 16;
    ndead = 0;
    mydead = true;
#pragma unroll
    for (int k = 0; k < B16; k++) {
        if (k >= b) continue;   // one uniform branch per step (it also bounds the scheduling region)
        const double p = rdlane(a[k], k), d0 = rdlane(dg, k);
        const bool dead = !(p > tol * d0) || !(d0 > 0.0);
        const double rp = dead ? 0.0 : rsq_nr(p), piv = p * rp;
        ndead += dead;
        if (ln == k) mydead = dead;
        if (ln >= k) a[k] = dead ? 0.0 : (ln == k ? piv : a[k] * rp);
        if (ln >= k && ln < B16) Rl[tri_off(ln) + k] = a[k];   // lanes >= B16 hold no column
        __syncthreads();
#pragma unroll
        for (int i = k + 1; i < B16; i++) {
            a[i] = fma(-Rl[tri_off(i) + k], a[k], a[i]);
            // at most 16 broadcast values in flight (register pressure)
            if ((i & 15) == 15) asm volatile("" ::: "memory");
        }
    }
}

// X = R^-1 by rows (X R = I) from the R in Rl: afterwards lane i holds row i
// of X (slots j >= i, zero below).
template <int NB>
__device__ __forceinline__ void inv_rows(double (&a)[NB * 16], int b, int ln, bool mydead, const double *Rl)
{
    constexpr int B16 = NB * 16;
#pragma unroll
    for (int j = 0; j < B16; j++) {
        if (j >= b) continue;
        const double rjj = Rl[tri_off(j) + j];
        const bool deadj = !(rjj > 0.0);
        const double ir = deadj ? 0.0 : rcp_nr(rjj);
        double s = 0.0;
#pragma unroll
        for (int l = 0; l < j; l++) {
            s = fma(-a[l], Rl[tri_off(j) + l], s);
            if ((l & 15) == 15) asm volatile("" ::: "memory");
        }
        if (ln <= j) a[j] = (deadj || mydead) ? 0.0 : (ln == j ? ir : s * ir);
        if (ln == j) {
#pragma unroll
            for (int l = 0; l < j; l++) a[l] = 0.0;
        }
        // R is read-only here: keep the scheduler from hoisting later steps' loads
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Stage the b x b column-major slot matrix M (ld SSA_BMAX) into Ms[l*MP + c]
// = M[l][c] (zero beyond b), 64 threads.
template <int NB> __device__ __forceinline__ void stage_rows(const double *M, int b, int ln, double *Ms)
{
    constexpr int B16 = NB * 16, MP = B16 + 1;
    for (int c = 0; c < B16; c++)
        if (ln < B16) Ms[ln * MP + c] = (ln < b && c < b) ? M[(long long)c * SSA_BMAX + ln] : 0.0;
}

// One wave per series: R = chol(G) and R^-1 of the summed Gram into the
// pass's slots (pass 1: R1, R1INV, colsum(Y) for Y, the dead-column count;
// pass 2: R2, R2INV).  Dead pivots as in k_chol.  Rows and columns b..NB*16-1
// of both slots are written as zeros.
template <int NB>
__global__ __launch_bounds__(64) void k_chol_w(const ssa_desc *D, const ssa_item *items, double *small, int which,
                                               int pass, double tol)
{
    constexpr int B16 = NB * 16;
    __shared__ double Rl[tri_off(B16)];
    const ssa_item it = items[blockIdx.x];
    const ssa_desc d = D[it.s];
    const int b = d.b, ln = threadIdx.x;
    const int tiles = which == SSA_PANEL_Y ? d.tilesY : d.tilesZ;
    double a[B16];
    double dg, cy;
    gram_sum_regs<NB>(small + d.off_part, tiles, b, ln, a, dg, cy);
    double *misc = slot(small, d, SSA_SLOT_MISC);
    if (pass == 1 && which == SSA_PANEL_Y && ln < b) misc[SSA_MISC_COLSUMY + ln] = cy;
    int ndead;
    bool mydead;
    chol_regs<NB>(a, dg, b, ln, tol, Rl, ndead, mydead);
    __syncthreads();
    double *Ro = slot(small, d, pass == 1 ? SSA_SLOT_R1 : SSA_SLOT_R2);
    double *Rio = slot(small, d, pass == 1 ? SSA_SLOT_R1INV : SSA_SLOT_R2INV);
    // column j of R from the triangle: lanes over rows, coalesced
    if (ln < B16) {
        for (int j = 0; j < B16; j++)
            Ro[(long long)j * SSA_BMAX + ln] = (ln <= j && j < b) ? Rl[tri_off(j) + ln] : 0.0;
    }
    inv_rows<NB>(a, b, ln, mydead, Rl);
    if (ln < B16) {
#pragma unroll
        for (int j = 0; j < B16; j++) Rio[(long long)j * SSA_BMAX + ln] = (ln < b && j < b) ? a[j] : 0.0;
    }
    // dead padding columns cannot occur (identity pivots), so ndead counts real columns only
    if (pass == 1 && ln == 0) misc[SSA_MISC_DEAD] = ndead;
}

// ---------------------------------------------------------------------------
// C = A B for upper-triangular b x b slot matrices (column-major, ld
// SSA_BMAX), FP64 MFMA.  One workgroup (4 waves) per series; block (I, J) of
// C needs k in [16 I, 16 J + 16) only; blocks below the diagonal are written
// as zeros so C is a clean upper-triangular matrix.
// ---------------------------------------------------------------------------
template <int NB>
__global__ __launch_bounds__(256) void k_trmm(const ssa_desc *D, const ssa_item *items, double *small, int sa,
                                              int sb, int sc)
{
    const ssa_item it = items[blockIdx.x];
    const ssa_desc d = D[it.s];
    const int b = d.b;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, cl = l & 15, kq = l >> 4;
    const double *A = small + d.off_small + (long long)sa * SSA_BMAX * SSA_BMAX;
    const double *B = small + d.off_small + (long long)sb * SSA_BMAX * SSA_BMAX;
    double *C = small + d.off_small + (long long)sc * SSA_BMAX * SSA_BMAX;
    for (int m = w; m < NB * NB; m += 4) {
        const int I = m / NB, J = m - I * NB;
        d4v acc = {0, 0, 0, 0};
        if (I <= J) {
            // every k step's operands loaded before the first product (steps
            // beyond 16 J + 16 carry zeros and add exact zeros)
            constexpr int KS = 4 * NB;
            double av[KS], bv[KS];
#pragma unroll
            for (int t = 0; t < KS; t++) {
                const int k0 = 16 * I + 4 * t;
                const bool in = k0 < 16 * J + 16;
                const int ra = 16 * I + cl, ka = k0 + kq;     // A[ra][ka]
                const int kb = k0 + kq, cb = 16 * J + cl;     // B[kb][cb]
                const bool ia = in && ra < b && ka < b, ib = in && kb < b && cb < b;
                const double x = A[ia ? (long long)ka * SSA_BMAX + ra : 0];
                const double y = B[ib ? (long long)cb * SSA_BMAX + kb : 0];
                av[t] = ia ? x : 0.0;
                bv[t] = ib ? y : 0.0;
            }
#pragma unroll
            for (int t = 0; t < KS; t++) acc = mfma64(av[t], bv[t], acc);
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = 16 * I + kq + 4 * v, j = 16 * J + cl;
            if (i < b && j < b) C[(long long)j * SSA_BMAX + i] = acc[v];
        }
    }
}

// ---------------------------------------------------------------------------
// Cholesky of the summed Gram (one workgroup per series; widths above 64).
// pass 1: R1 = chol(G), R1^-1 -> slots; panel Y also stores its column sums.
// pass 2: R2 = chol(G2), R2^-1 -> slots.
// A pivot below tol * (original diagonal) marks a numerically dependent
// column: its row of R and column of R^-1 are zero (the column of Q is zero).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_chol(const ssa_desc *D, const ssa_item *items, double *small, int which,
                                              int pass, double tol)
{
    extern __shared__ double sm[];
    const int b = D[items[blockIdx.x].s].b;
    const int LDP = b + 1;             // padded row length
    double *A = sm;                    // b x LDP, row-major (upper part used)
    double *X = sm + b * LDP;          // b x LDP
    __shared__ double diag0[SSA_BMAX];
    __shared__ int dead[SSA_BMAX];
    __shared__ double piv;
    const ssa_item it = items[blockIdx.x];
    const ssa_desc d = D[it.s];
    const int tiles = which == SSA_PANEL_Y ? d.tilesY : d.tilesZ;
    const double *part = small + d.off_part;
    const int stride = b * b + b;
    // sum partials (fixed order -> deterministic)
    for (int idx = threadIdx.x; idx < b * b; idx += 256) {
        double s = 0;
        for (int t = 0; t < tiles; t++) s += part[(long long)t * stride + idx];
        const int i = idx / b, j = idx - i * b;
        A[i * LDP + j] = s;
    }
    if (pass == 1 && which == SSA_PANEL_Y && threadIdx.x < b) {
        double s = 0;
        for (int t = 0; t < tiles; t++) s += part[(long long)t * stride + b * b + threadIdx.x];
        slot(small, d, SSA_SLOT_MISC)[SSA_MISC_COLSUMY + threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x < b) diag0[threadIdx.x] = A[threadIdx.x * LDP + threadIdx.x];
    __syncthreads();
    for (int j = 0; j < b; j++) {
        if (threadIdx.x == 0) {
            const double p = A[j * LDP + j];
            const int dd = !(p > tol * diag0[j]) || !(diag0[j] > 0.0);
            dead[j] = dd;
            piv = dd ? 0.0 : sqrt(p);
        }
        __syncthreads();
        const double r = piv;
        for (int l = j + threadIdx.x; l < b; l += 256)
            A[j * LDP + l] = (r > 0.0) ? ((l == j) ? r : A[j * LDP + l] / r) : 0.0;
        __syncthreads();
        const int m = b - j - 1;
        for (int idx = threadIdx.x; idx < m * m; idx += 256) {
            const int i = j + 1 + idx / m, l = j + 1 + idx % m;
            if (l >= i) A[i * LDP + l] -= A[j * LDP + i] * A[j * LDP + l];
        }
        __syncthreads();
    }
    // zero the strict lower triangle
    for (int idx = threadIdx.x; idx < b * b; idx += 256) {
        const int i = idx / b, j = idx - i * b;
        if (j < i) A[i * LDP + j] = 0.0;
    }
    __syncthreads();
    // X = R^-1 (upper), column by column
    if (threadIdx.x < b) {
        const int j = threadIdx.x;
        for (int i = b - 1; i >= 0; i--) X[i * LDP + j] = 0.0;
        if (!dead[j]) {
            for (int i = j; i >= 0; i--) {
                if (dead[i]) { X[i * LDP + j] = 0.0; continue; }
                double s = (i == j) ? 1.0 : 0.0;
                for (int l = i + 1; l <= j; l++) s -= A[i * LDP + l] * X[l * LDP + j];
                X[i * LDP + j] = s / A[i * LDP + i];
            }
        }
    }
    __syncthreads();
    double *misc = slot(small, d, SSA_SLOT_MISC);
    double *Ro = slot(small, d, pass == 1 ? SSA_SLOT_R1 : SSA_SLOT_R2);
    double *Rio = slot(small, d, pass == 1 ? SSA_SLOT_R1INV : SSA_SLOT_R2INV);
    for (int idx = threadIdx.x; idx < b * b; idx += 256) {
        const int j = idx / b, i = idx - j * b;   // column-major store
        Ro[(long long)j * SSA_BMAX + i] = A[i * LDP + j];
        Rio[(long long)j * SSA_BMAX + i] = X[i * LDP + j];
    }
    if (pass == 1 && threadIdx.x == 0) {
        int nd = 0;
        for (int j = 0; j < b; j++) nd += dead[j];
        misc[SSA_MISC_DEAD] = nd;
    }
}

// ---------------------------------------------------------------------------
// out (rows x nc) = in (rows x b) * M (b x nc, col-major ld BMAX) as
// out^T = M^T in^T: FP64 MFMA with FP64 accumulation, T-typed panels, rows in
// lanes so that each store is 16 consecutive rows of a column.  item (s, tile
// of SSA_GRAM_ROWS rows), 8 waves sharing the staged M; in place allowed (a
// wave loads all b columns of its rows before it stores them).  ncols_k: 0 ->
// nc = b and M upper triangular (R^-1), 1 -> nc = k, M full.
// ---------------------------------------------------------------------------
constexpr int APPLY_WAVES = 8;

template <typename T, int NB>
__global__ __launch_bounds__(64 * APPLY_WAVES) void k_apply(const ssa_desc *D, const ssa_item *items, T *bulk,
                                                            const double *small, int src, int dst, int mslot,
                                                            int ncols_k)
{
    extern __shared__ double sm[];
    constexpr int VR = RowVec<T>::VR, SL = 16 * VR, KK = NB * 4, MP = small_ld<NB>();
    const ssa_item it = items[blockIdx.x];
    const ssa_desc d = D[it.s];
    const Panel<T> Pi = panel_of<T>(bulk, d, src);
    const Panel<T> Po = panel_of<T>(bulk, d, dst);
    const int b = d.b, nc = ncols_k ? d.k : d.b;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, cl = l & 15, kq = l >> 4;
    // gridDim.y (a power of two) workgroups share the tile in equal row ranges;
    // every output row is formed the same way whatever its range
    const int sub = SSA_GRAM_ROWS / (int)gridDim.y;
    const int r_begin = it.x * SSA_GRAM_ROWS + (int)blockIdx.y * sub;
    const int r_end = min(Pi.rows, r_begin + sub);
    T cur[KK][VR], nxt[KK][VR];
    auto load = [&](int r0, T (&a)[KK][VR]) {
#pragma unroll
        for (int kk = 0; kk < KK; kk++) {
            const int c = 4 * kk + kq, r = r0 + VR * cl;
            if (c < b && r < r_end) load_vr(Pi.p + (long long)c * Pi.ld + r, r_end - r, a[kk]);
            else {
#pragma unroll
                for (int v = 0; v < VR; v++) a[kk][v] = (T)0;
            }
        }
    };
    int r0 = r_begin + SL * w;
    load(r0, cur);       // the first rows are in flight while M is staged
    double *Ms = sm;
    stage_small<NB, 64 * APPLY_WAVES>(small + d.off_small + (long long)mslot * SSA_BMAX * SSA_BMAX, b, nc, Ms);
    __syncthreads();
    for (; r0 < r_end; r0 += APPLY_WAVES * SL) {
        load(r0 + APPLY_WAVES * SL, nxt);
        const int r = r0 + VR * cl;
#pragma unroll
        for (int cb = 0; cb < NB; cb++) {
            if (cb * 16 >= nc) break;
            d4v t[VR];
#pragma unroll
            for (int rb = 0; rb < VR; rb++) t[rb] = d4v{0, 0, 0, 0};
#pragma unroll
            for (int kk = 0; kk < KK; kk++) {
                // an upper-triangular M has no rows below the column block's last column
                if (!ncols_k && kk >= 4 * (cb + 1)) break;
                const double mv = Ms[(4 * kk + kq) * MP + cb * 16 + cl];
#pragma unroll
                for (int rb = 0; rb < VR; rb++) t[rb] = mfma64(mv, (double)cur[kk][rb], t[rb]);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int c = cb * 16 + kq + 4 * v;
                if (c < nc && r < r_end) {
                    T o[VR];
#pragma unroll
                    for (int rb = 0; rb < VR; rb++) o[rb] = (T)t[rb][v];
                    store_vr(Po.p + (long long)c * Po.ld + r, r_end - r, o);
                }
            }
        }
#pragma unroll
        for (int kk = 0; kk < KK; kk++)
#pragma unroll
            for (int v = 0; v < VR; v++) cur[kk][v] = nxt[kk][v];
    }
}

// ---------------------------------------------------------------------------
// FP32 panels, FP32 MFMA (v_mfma_f32_16x16x4f32) with FP32 accumulation: the
// mixed form's apply.  Same item, wave and row-in-lane structure as k_apply.
// M (FP64 in the small area) is rounded to FP32 when staged; the panel is
// FP32 already, so each product is exact in the FP32 MFMA and the only new
// rounding is the accumulation over b terms, of the same order as the FP32
// rounding the input panel already carries.  Operand layout: lane l holds
// A[l%16][l/16] and B[l/16][l%16]; accumulator value v of lane l is
// D[4*(l/16) + v][l%16].
// ---------------------------------------------------------------------------
template <int NB>
__global__ __launch_bounds__(64 * APPLY_WAVES) void k_apply_f32(const ssa_desc *D, const ssa_item *items,
                                                                float *bulk, const double *small, int src,
                                                                int dst, int mslot, int ncols_k)
{
    extern __shared__ float smf[];
    constexpr int VR = 2, SL = 16 * VR, KK = NB * 4, MP = small_ld<NB>(), B16 = NB * 16;
    const ssa_item it = items[blockIdx.x];
    const ssa_desc d = D[it.s];
    const Panel<float> Pi = panel_of<float>(bulk, d, src);
    const Panel<float> Po = panel_of<float>(bulk, d, dst);
    const int b = d.b, nc = ncols_k ? d.k : d.b;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63, cl = l & 15, kq = l >> 4;
    const int sub = SSA_GRAM_ROWS / (int)gridDim.y;     // as in k_apply
    const int r_begin = it.x * SSA_GRAM_ROWS + (int)blockIdx.y * sub;
    const int r_end = min(Pi.rows, r_begin + sub);
    float cur[KK][VR], nxt[KK][VR];
    auto load = [&](int r0, float (&a)[KK][VR]) {
#pragma unroll
        for (int kk = 0; kk < KK; kk++) {
            const int c = 4 * kk + kq, r = r0 + VR * cl;
            if (c < b && r < r_end) load_vr(Pi.p + (long long)c * Pi.ld + r, r_end - r, a[kk]);
            else {
#pragma unroll
                for (int v = 0; v < VR; v++) a[kk][v] = 0.0f;
            }
        }
    };
    int r0 = r_begin + SL * w;
    load(r0, cur);       // the first rows are in flight while M is staged
    float *Ms = smf;
    stage_small<NB, 64 * APPLY_WAVES, float>(small + d.off_small + (long long)mslot * SSA_BMAX * SSA_BMAX, b, nc, Ms);
    __syncthreads();
    for (; r0 < r_end; r0 += APPLY_WAVES * SL) {
        load(r0 + APPLY_WAVES * SL, nxt);
        const int r = r0 + VR * cl;
#pragma unroll
        for (int cb = 0; cb < NB; cb++) {
            if (cb * 16 >= nc) break;
            f4v t[VR];
#pragma unroll
            for (int rb = 0; rb < VR; rb++) t[rb] = f4v{0, 0, 0, 0};
#pragma unroll
            for (int kk = 0; kk < KK; kk++) {
                // an upper-triangular M has no rows below the column block's last column
                if (!ncols_k && kk >= 4 * (cb + 1)) break;
                const float mv = Ms[(4 * kk + kq) * MP + cb * 16 + cl];
#pragma unroll
                for (int rb = 0; rb < VR; rb++)
                    t[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(mv, cur[kk][rb], t[rb], 0, 0, 0);
            }
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int c = cb * 16 + 4 * kq + v;
                if (c < nc && r < r_end) {
                    float o[VR];
#pragma unroll
                    for (int rb = 0; rb < VR; rb++) o[rb] = t[rb][v];
                    store_vr(Po.p + (long long)c * Po.ld + r, r_end - r, o);
                }
            }
        }
#pragma unroll
        for (int kk = 0; kk < KK; kk++)
#pragma unroll
            for (int v = 0; v < VR; v++) cur[kk][v] = nxt[kk][v];
    }
}

// ---------------------------------------------------------------------------
// SVD of R (b x b, FP64) by one-sided Jacobi on R^T with a round-robin
// parallel order: R^T V = W S, so the converged columns of R^T, normalised,
// are the right singular vectors W of R (R W = U S) and no rotation has to be
// accumulated.  Sorted descending; signs chosen so that colsum(Q) . W_i >= 0,
// i.e. sum(U[:,i]) >= 0 for U = Q W.
// Outputs: WK = W[:, :k] (signed), WKS = WK S^-1, sigma, lambda, total variance.
// ---------------------------------------------------------------------------
template <int NB>
__global__ __launch_bounds__(256) void k_svd(const ssa_desc *D, const ssa_item *items, double *small, int r_slot,
                                             int rinv_slot, double early)
{
    extern __shared__ double sm[];
    // Column j at A + j*LDP, rows 0..16 NB - 1 (zero at and beyond b), so the
    // 8 lanes of a group load and store a whole column without masks.  LDP is
    // an odd multiple of 8 doubles: the 8-double row blocks of columns c and
    // c + 1, c + 2, c + 3 start 16 banks apart, and the groups of one
    // half-wave, which hold consecutive pair slots and hence consecutive
    // column numbers, touch disjoint banks.
    constexpr int ROWS = 16 * NB, LDP = ROWS + 8;
    double *A = sm;
    __shared__ int perm[SSA_BMAX + 1];
    __shared__ double S[SSA_BMAX + 1];
    __shared__ int rotated, big;
    const ssa_item it = items[blockIdx.x];
    const ssa_desc d = D[it.s];
    const int b = d.b, bp = (b + 1) & ~1;
    const double *R = small + d.off_small + (long long)r_slot * SSA_BMAX * SSA_BMAX;
    // A = R^T: column j is row j of R (R column-major, ld SSA_BMAX); all loads
    // are issued before the first LDS store (consecutive threads read
    // consecutive j of one row i of R)
    {
        constexpr int PER = ROWS * ROWS / 256;
        double v[PER];
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int idx = (int)threadIdx.x + 256 * q, i = idx / ROWS, j = idx - i * ROWS;
            const bool in = i < b && j < b;
            const double x = R[in ? (long long)i * SSA_BMAX + j : 0];
            v[q] = in ? x : 0.0;
        }
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int idx = (int)threadIdx.x + 256 * q, i = idx / ROWS, j = idx - i * ROWS;
            A[j * LDP + i] = v[q];
        }
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nwave = blockDim.x >> 6;
    // 32 groups of 8 lanes; group g rotates pair g of the round (bp/2 <= 40 pairs:
    // groups take pairs g and g + 32), lanes stride the b rows
    const int grp = threadIdx.x >> 3, gl = threadIdx.x & 7;
    // A sweep rotates every pair whose cosine exceeds 1e-15.  Convergence is
    // quadratic: after a sweep in which no pair's cosine exceeded `early`, the
    // remaining cosines are of order early^2, so the sweep that would only
    // confirm convergence is skipped.
    int sweeps = 0;
    while (sweeps < 40) {
        if (threadIdx.x == 0) rotated = big = 0;
        __syncthreads();
        int any_local = 0, big_local = 0;
        for (int round = 0; round < bp - 1; round++) {
            for (int pi = grp; pi < bp / 2; pi += 32) {
                // circle-method pairing without integer division
                int p, q;
                if (pi == 0) { p = round; q = bp - 1; }
                else {
                    p = round + pi;
                    if (p >= bp - 1) p -= bp - 1;
                    q = round - pi;
                    if (q < 0) q += bp - 1;
                }
                // The rotation is defined on the ordered pair (min, max) of the two
                // column numbers; columns are loaded in slot order (x = p, y = q) and
                // a swapped pair exchanges the roles of the norms and the sign of s.
                const bool sw = p > q;
                // rows gl + 8m of both columns, all loads issued up front
                constexpr int MB = 2 * NB;     // rows per lane: 16 NB rows over 8 lanes
                double x[MB], y[MB];
#pragma unroll
                for (int m = 0; m < MB; m++) {
                    x[m] = A[p * LDP + gl + 8 * m];
                    y[m] = A[q * LDP + gl + 8 * m];
                }
                double xx = 0, yy = 0, ga = 0;
#pragma unroll
                for (int m = 0; m < MB; m++) {
                    xx = fma(x[m], x[m], xx);
                    yy = fma(y[m], y[m], yy);
                    ga = fma(x[m], y[m], ga);
                }
                xx = sum8(xx);
                yy = sum8(yy);
                ga = sum8(ga);
                const double al = sw ? yy : xx, be = sw ? xx : yy;
                // cosine tests squared: |ga| > e sqrt(al be)  <=>  ga^2 > e^2 al be
                const double ab = al * be, g2 = ga * ga;
                if (ga != 0.0 && g2 > early * early * ab) big_local = 1;
                if (ga != 0.0 && g2 > 1e-30 * ab) {
                    // The rotation of tan(2 theta) = 2 ga / (be - al), |theta| <= pi/4
                    // (Rutishauser's t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2)),
                    // zeta = (be - al) / (2 ga)), from two refined reciprocal square
                    // roots in series instead of four reciprocals and roots:
                    //   r = sqrt(d^2 + 4 ga^2), cos 2theta = |d| / r,
                    //   c = sqrt((1 + cos 2theta) / 2), s = sign(zeta) |ga| / (r c).
                    const double d = be - al;
                    const double rinv = rsq_nr(fma(d, d, 4.0 * g2));
                    const double hc = fma(0.5 * fabs(d), rinv, 0.5);
                    const double ic = rsq_nr(hc);
                    const double sg = (d == 0.0 || ((d > 0.0) == (ga > 0.0))) ? 1.0 : -1.0;
                    const double c = hc * ic, s = (sw ? -sg : sg) * fabs(ga) * rinv * ic;
                    // (min, max) <- (c min - s max, s min + c max), written for (p, q)
#pragma unroll
                    for (int m = 0; m < MB; m++) {
                        A[p * LDP + gl + 8 * m] = c * x[m] - s * y[m];
                        A[q * LDP + gl + 8 * m] = s * x[m] + c * y[m];
                    }
                    any_local = 1;
                }
            }
            __syncthreads();
        }
        if (any_local) rotated = 1;
        if (big_local) big = 1;
        __syncthreads();
        const int any = rotated, any_big = big;
        __syncthreads();
        sweeps++;
        if (!any || !any_big) break;
    }
    // singular values = column norms
    for (int j = wave; j < b; j += nwave) {
        double s = 0;
        for (int i = lane; i < b; i += 64) s += A[j * LDP + i] * A[j * LDP + i];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0) S[j] = sqrt(s);
    }
    __syncthreads();
    if (threadIdx.x < b) {
        // a NaN norm (non-finite input) ranks below every number, so the
        // ranks stay a permutation of 0..b-1 and perm indexes A in bounds
        const int i = threadIdx.x;
        const double si = S[i] == S[i] ? S[i] : -1.0;
        int rank = 0;
        for (int j = 0; j < b; j++) {
            const double sj = S[j] == S[j] ? S[j] : -1.0;
            rank += (sj > si) || (sj == si && j < i);
        }
        perm[rank] = i;
    }
    __syncthreads();
    double *misc = slot(small, d, SSA_SLOT_MISC);
    double *WK = slot(small, d, SSA_SLOT_WK), *WKS = slot(small, d, SSA_SLOT_WKS);
    // colsum(Q) = colsum(Y) R_Y^-1 of the final orthonormalisation of Y
    if (threadIdx.x < b) {
        const double *Ri = small + d.off_small + (long long)rinv_slot * SSA_BMAX * SSA_BMAX;
        const int j = threadIdx.x;
        double cq = 0.0;
        for (int i = 0; i <= j; i++) cq += misc[SSA_MISC_COLSUMY + i] * Ri[(long long)j * SSA_BMAX + i];
        misc[SSA_MISC_COLSUMQ + j] = cq;
    }
    __syncthreads();
    const double smax = S[perm[0]];
    for (int j = wave; j < d.k; j += nwave) {
        const int src = perm[j];
        double dot = 0;
        for (int i = lane; i < b; i += 64) dot += misc[SSA_MISC_COLSUMQ + i] * A[src * LDP + i];
        for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o);
        const double sg = dot < 0.0 ? -1.0 : 1.0;
        const double sv = S[src];
        const double inv = (sv > 1e-300 && sv > 1e-15 * smax) ? 1.0 / sv : 0.0;
        // right singular vector of R: the normalised converged column of R^T
        const double wn = sv > 1e-300 ? sg / sv : 0.0;
        for (int i = lane; i < b; i += 64) {
            const double w = wn * A[src * LDP + i];
            WK[(long long)j * SSA_BMAX + i] = w;
            WKS[(long long)j * SSA_BMAX + i] = w * inv;
        }
        if (lane == 0) {
            misc[SSA_MISC_SIGMA + j] = sv;
            misc[SSA_MISC_LAMBDA + j] = sv * sv;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tv = 0;
        for (int j = 0; j < d.k; j++) tv += S[perm[j]] * S[perm[j]];
        misc[SSA_MISC_TOTVAR] = tv;
        misc[SSA_MISC_SWEEPS] = sweeps;
    }
}

// splits of a reconstruction item: at most nsplit, and no more than the staging
// area has group slots for (stage_cap pieces of AL * SK * n)
__device__ __forceinline__ int recon_splits(const ssa_desc &d, int nsplit)
{
    const long long slots = d.stage_cap / ((long long)d.AL * d.SK * d.n);
    return (int)max(1LL, min((long long)nsplit, slots));
}

// ---------------------------------------------------------------------------
// reconstruction pieces.  Item (s, gid, a, sk): the convolution of the row
// chunk a of channel m = gdesc[gid].z (U rows) with the column chunk sk
// (V rows), summed over the group's components:
//   piece = sum_g (s_g u_g,a) (*) v_g,sk = Im( T+( sum_g T-(z_g)^2 ) ) / (2n),
//   z_g = sqrt(s_g) (u_g,a + i v_g,sk)
// (the imaginary part of the self-convolution of a + ib is 2 a (*) b).
// ---------------------------------------------------------------------------
template <typename T, int LOGN>
__global__ __launch_bounds__((FftGeom<T, LOGN>::NT), (FftGeom<T, LOGN>::OCC)) void k_recon(const ssa_desc *__restrict__ D,
                                                              const ssa_item *__restrict__ items, T *bulk,
                                                              const double *__restrict__ small,
                                                              const int4 *__restrict__ gdesc,
                                                              const int *__restrict__ glist,
                                                              const C2<T> *__restrict__ tw, int scale_sigma,
                                                              int limit_gsize, int nsplit, int n_items)
{
    using G = FftGeom<T, LOGN>;
    constexpr int E = G::E, NT = G::NT;
    constexpr bool SPLIT = G::SPLIT;
    __shared__ T lre[G::LDS];
    __shared__ T lim[SPLIT ? 1 : G::LDS];
    const int item = xcd_item(n_items);
    if (item < 0) return;
    const ssa_item it = items[item];
    const ssa_desc d = D[it.s];
    // item.x = group * S + split: split j of S sums entries [ce j / S,
    // ce (j+1) / S) of the group's list (ce entries: the whole list, or with
    // limit_gsize at most min(k, gsize)) into its own staged piece (the
    // overlap-add sums the pieces); S = 1 sums the whole list
    const int S = recon_splits(d, nsplit);
    const int4 gd = gdesc[it.x / S];
    const int a = it.y, sk = it.z, m = gd.z;
    int g0 = 0, g1 = gd.y;
    if (S > 1) {
        int ce = gd.y;
        if (limit_gsize) ce = min(min(ce, d.k), d.gsize);
        const int j = it.x % S;
        g0 = ce * j / S;
        g1 = ce * (j + 1) / S;
    }
    const int urow0 = m * d.L + a * d.CL, ulen = min(d.CL, d.L - a * d.CL);
    const int vrow0 = sk * d.CK, vlen = min(d.CK, d.K - sk * d.CK);
    const double *sig = small + d.off_small + (long long)SSA_SLOT_MISC * SSA_BMAX * SSA_BMAX + SSA_MISC_SIGMA;
    C2<T> acc[E], z[E];
#pragma unroll
    for (int e = 0; e < E; e++) acc[e] = C2<T>{0, 0};
    for (int g = g0; g < g1; g++) {
        const int idx = glist[gd.x + g];
        if (idx < 0 || idx >= d.k || (limit_gsize && idx >= d.gsize)) continue;
        const T sc = (T)(scale_sigma ? sqrt(fabs(sig[idx])) : 1.0);
        const T *u = bulk + d.off_U + (long long)idx * d.ML + urow0;
        const T *v = bulk + d.off_V + (long long)idx * d.K + vrow0;
        // every load unconditional (positions past the chunk read its last
        // element, ulen, vlen >= 1) and masked afterwards: a load under a
        // branch whose result is used inside it waits for that load before
        // the next one issues, which serialised the 2E loads of a component.
        // The opaque thread index keeps the clamped offsets from being
        // hoisted out of the component loop (they would be spilled).
        int tq = threadIdx.x;
        asm volatile("" : "+v"(tq));
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int q = tq + e * NT;
            const T uu = u[min(q, ulen - 1)], vv = v[min(q, vlen - 1)];
            z[e].x = (q < ulen) ? sc * uu : (T)0;
            z[e].y = (q < vlen) ? sc * vv : (T)0;
        }
        fft<T, LOGN, -1, SPLIT>(z, lre, lim, tw);
#pragma unroll
        for (int e = 0; e < E; e++) {
            acc[e].x += z[e].x * z[e].x - z[e].y * z[e].y;
            acc[e].y += (T)2 * z[e].x * z[e].y;
        }
    }
    fft<T, LOGN, 1, SPLIT>(acc, lre, lim, tw);
    const int plen = ulen + vlen - 1;
    T *st = bulk + d.off_stage + (((long long)it.x * d.AL + a) * d.SK + sk) * G::N;
    const T sc = (T)0.5 / (T)G::N;
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int q = threadIdx.x + e * NT;
        if (q < plen) st[q] = acc[e].y * sc;
    }
}

// ---------------------------------------------------------------------------
// component spectra cache (MKL/ssa_opt.h:2110-2323 caches FFT(sigma u) and FFT(v)
// per component; here per chunk, unscaled, the sigma applied when summing).
// Item (s, pair p, chunk c): x = chunk c of component 2p, y = the same chunk of
// component 2p + 1 (zero beyond k); Z = T-(x + i y), and since x and y are real,
// X[j] = (Z[j] + conj Z[n-j]) / 2 and Y[j] = (Z[j] - conj Z[n-j]) / (2i).  The
// mirror values come through LDS one plane at a time.  Only the n/2 + 1 values
// j <= n/2 are stored (the rest are conjugates).
// ---------------------------------------------------------------------------
template <typename T, int LOGN>
__global__ __launch_bounds__((FftGeom<T, LOGN>::NT), (FftGeom<T, LOGN>::OCC)) void k_cache_build(
    const ssa_desc *__restrict__ D, const ssa_item *__restrict__ items, T *bulk, const C2<T> *__restrict__ tw)
{
    using G = FftGeom<T, LOGN>;
    constexpr int E = G::E, NT = G::NT, N = G::N, H = N / 2 + 1;
    constexpr bool SPLIT = G::SPLIT;
    __shared__ T lre[G::LDS];
    __shared__ T lim[SPLIT ? 1 : G::LDS];
    const ssa_item it = items[blockIdx.x];
    const ssa_desc d = D[it.s];
    const int g0 = 2 * it.x, g1 = g0 + 1, c = it.y, nLc = d.M * d.AL;
    const bool two = g1 < d.k;
    const T *src;
    long long ld;
    int len;
    if (c < nLc) {
        const int m = c / d.AL, a = c - m * d.AL;
        src = bulk + d.off_U + (long long)m * d.L + (long long)a * d.CL;
        ld = d.ML;
        len = min(d.CL, d.L - a * d.CL);
    } else {
        const int sk = c - nLc;
        src = bulk + d.off_V + (long long)sk * d.CK;
        ld = d.K;
        len = min(d.CK, d.K - sk * d.CK);
    }
    const T *x0 = src + (long long)g0 * ld, *x1 = src + (long long)(two ? g1 : g0) * ld;
    C2<T> z[E];
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int q = threadIdx.x + e * NT;
        z[e].x = q < len ? x0[q] : (T)0;
        z[e].y = (two && q < len) ? x1[q] : (T)0;
    }
    fft<T, LOGN, -1, SPLIT>(z, lre, lim, tw);
    // mirror values Z[n - j] through LDS, real plane then imaginary plane
    T mr[E], mi[E];
#pragma unroll
    for (int ph = 0; ph < 2; ph++) {
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; e++) lre[threadIdx.x + e * NT] = ph ? z[e].y : z[e].x;
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int j = threadIdx.x + e * NT;
            const T v = lre[(N - j) & (N - 1)];
            if (ph) mi[e] = v;
            else mr[e] = v;
        }
    }
    const int nC = nLc + d.SK;
    C2<T> *out0 = reinterpret_cast<C2<T> *>(bulk + d.off_cache) + ((long long)g0 * nC + c) * H;
    C2<T> *out1 = reinterpret_cast<C2<T> *>(bulk + d.off_cache) + ((long long)g1 * nC + c) * H;
    const T h = (T)0.5;
#pragma unroll
    for (int e = 0; e < E; e++) {
        const int j = threadIdx.x + e * NT;
        if (j < H) {
            out0[j] = C2<T>{h * (z[e].x + mr[e]), h * (z[e].y - mi[e])};
            if (two) out1[j] = C2<T>{h * (z[e].y + mi[e]), h * (mr[e] - z[e].x)};
        }
    }
}

// ---------------------------------------------------------------------------
// reconstruction pieces from the cache.  Piece P = (gid AL + a) SK + sk of the
// launch: sum over the group's components of s_g U_g,a[j] V_g,sk[j] (the
// product of two real-signal spectra is Hermitian, so j > n/2 reads n - j and
// conjugates).  Item (s, P0): pieces P0 and P0 + 1 share one inverse
// transform, T+(A + i B) = a + i b for real a and b; each is staged where
// k_recon stages a piece with one split.
// ---------------------------------------------------------------------------
template <typename T, int LOGN>
__global__ __launch_bounds__((FftGeom<T, LOGN>::NT), (FftGeom<T, LOGN>::OCC)) void k_recon_cached(
    const ssa_desc *__restrict__ D, const ssa_item *__restrict__ items, T *bulk, const double *__restrict__ small,
    const int4 *__restrict__ gdesc, const int *__restrict__ glist, const C2<T> *__restrict__ tw, int scale_sigma,
    int n_pieces, int n_items)
{
    using G = FftGeom<T, LOGN>;
    constexpr int E = G::E, NT = G::NT, N = G::N, H = N / 2 + 1;
    constexpr bool SPLIT = G::SPLIT;
    __shared__ T lre[G::LDS];
    __shared__ T lim[SPLIT ? 1 : G::LDS];
    const int item = xcd_item(n_items);
    if (item < 0) return;
    const ssa_item it = items[item];
    const ssa_desc d = D[it.s];
    const int nC = d.M * d.AL + d.SK, per = d.AL * d.SK;
    const double *sig = small + d.off_small + (long long)SSA_SLOT_MISC * SSA_BMAX * SSA_BMAX + SSA_MISC_SIGMA;
    const C2<T> *cache = reinterpret_cast<const C2<T> *>(bulk + d.off_cache);
    C2<T> acc[E];
#pragma unroll
    for (int e = 0; e < E; e++) acc[e] = C2<T>{0, 0};
    int plen[2] = {0, 0};
    for (int h = 0; h < 2; h++) {
        const int P = it.x + h;
        if (P >= n_pieces) break;
        const int gid = P / per, rem = P - gid * per, a = rem / d.SK, sk = rem - a * d.SK;
        const int4 gd = gdesc[gid];
        const int cu = gd.z * d.AL + a, cv = d.M * d.AL + sk;
        plen[h] = min(d.CL, d.L - a * d.CL) + min(d.CK, d.K - sk * d.CK) - 1;
        C2<T> p[E];
#pragma unroll
        for (int e = 0; e < E; e++) p[e] = C2<T>{0, 0};
        for (int g = 0; g < gd.y; g++) {
            const int idx = glist[gd.x + g];
            if (idx < 0 || idx >= d.k) continue;
            const T sc = (T)(scale_sigma ? sig[idx] : 1.0);
            const C2<T> *cu_p = cache + ((long long)idx * nC + cu) * H;
            const C2<T> *cv_p = cache + ((long long)idx * nC + cv) * H;
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int j = threadIdx.x + e * NT;
                const int jj = j <= N / 2 ? j : N - j;
                const C2<T> uu = cu_p[jj], vv = cv_p[jj];
                p[e].x += sc * (uu.x * vv.x - uu.y * vv.y);
                p[e].y += sc * (uu.x * vv.y + uu.y * vv.x);
            }
        }
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int j = threadIdx.x + e * NT;
            const T py = j <= N / 2 ? p[e].y : -p[e].y;   // conjugate beyond n/2
            // piece 0 into the real part, piece 1 (times i) into the imaginary part
            if (h == 0) acc[e] = C2<T>{p[e].x, py};
            else acc[e] = C2<T>{acc[e].x - py, acc[e].y + p[e].x};
        }
    }
    fft<T, LOGN, 1, SPLIT>(acc, lre, lim, tw);
    const T sc = (T)1 / (T)N;
    for (int h = 0; h < 2; h++) {
        const int P = it.x + h;
        if (P >= n_pieces) break;
        T *st = bulk + d.off_stage + (long long)P * N;
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int q = threadIdx.x + e * NT;
            if (q < plen[h]) st[q] = (h ? acc[e].y : acc[e].x) * sc;
        }
    }
}

// overlap-add of the pieces of group gid and diagonal-averaging weights.
// item (s, gid, t tile of `tile` samples); weight 0: 1/c_t, 1: sqrt(1/c_t).
// A sample sums its pieces in (split, L chunk, K chunk) order.  ALL: the
// loads of all pieces, eight at a time (clamped positions, masked
// afterwards), for short launches, whose time is the latency of one item;
// otherwise only the pieces that cover the sample are read (less traffic
// when a launch fills the device).
template <typename T, bool ALL>
__global__ __launch_bounds__(256) void k_olap(const ssa_desc *D, const ssa_item *items, const int4 *gdesc, T *bulk,
                                              int weight, int nsplit, int tile)
{
    const ssa_item it = items[blockIdx.x];
    const ssa_desc d = D[it.s];
    const int n = d.n, S = recon_splits(d, nsplit), AL = d.AL, SK = d.SK, np = S * AL * SK;
    const T *stage = bulk + d.off_stage + (long long)it.x * S * AL * SK * n;
    for (int t = it.y * tile + threadIdx.x; t < min(d.N, (it.y + 1) * tile); t += 256) {
        double s = 0;
        if constexpr (ALL) {
            int a = 0, sk = 0;   // chunks of piece p0 + r (the split is p / (AL SK))
            for (int p0 = 0; p0 < np; p0 += 8) {
                T v[8];
                bool ok[8];
#pragma unroll
                for (int r = 0; r < 8; r++) {
                    const int pc = min(p0 + r, np - 1);
                    const int q = t - (a * d.CL + sk * d.CK);
                    const int plen = min(d.CL, d.L - a * d.CL) + min(d.CK, d.K - sk * d.CK) - 1;
                    ok[r] = p0 + r < np && q >= 0 && q < plen;
                    v[r] = stage[(long long)pc * n + min(max(q, 0), n - 1)];
                    if (++sk == SK) {
                        sk = 0;
                        if (++a == AL) a = 0;
                    }
                }
#pragma unroll
                for (int r = 0; r < 8; r++)
                    if (ok[r]) s += (double)v[r];
            }
        } else {
            for (int j = 0; j < S; j++)
                for (int a = 0; a < AL; a++) {
                    const int ulen = min(d.CL, d.L - a * d.CL);
                    for (int sk = 0; sk < SK; sk++) {
                        const int q = t - (a * d.CL + sk * d.CK);
                        const int vlen = min(d.CK, d.K - sk * d.CK);
                        if (q >= 0 && q < ulen + vlen - 1) s += (double)stage[((long long)(j * AL + a) * SK + sk) * n + q];
                    }
                }
        }
        const int c = min(min(t + 1, d.L), min(d.K, d.N - t));
        const double w = 1.0 / (double)c;
        bulk[d.off_out + (long long)it.x * d.N + t] = (T)(s * (weight ? sqrt(w) : w));
    }
}

// ---------------------------------------------------------------------------
// The first and last SSA_EDGE samples of a group's reconstruction by direct
// anti-diagonal sums, overwriting what k_olap wrote there:
//   out[t] = w_t * sum_g s_g sum_{i + j = t} u_g[i] v_g[j],  i < L, j < K.
// A transform-based convolution leaves an absolute error of about eps * s_g
// |u_g| |v_g| at every output sample; diagonal averaging divides it by the
// count min(t + 1, L, K, N - t), which is small only near the two ends, so the
// ends carry almost all of the error of an FP32 reconstruction.  There the sums
// have at most SSA_EDGE terms each.  Item (s, gid); products and sums in FP64.
// The group's components are staged EDGE_CG at a time: every load of a stage
// is issued before its barrier (index-clamped and masked, not under
// branches), so a group pays one memory latency per EDGE_CG components.  The
// P thread groups of 2 SSA_EDGE threads each sum EDGE_CG / P of a stage's
// components (two chains side by side, terms read eight at a time at
// constant LDS offsets); the first group then adds the stage's sums in
// component order.  Each sample is summed in component order, terms in index
// order, whatever the staging and P: P = 4 for a launch of few items (one
// series: its time is one item's latency), P = 1 for a batch (more items per
// CU).
// scale_sigma, weight, limit_gsize and the group layout as in k_recon/k_olap.
// ---------------------------------------------------------------------------
constexpr int EDGE_CG = 8;

template <typename T, int P>
__global__ __launch_bounds__(P * 2 * SSA_EDGE) void k_edge(const ssa_desc *__restrict__ D,
                                                           const ssa_item *__restrict__ items,
                                                           const int4 *__restrict__ gdesc,
                                                           const int *__restrict__ glist, T *bulk,
                                                           const double *__restrict__ small, int scale_sigma,
                                                           int weight, int limit_gsize)
{
    constexpr int E = SSA_EDGE, CG = EDGE_CG, NTH = P * 2 * SSA_EDGE, CPG = CG / P;
    static_assert(CG % (2 * P) == 0, "each thread group sums whole pairs of a stage's components");
    __shared__ double uh[CG][E], vh[CG][E], ut[CG][E], vt[CG][E];
    __shared__ double ss[CG][2 * E];
    __shared__ int gl[NTH];
    const ssa_item it = items[blockIdx.x];
    const ssa_desc d = D[it.s];
    const int4 gd = gdesc[it.x];
    const int L = d.L, K = d.K, N = d.N, m = gd.z;
    // thread group p sums components p + P c, c < CPG, of a stage; within a
    // group, thread j < E: head sample t = j; thread E + j: tail sample t = N - E + j
    const int p = threadIdx.x / (2 * E), tt = threadIdx.x % (2 * E);
    const int j = tt & (E - 1);
    const bool tail = tt >= E;
    const int t = tail ? N - E + j : j;
    // a sample both ranges contain (N < 2E) is written by the head thread
    const bool mine = t >= 0 && t < N && (!tail || t >= E);
    const int lo = max(0, t - K + 1), hi = min(t, L - 1);
    const double *sig = small + d.off_small + (long long)SSA_SLOT_MISC * SSA_BMAX * SSA_BMAX + SSA_MISC_SIGMA;
    // rows of u and v this thread stages: head rows [0, E) and tail rows
    // [L - E, L) of u, [0, E) and [K - E, K) of v (clamped into the columns;
    // rows outside them are staged as zero)
    const int ur = tail ? L - E + j : j, vr = tail ? K - E + j : j;
    const bool uin = ur >= 0 && ur < L, vin = vr >= 0 && vr < K;
    const int urc = min(max(ur, 0), L - 1), vrc = min(max(vr, 0), K - 1);
    const T *ubase = bulk + d.off_U + (long long)m * L + urc;
    const T *vbase = bulk + d.off_V + vrc;
    const int uo = tail ? L - E : 0, vo = tail ? K - E : 0, nt = hi - lo + 1;
    double acc = 0.0;
    // the group's list in blocks of NTH entries: component indices in gl, -1
    // for one that is skipped
    for (int b0 = 0; b0 < gd.y; b0 += NTH) {
        const int ng = min(NTH, gd.y - b0);
        __syncthreads();   // the previous block's indices are no longer read
        if ((int)threadIdx.x < ng) {
            const int idx = glist[gd.x + b0 + threadIdx.x];
            gl[threadIdx.x] = (idx < 0 || idx >= d.k || (limit_gsize && idx >= d.gsize)) ? -1 : idx;
        }
        for (int g0 = 0; g0 < ng; g0 += CG) {
            __syncthreads();   // gl is written; the previous stage's values are no longer read
            // this thread's components of the stage: cs[c] = p + P c
            int id[CPG];
            {
                double a[CPG], bv[CPG];
#pragma unroll
                for (int c = 0; c < CPG; c++) {
                    const int cc = p + P * c;
                    id[c] = g0 + cc < ng ? gl[g0 + cc] : -1;
                    const int ci = max(id[c], 0);
                    a[c] = (double)ubase[(long long)ci * d.ML];
                    bv[c] = (double)vbase[(long long)ci * K];
                }
#pragma unroll
                for (int c = 0; c < CPG; c++) {
                    const int cc = p + P * c;
                    const double ua = id[c] >= 0 && uin ? a[c] : 0.0, va = id[c] >= 0 && vin ? bv[c] : 0.0;
                    if (!tail) {
                        uh[cc][j] = ua;
                        vh[cc][j] = va;
                    } else {
                        ut[cc][j] = ua;
                        vt[cc][j] = va;
                    }
                }
            }
            __syncthreads();
            if (mine) {
                // head (t < E): both indices below E; tail (t >= N - E): i >= L - E
                // and t - i >= K - E.  Terms i = lo..hi in order: term lo + ii
                // reads up[ii] and vq[-ii] (constant offsets within a group of
                // 8); the last, partial group reads clamped positions and adds
                // only its terms.
#pragma unroll
                for (int c = 0; c < CPG; c += 2) {
                    if (id[c] < 0 && id[c + 1] < 0) continue;
                    const int c0 = p + P * c, c1 = p + P * (c + 1);
                    const double *up0 = (tail ? ut[c0] : uh[c0]) + (lo - uo);
                    const double *vq0 = (tail ? vt[c0] : vh[c0]) + (t - lo - vo);
                    const double *up1 = (tail ? ut[c1] : uh[c1]) + (lo - uo);
                    const double *vq1 = (tail ? vt[c1] : vh[c1]) + (t - lo - vo);
                    double s0 = 0.0, s1 = 0.0;
                    int ii = 0;
                    for (; ii + 8 <= nt; ii += 8) {
                        double x0[8], y0[8], x1[8], y1[8];
#pragma unroll
                        for (int r = 0; r < 8; r++) {
                            x0[r] = up0[ii + r];
                            y0[r] = (vq0 - ii - 7)[7 - r];
                            x1[r] = up1[ii + r];
                            y1[r] = (vq1 - ii - 7)[7 - r];
                        }
#pragma unroll
                        for (int r = 0; r < 8; r++) {
                            s0 = fma(x0[r], y0[r], s0);
                            s1 = fma(x1[r], y1[r], s1);
                        }
                    }
                    if (ii < nt) {
                        double x0[8], y0[8], x1[8], y1[8];
#pragma unroll
                        for (int r = 0; r < 8; r++) {
                            const int q = min(ii + r, nt - 1);
                            x0[r] = up0[q];
                            y0[r] = vq0[-q];
                            x1[r] = up1[q];
                            y1[r] = vq1[-q];
                        }
#pragma unroll
                        for (int r = 0; r < 8; r++)
                            if (ii + r < nt) {
                                s0 = fma(x0[r], y0[r], s0);
                                s1 = fma(x1[r], y1[r], s1);
                            }
                    }
                    ss[c0][tt] = s0;
                    ss[c1][tt] = s1;
                }
            }
            __syncthreads();
            if (p == 0 && mine) {
#pragma unroll
                for (int c = 0; c < CG; c++) {
                    const int idc = g0 + c < ng ? gl[g0 + c] : -1;
                    if (idc >= 0) acc = fma(scale_sigma ? sig[idc] : 1.0, ss[c][tt], acc);
                }
            }
        }
    }
    if (p != 0 || !mine) return;
    const int c = min(min(t + 1, L), min(K, N - t));
    const double w = 1.0 / (double)c;
    bulk[d.off_out + (long long)it.x * N + t] = (T)(acc * (weight ? sqrt(w) : w));
}

// every series' step output (reconstruction + h forecasts) into its gather slot
template <typename T>
__global__ __launch_bounds__(256) void k_gather(const ssa_desc *__restrict__ D, const ssa_item *__restrict__ items,
                                                T *bulk, int h)
{
    const ssa_desc d = D[items[blockIdx.x].s];
    const T *src = bulk + d.off_out;
    T *dst = bulk + d.off_gather;
    for (int t = threadIdx.x; t < d.N + h; t += 256) dst[t] = src[t];
}

// The results one decomposition hands to the host: the FP64 block of singular
// values, eigenvalues and totals (misc_n values), then un values of U and vn
// of V in the panel type, packed into one buffer (page-locked host memory,
// written through its device address).
template <typename T>
__global__ __launch_bounds__(256) void k_pack_results(const double *__restrict__ misc, int misc_n,
                                                      const T *__restrict__ U, long long un,
                                                      const T *__restrict__ V, long long vn, char *out)
{
    double *om = reinterpret_cast<double *>(out);
    T *ouv = reinterpret_cast<T *>(out + 8LL * misc_n);
    const long long n = misc_n + un + vn, step = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += step) {
        if (i < misc_n) om[i] = misc[i];
        else if (i < misc_n + un) ouv[i - misc_n] = U[i - misc_n];
        else ouv[i - misc_n] = V[i - misc_n - un];
    }
}

// W-correlation normalisation: W_ij = G_ij / sqrt(G_ii G_jj), rows whose
// sigma * ||h||_w <= 1e-12 are zero (MKL/ssa_opt.h:2466-2471); row-major k x k.
__global__ __launch_bounds__(256) void k_wnorm(const ssa_desc *D, const ssa_item *items, double *small)
{
    const ssa_item it = items[blockIdx.x];
    const ssa_desc d = D[it.s];
    const int k = d.k, tiles = (d.N + SSA_GRAM_ROWS - 1) / SSA_GRAM_ROWS;
    const double *part = small + d.off_part;
    const double *sig = slot(small, d, SSA_SLOT_MISC) + SSA_MISC_SIGMA;
    double *W = slot(small, d, SSA_SLOT_W);
    __shared__ double dg[SSA_BMAX];
    const int stride = k * k + k;
    if (threadIdx.x < k) {
        double s = 0;
        for (int t = 0; t < tiles; t++) s += part[(long long)t * stride + threadIdx.x * k + threadIdx.x];
        dg[threadIdx.x] = s;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < k * k; idx += 256) {
        const int i = idx / k, j = idx - i * k;
        double s = 0;
        for (int t = 0; t < tiles; t++) s += part[(long long)t * stride + idx];
        const double ni = sqrt(dg[i]), nj = sqrt(dg[j]);
        const bool zi = !(fabs(sig[i]) * ni > 1e-12), zj = !(fabs(sig[j]) * nj > 1e-12);
        W[idx] = (zi || zj) ? 0.0 : s / (ni * nj);
    }
}

// LRF (verticality, coefficients) and the h-step recurrence on the group's
// reconstruction (at off_out, gid 0); forecasts written after it.
template <typename T>
__global__ __launch_bounds__(256) void k_forecast(ssa_desc *D, const ssa_item *items, const int4 *gdesc,
                                                  const int *glist, T *bulk, double *small, int h, int limit_gsize,
                                                  int vmode)
{
    __shared__ double red[32];
    __shared__ double pis[SSA_BMAX];
    __shared__ int ids[SSA_BMAX];
    const ssa_item it = items[blockIdx.x];
    ssa_desc *dp = D + it.s;
    const ssa_desc d = *dp;
    const int4 gd = gdesc[0];
    const int L = d.L;
    int n = gd.y < SSA_BMAX ? gd.y : SSA_BMAX;
    if (limit_gsize && n > d.gsize) n = d.gsize;
    const T *U = bulk + d.off_U;
    if (threadIdx.x == 0) {
        for (int g = 0; g < n && g < SSA_BMAX; g++) ids[g] = glist[gd.x + g];
    }
    __syncthreads();
    if (threadIdx.x < n) pis[threadIdx.x] = (double)U[(long long)ids[threadIdx.x] * d.ML + L - 1];
    __syncthreads();
    double nu = 0;
    for (int g = 0; g < n; g++) nu += pis[g] * pis[g];
    if (!(nu < 1.0 - 1e-10)) {
        if (threadIdx.x == 0) dp->status = 2;
        return;
    }
    const double sc = 1.0 / (1.0 - nu);
    // R and the rolling window live in the staging area (FP64 stored as T pairs is avoided:
    // use the small arena's partial region, sized for >= L + h doubles)
    double *R = small + d.off_part;
    double *buf = R + L;
    const T *rec = bulk + d.off_out;
    for (int j = threadIdx.x; j < L - 1; j += 256) {
        if (!vmode) {
            double s = 0;
            for (int g = 0; g < n; g++) s += pis[g] * (double)U[(long long)ids[g] * d.ML + j];
            R[j] = s * sc;
        }
        buf[j] = (double)rec[d.N - (L - 1) + j];
    }
    __syncthreads();
    T *out = bulk + d.off_out + d.N;
    if (vmode) {
        // V-forecast (MKL/ssa_opt.h:2906-2965): each step projects the last
        // L-1 values on every eigenvector of the group, x = sc * sum_g pi_g
        // (U_g[0:L-1] . z); wave w takes components w, w+4, ...
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        for (int t = 0; t < h; t++) {
            double x = 0;
            for (int g = w; g < n; g += 4) {
                const T *u = U + (long long)ids[g] * d.ML;
                double dot = 0;
                for (int j = lane; j < L - 1; j += 64) dot += (double)u[j] * buf[t + j];
                for (int o = 32; o > 0; o >>= 1) dot += __shfl_xor(dot, o);
                x += pis[g] * dot;
            }
            if (lane == 0) red[w] = x;
            __syncthreads();
            if (threadIdx.x == 0) {
                const double v = (red[0] + red[1] + red[2] + red[3]) * sc;
                buf[L - 1 + t] = v;
                out[t] = (T)v;
            }
            __syncthreads();
        }
        return;
    }
    for (int t = 0; t < h; t++) {
        double s = 0, dummy = 0;
        for (int j = threadIdx.x; j < L - 1; j += 256) s += R[j] * buf[t + j];
        block_sum2<T>(s, dummy, red);
        if (threadIdx.x == 0) {
            buf[L - 1 + t] = s;
            out[t] = (T)s;
        }
        __syncthreads();
    }
}

}  // namespace

// ===========================================================================
// launchers
// ===========================================================================
extern "C" {

// Twiddle tables live in static device arrays (one copy per device, filled
// once from a host table built in long double), so no entry point allocates
// device memory for them.  FP64: the per-pass tables (tw_offset); FP32: the
// base table exp(+2 pi i m / 2^TW_BASE_LOG).
__device__ __align__(256) C2<double> tw_table64[TW_ENTRIES];
__device__ __align__(256) C2<float> tw_table32[1 << TW_BASE_LOG];

constexpr int TW_MAX_DEVICES = 64;
static C2<double> h_tw64[TW_ENTRIES];
static C2<float> h_tw32[1 << TW_BASE_LOG];
static void *tw_dev[2][TW_MAX_DEVICES];
static std::mutex tw_lock;

static void build_host_twiddles()
{
    const long double two_pi = 6.283185307179586476925286766559005768L;
    const long long n32 = 1LL << TW_BASE_LOG;
    for (long long m = 0; m < n32; m++) {
        const long double a = two_pi * (long double)m / (long double)n32;
        h_tw32[m] = C2<float>{(float)cosl(a), (float)sinl(a)};
    }
    for (int rlog = 1; rlog <= 4; rlog++)
        for (int nslog = 0; nslog <= TW_NSLOG_MAX; nslog++) {
            const long long base = tw_offset(nslog, rlog);
            const int NS = 1 << nslog, R = 1 << rlog;
            for (int kk = 0; kk < NS; kk++)
                for (int r = 1; r < R; r++) {
                    // exp(+2 pi i r kk / (NS R))
                    const long double a = two_pi * (long double)(r * kk) / (long double)(NS * R);
                    h_tw64[base + (long long)(r - 1) * NS + kk] = C2<double>{(double)cosl(a), (double)sinl(a)};
                }
        }
}

int ssa_dev_twiddles(int prec, void **table_out, void *stream)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= TW_MAX_DEVICES) return -1;
    const int pi = prec == SSA_PREC_F64 ? 0 : 1;
    std::lock_guard<std::mutex> guard(tw_lock);
    if (!tw_dev[pi][dev]) {
        static bool host_built = false;
        if (!host_built) {
            build_host_twiddles();
            host_built = true;
        }
        void *p = nullptr;
        e = pi == 0 ? hipGetSymbolAddress(&p, HIP_SYMBOL(tw_table64)) : hipGetSymbolAddress(&p, HIP_SYMBOL(tw_table32));
        if (e != hipSuccess) return (int)e;
        // on the caller's stream: work on the null stream would occupy a
        // hardware queue of its own, which the engines' streams then share
        hipStream_t st = (hipStream_t)stream;
        e = pi == 0 ? hipMemcpyAsync(p, h_tw64, sizeof(h_tw64), hipMemcpyHostToDevice, st)
                    : hipMemcpyAsync(p, h_tw32, sizeof(h_tw32), hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) return (int)e;
        tw_dev[pi][dev] = p;
    }
    *table_out = tw_dev[pi][dev];
    return 0;
}

#define SSA_LGN_CASES(X) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13)

static bool lgn_ok(int prec, int lgn) { (void)prec; return lgn >= 6 && lgn <= 13; }

int ssa_dev_spectra(int prec, int lgn, const ssa_desc *D, const ssa_item *items, int n_items, void *bulk,
                    const void *tw, void *stream)
{
    if (n_items <= 0) return 0;
    if (!lgn_ok(prec, lgn)) return -2;
    hipStream_t st = (hipStream_t)stream;
#define CASE(L)                                                                                               \
    case L:                                                                                                   \
        if (prec == SSA_PREC_F64)                                                                             \
            hipLaunchKernelGGL((k_spectra<double, L>), dim3(n_items),                         \
                               dim3(FftGeom<double, L>::NT), 0, st, D, items, (double *)bulk, \
                               (const C2<double> *)tw);                                                       \
        else                                                                                                  \
            hipLaunchKernelGGL((k_spectra<float, L>), dim3(n_items), dim3(FftGeom<float, L>::NT), 0, st, D,   \
                               items, (float *)bulk, (const C2<float> *)tw);                                  \
        break;
    switch (lgn) { SSA_LGN_CASES(CASE) default: return -2; }
#undef CASE
    return (int)hipGetLastError();
}

int ssa_dev_matvec(int prec, int lgn, int dir, int in_panel, int out_panel, int wide, const ssa_desc *D,
                   const ssa_item *items, int n_items, void *bulk, const void *tw, void *stream)
{
    if (n_items <= 0) return 0;
    if (!lgn_ok(prec, lgn)) return -2;
    hipStream_t st = (hipStream_t)stream;
#define LAUNCH(T, L, WD)                                                                                      \
    hipLaunchKernelGGL((k_matvec<T, L, WD>), dim3(xcd_grid(n_items)), dim3(FftGeom<T, L, WD>::NT), 0, st, D,  \
                       items, (T *)bulk, (const C2<T> *)tw, dir, in_panel, out_panel, n_items)
#define CASE(L)                                                                                               \
    case L:                                                                                                   \
        if (prec == SSA_PREC_F64) {                                                                           \
            if (wide) LAUNCH(double, L, true);                                                                \
            else LAUNCH(double, L, false);                                                                    \
        } else {                                                                                              \
            if (wide) LAUNCH(float, L, true);                                                                 \
            else LAUNCH(float, L, false);                                                                     \
        }                                                                                                     \
        break;
    switch (lgn) { SSA_LGN_CASES(CASE) default: return -2; }
#undef CASE
#undef LAUNCH
    return (int)hipGetLastError();
}

int ssa_dev_mv_spec(int prec, int lgn, int dir, int in_panel, int wide, const ssa_desc *D, const ssa_item *items,
                    int n_items, void *bulk, const void *tw, void *stream)
{
    if (n_items <= 0) return 0;
    if (!lgn_ok(prec, lgn)) return -2;
    hipStream_t st = (hipStream_t)stream;
#define LAUNCH(T, L, WD)                                                                                      \
    hipLaunchKernelGGL((k_mv_spec<T, L, WD>), dim3(xcd_grid(n_items)), dim3(FftGeom<T, L, WD>::NT), 0, st, D, \
                       items, (T *)bulk, (const C2<T> *)tw, dir, in_panel, n_items)
#define CASE(L)                                                                                               \
    case L:                                                                                                   \
        if (prec == SSA_PREC_F64) {                                                                           \
            if (wide) LAUNCH(double, L, true);                                                                \
            else LAUNCH(double, L, false);                                                                    \
        } else {                                                                                              \
            if (wide) LAUNCH(float, L, true);                                                                 \
            else LAUNCH(float, L, false);                                                                     \
        }                                                                                                     \
        break;
    switch (lgn) { SSA_LGN_CASES(CASE) default: return -2; }
#undef CASE
#undef LAUNCH
    return (int)hipGetLastError();
}

int ssa_dev_mv_acc(int prec, int lgn, int dir, int out_panel, int wide, const ssa_desc *D, const ssa_item *items,
                   int n_items, void *bulk, const void *tw, void *stream)
{
    if (n_items <= 0) return 0;
    if (!lgn_ok(prec, lgn)) return -2;
    hipStream_t st = (hipStream_t)stream;
#define LAUNCH(T, L, WD)                                                                                      \
    hipLaunchKernelGGL((k_mv_acc<T, L, WD>), dim3(xcd_grid(n_items)), dim3(FftGeom<T, L, WD>::NT), 0, st, D,  \
                       items, (T *)bulk, (const C2<T> *)tw, dir, out_panel, n_items)
#define CASE(L)                                                                                               \
    case L:                                                                                                   \
        if (prec == SSA_PREC_F64) {                                                                           \
            if (wide) LAUNCH(double, L, true);                                                                \
            else LAUNCH(double, L, false);                                                                    \
        } else {                                                                                              \
            if (wide) LAUNCH(float, L, true);                                                                 \
            else LAUNCH(float, L, false);                                                                     \
        }                                                                                                     \
        break;
    switch (lgn) { SSA_LGN_CASES(CASE) default: return -2; }
#undef CASE
#undef LAUNCH
    return (int)hipGetLastError();
}

int ssa_dev_omega(int prec, const ssa_desc *D, const ssa_item *items, int n_items, void *bulk, long long call,
                  void *stream)
{
    if (n_items <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    // about 1024 workgroups: a launch of few row tiles (one short series)
    // splits each tile's column quads over up to SSA_BMAX / 4 workgroups
    int py = 1024 / n_items;
    py = py < 1 ? 1 : py > SSA_BMAX / 4 ? SSA_BMAX / 4 : py;
    const dim3 grid(n_items, py);
    if (prec == SSA_PREC_F64)
        hipLaunchKernelGGL(k_omega<double>, grid, dim3(256), 0, st, D, items, (double *)bulk, call);
    else hipLaunchKernelGGL(k_omega<float>, grid, dim3(256), 0, st, D, items, (float *)bulk, call);
    return (int)hipGetLastError();
}

extern "C++" {
template <typename T, int NB>
static void launch_gram(int panel, int use_m, const ssa_desc *D, const ssa_item *items, int n_items, void *bulk,
                        double *small, hipStream_t st, bool acc32 = false)
{
    constexpr int NBLK = n_upper_blocks<NB>(), B16 = NB * 16, MP = small_ld<NB>();
    const size_t red = (size_t)(NBLK * 256 + NB * 16) * sizeof(double);
    const size_t ms = (size_t)B16 * MP * sizeof(double);
    // a launch of few row tiles (short series) deals each tile's upper blocks
    // over up to NBLK workgroups
    const int py = n_items >= 512 || NBLK == 1 ? 1 : (1024 / n_items < NBLK ? 1024 / n_items : NBLK);
    const dim3 grid(n_items, py);
    if constexpr (sizeof(T) == 4) {
        if (acc32 && !use_m) {
            if (py > 1)
                hipLaunchKernelGGL((k_gram<T, NB, 0, true, true>), grid, dim3(256), red, st, D, items,
                                   (const T *)bulk, small, panel);
            else
                hipLaunchKernelGGL((k_gram<T, NB, 0, false, true>), grid, dim3(256), red, st, D, items,
                                   (const T *)bulk, small, panel);
            return;
        }
    }
    // Y and Z columns are padded (ld a multiple of SSA_PANEL_ALIGN): branch-free loads
    const bool pad = panel == SSA_PANEL_Y || panel == SSA_PANEL_Z;
    if (py > 1) {
        if (use_m)
            hipLaunchKernelGGL((k_gram<T, NB, 1, true>), grid, dim3(256), red > ms ? red : ms, st, D, items,
                               (const T *)bulk, small, panel);
        else if (pad)
            hipLaunchKernelGGL((k_gram<T, NB, 0, true, false, true>), grid, dim3(256), red, st, D, items,
                               (const T *)bulk, small, panel);
        else
            hipLaunchKernelGGL((k_gram<T, NB, 0, true>), grid, dim3(256), red, st, D, items, (const T *)bulk,
                               small, panel);
    } else if (use_m)
        hipLaunchKernelGGL((k_gram<T, NB, 1, false>), grid, dim3(256), red > ms ? red : ms, st, D, items,
                           (const T *)bulk, small, panel);
    else if (pad)
        hipLaunchKernelGGL((k_gram<T, NB, 0, false, false, true>), grid, dim3(256), red, st, D, items,
                           (const T *)bulk, small, panel);
    else
        hipLaunchKernelGGL((k_gram<T, NB, 0, false>), grid, dim3(256), red, st, D, items, (const T *)bulk, small,
                           panel);
}
}  // extern "C++"

int ssa_dev_gram(int prec, int panel, int use_m, int nb, const ssa_desc *D, const ssa_item *items, int n_items,
                 void *bulk, double *small, void *stream)
{
    if (n_items <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
#define G(NB)                                                                                   \
    case NB:                                                                                    \
        if (prec == SSA_PREC_F64) launch_gram<double, NB>(panel, use_m, D, items, n_items, bulk, small, st); \
        else launch_gram<float, NB>(panel, use_m, D, items, n_items, bulk, small, st);          \
        break;
    switch (nb) {
        G(1) G(2) G(3) G(4) G(5)
    default: return -1;
    }
#undef G
    return (int)hipGetLastError();
}

int ssa_dev_chol(int panel, int pass, int bmax, double tol, const ssa_desc *D, const ssa_item *items, int n_items,
                 double *small, void *stream)
{
    if (n_items <= 0) return 0;
    const size_t lds = 2 * (size_t)bmax * (bmax + 1) * sizeof(double);
    static int attr = 0;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)k_chol, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(2 * SSA_BMAX * LDP * sizeof(double)));
        if (e != hipSuccess) return (int)e;
        attr = 1;
    }
    if (bmax <= 64) {
        // one wave per series, the matrix in registers
        hipStream_t st = (hipStream_t)stream;
        switch (SSA_NB(bmax)) {
        case 1: hipLaunchKernelGGL(k_chol_w<1>, dim3(n_items), dim3(64), 0, st, D, items, small, panel, pass, tol); break;
        case 2: hipLaunchKernelGGL(k_chol_w<2>, dim3(n_items), dim3(64), 0, st, D, items, small, panel, pass, tol); break;
        case 3: hipLaunchKernelGGL(k_chol_w<3>, dim3(n_items), dim3(64), 0, st, D, items, small, panel, pass, tol); break;
        default: hipLaunchKernelGGL(k_chol_w<4>, dim3(n_items), dim3(64), 0, st, D, items, small, panel, pass, tol); break;
        }
        return (int)hipGetLastError();
    }
    hipLaunchKernelGGL(k_chol, dim3(n_items), dim3(256), lds, (hipStream_t)stream, D, items, small, panel, pass,
                       tol);
    return (int)hipGetLastError();
}

extern "C++" {
template <typename T, int NB>
static void launch_apply(int src, int dst, int slot, int ncols_k, const ssa_desc *D, const ssa_item *items,
                         int n_items, void *bulk, const double *small, hipStream_t st)
{
    // FP32 panels take the FP32 MFMA unless SSA_AMD_APPLY_F64=1 asks for FP64 accumulation
    static const int apply_f64 = [] {
        const char *v = getenv("SSA_AMD_APPLY_F64");
        return v && atoi(v) != 0;
    }();
    // a launch of few row tiles (short series) splits each tile over up to 8
    // workgroups of 256 rows, so that more CUs share it
    int py = 1;
    while (py < 8 && n_items * py * 2 <= 1024) py *= 2;
    const dim3 grid(n_items, py);
    if (sizeof(T) == 4 && !apply_f64) {
        const size_t lds = (size_t)NB * 16 * small_ld<NB>() * sizeof(float);
        hipLaunchKernelGGL((k_apply_f32<NB>), grid, dim3(64 * APPLY_WAVES), lds, st, D, items,
                           (float *)bulk, small, src, dst, slot, ncols_k);
        return;
    }
    const size_t lds = (size_t)NB * 16 * small_ld<NB>() * sizeof(double);
    hipLaunchKernelGGL((k_apply<T, NB>), grid, dim3(64 * APPLY_WAVES), lds, st, D, items, (T *)bulk, small,
                       src, dst, slot, ncols_k);
}
}  // extern "C++"

int ssa_dev_apply(int prec, int src, int dst, int slot, int ncols_k, int nb, const ssa_desc *D,
                  const ssa_item *items, int n_items, void *bulk, double *small, void *stream)
{
    if (n_items <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
#define A(NB)                                                                                                  \
    case NB:                                                                                                   \
        if (prec == SSA_PREC_F64) launch_apply<double, NB>(src, dst, slot, ncols_k, D, items, n_items, bulk, small, st); \
        else launch_apply<float, NB>(src, dst, slot, ncols_k, D, items, n_items, bulk, small, st);             \
        break;
    switch (nb) {
        A(1) A(2) A(3) A(4) A(5)
    default: return -1;
    }
#undef A
    return (int)hipGetLastError();
}

int ssa_dev_trmm(int nb, int slot_a, int slot_b, int slot_c, const ssa_desc *D, const ssa_item *items, int n_items,
                 double *small, void *stream)
{
    if (n_items <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    switch (nb) {
    case 1: hipLaunchKernelGGL(k_trmm<1>, dim3(n_items), dim3(256), 0, st, D, items, small, slot_a, slot_b, slot_c); break;
    case 2: hipLaunchKernelGGL(k_trmm<2>, dim3(n_items), dim3(256), 0, st, D, items, small, slot_a, slot_b, slot_c); break;
    case 3: hipLaunchKernelGGL(k_trmm<3>, dim3(n_items), dim3(256), 0, st, D, items, small, slot_a, slot_b, slot_c); break;
    case 4: hipLaunchKernelGGL(k_trmm<4>, dim3(n_items), dim3(256), 0, st, D, items, small, slot_a, slot_b, slot_c); break;
    case 5: hipLaunchKernelGGL(k_trmm<5>, dim3(n_items), dim3(256), 0, st, D, items, small, slot_a, slot_b, slot_c); break;
    default: return -1;
    }
    return (int)hipGetLastError();
}

int ssa_dev_svd(int bmax, int r_slot, int rinv_slot, double early, const ssa_desc *D, const ssa_item *items,
                int n_items, double *small, void *stream)
{
    if (n_items <= 0) return 0;
    // 16 NB columns of 16 NB + 8 doubles
    const size_t lds = (size_t)16 * SSA_NB(bmax) * (16 * SSA_NB(bmax) + 8) * sizeof(double);
    hipStream_t st = (hipStream_t)stream;
    switch (SSA_NB(bmax)) {
    case 1: hipLaunchKernelGGL(k_svd<1>, dim3(n_items), dim3(256), lds, st, D, items, small, r_slot, rinv_slot, early); break;
    case 2: hipLaunchKernelGGL(k_svd<2>, dim3(n_items), dim3(256), lds, st, D, items, small, r_slot, rinv_slot, early); break;
    case 3: hipLaunchKernelGGL(k_svd<3>, dim3(n_items), dim3(256), lds, st, D, items, small, r_slot, rinv_slot, early); break;
    case 4: hipLaunchKernelGGL(k_svd<4>, dim3(n_items), dim3(256), lds, st, D, items, small, r_slot, rinv_slot, early); break;
    case 5: hipLaunchKernelGGL(k_svd<5>, dim3(n_items), dim3(256), lds, st, D, items, small, r_slot, rinv_slot, early); break;
    default: return -1;
    }
    return (int)hipGetLastError();
}

int ssa_dev_recon(int prec, int lgn, int scale_sigma, int limit_gsize, int nsplit, const ssa_desc *D,
                  const ssa_item *items, int n_items, const int *gdesc, const int *glist, void *bulk,
                  const double *small, const void *tw, void *stream)
{
    if (n_items <= 0) return 0;
    if (!lgn_ok(prec, lgn)) return -2;
    hipStream_t st = (hipStream_t)stream;
#define CASE(L)                                                                                               \
    case L:                                                                                                   \
        if (prec == SSA_PREC_F64)                                                                             \
            hipLaunchKernelGGL((k_recon<double, L>), dim3(xcd_grid(n_items)), dim3(FftGeom<double, L>::NT), 0, \
                               st, D, items, (double *)bulk, small, (const int4 *)gdesc, glist,               \
                               (const C2<double> *)tw, scale_sigma, limit_gsize, nsplit, n_items);                    \
        else                                                                                                  \
            hipLaunchKernelGGL((k_recon<float, L>), dim3(xcd_grid(n_items)), dim3(FftGeom<float, L>::NT), 0,   \
                               st, D, items, (float *)bulk, small, (const int4 *)gdesc, glist,                \
                               (const C2<float> *)tw, scale_sigma, limit_gsize, nsplit, n_items);                    \
        break;
    switch (lgn) { SSA_LGN_CASES(CASE) default: return -2; }
#undef CASE
    return (int)hipGetLastError();
}

int ssa_dev_olap(int prec, int weight, int nsplit, int tile, const ssa_desc *D, const ssa_item *items,
                 int n_items, const int *gdesc, void *bulk, void *stream)
{
    if (n_items <= 0) return 0;
    if (tile < 256 || tile % 256) return -1;
    // a launch of few items reads every piece ahead (its time is one item's latency)
    const bool all = n_items < 1024;
    const hipStream_t st = (hipStream_t)stream;
    if (prec == SSA_PREC_F64) {
        if (all) hipLaunchKernelGGL((k_olap<double, true>), dim3(n_items), dim3(256), 0, st, D, items,
                                    (const int4 *)gdesc, (double *)bulk, weight, nsplit, tile);
        else hipLaunchKernelGGL((k_olap<double, false>), dim3(n_items), dim3(256), 0, st, D, items,
                                (const int4 *)gdesc, (double *)bulk, weight, nsplit, tile);
    } else {
        if (all) hipLaunchKernelGGL((k_olap<float, true>), dim3(n_items), dim3(256), 0, st, D, items,
                                    (const int4 *)gdesc, (float *)bulk, weight, nsplit, tile);
        else hipLaunchKernelGGL((k_olap<float, false>), dim3(n_items), dim3(256), 0, st, D, items,
                                (const int4 *)gdesc, (float *)bulk, weight, nsplit, tile);
    }
    return (int)hipGetLastError();
}

int ssa_dev_cache_build(int prec, int lgn, const ssa_desc *D, const ssa_item *items, int n_items, void *bulk,
                        const void *tw, void *stream)
{
    if (n_items <= 0) return 0;
    if (!lgn_ok(prec, lgn)) return -2;
    hipStream_t st = (hipStream_t)stream;
#define CASE(L)                                                                                               \
    case L:                                                                                                   \
        if (prec == SSA_PREC_F64)                                                                             \
            hipLaunchKernelGGL((k_cache_build<double, L>), dim3(n_items), dim3(FftGeom<double, L>::NT), 0, st, \
                               D, items, (double *)bulk, (const C2<double> *)tw);                             \
        else                                                                                                  \
            hipLaunchKernelGGL((k_cache_build<float, L>), dim3(n_items), dim3(FftGeom<float, L>::NT), 0, st,   \
                               D, items, (float *)bulk, (const C2<float> *)tw);                               \
        break;
    switch (lgn) { SSA_LGN_CASES(CASE) default: return -2; }
#undef CASE
    return (int)hipGetLastError();
}

int ssa_dev_recon_cached(int prec, int lgn, int scale_sigma, const ssa_desc *D, const ssa_item *items, int n_items,
                         int n_pieces, const int *gdesc, const int *glist, void *bulk, const double *small,
                         const void *tw, void *stream)
{
    if (n_items <= 0) return 0;
    if (!lgn_ok(prec, lgn)) return -2;
    hipStream_t st = (hipStream_t)stream;
#define CASE(L)                                                                                               \
    case L:                                                                                                   \
        if (prec == SSA_PREC_F64)                                                                             \
            hipLaunchKernelGGL((k_recon_cached<double, L>), dim3(xcd_grid(n_items)), dim3(FftGeom<double, L>::NT), \
                               0, st, D, items, (double *)bulk, small, (const int4 *)gdesc, glist,            \
                               (const C2<double> *)tw, scale_sigma, n_pieces, n_items);                       \
        else                                                                                                  \
            hipLaunchKernelGGL((k_recon_cached<float, L>), dim3(xcd_grid(n_items)), dim3(FftGeom<float, L>::NT), \
                               0, st, D, items, (float *)bulk, small, (const int4 *)gdesc, glist,             \
                               (const C2<float> *)tw, scale_sigma, n_pieces, n_items);                        \
        break;
    switch (lgn) { SSA_LGN_CASES(CASE) default: return -2; }
#undef CASE
    return (int)hipGetLastError();
}

int ssa_dev_edge(int prec, int scale_sigma, int weight, int limit_gsize, const ssa_desc *D, const ssa_item *items,
                 int n_items, const int *gdesc, const int *glist, void *bulk, const double *small, void *stream)
{
    if (n_items <= 0) return 0;
    const hipStream_t st = (hipStream_t)stream;
    // a launch of few items spreads each item's components over four thread groups
#define EDGE_LAUNCH(T, P)                                                                                     \
    hipLaunchKernelGGL((k_edge<T, P>), dim3(n_items), dim3(P * 2 * SSA_EDGE), 0, st, D, items,                \
                       (const int4 *)gdesc, glist, (T *)bulk, small, scale_sigma, weight, limit_gsize)
    const bool few = n_items < 256;
    if (prec == SSA_PREC_F64) {
        if (few) EDGE_LAUNCH(double, 4);
        else EDGE_LAUNCH(double, 1);
    } else {
        if (few) EDGE_LAUNCH(float, 4);
        else EDGE_LAUNCH(float, 1);
    }
#undef EDGE_LAUNCH
    return (int)hipGetLastError();
}

int ssa_dev_gather(int prec, int h, const ssa_desc *D, const ssa_item *items, int n_items, void *bulk, void *stream)
{
    if (n_items <= 0) return 0;
    if (prec == SSA_PREC_F64)
        hipLaunchKernelGGL(k_gather<double>, dim3(n_items), dim3(256), 0, (hipStream_t)stream, D, items,
                           (double *)bulk, h);
    else
        hipLaunchKernelGGL(k_gather<float>, dim3(n_items), dim3(256), 0, (hipStream_t)stream, D, items,
                           (float *)bulk, h);
    return (int)hipGetLastError();
}

int ssa_dev_pack_results(int prec, const double *misc, int misc_n, const void *U, long long un, const void *V,
                         long long vn, void *out, void *stream)
{
    const long long n = (long long)misc_n + un + vn;
    if (misc_n < 0 || un < 0 || vn < 0 || n <= 0) return -1;
    const long long blocks = (n + 255) / 256;
    const dim3 grid((unsigned)(blocks < 256 ? blocks : 256));
    if (prec == SSA_PREC_F64)
        hipLaunchKernelGGL(k_pack_results<double>, grid, dim3(256), 0, (hipStream_t)stream, misc, misc_n,
                           (const double *)U, un, (const double *)V, vn, (char *)out);
    else
        hipLaunchKernelGGL(k_pack_results<float>, grid, dim3(256), 0, (hipStream_t)stream, misc, misc_n,
                           (const float *)U, un, (const float *)V, vn, (char *)out);
    return (int)hipGetLastError();
}

int ssa_dev_wcorr(int prec, int nb, const ssa_desc *D, const ssa_item *tile_items, int n_tiles,
                  const ssa_item *series_items, int n_series, void *bulk, double *small, void *stream)
{
    if (n_tiles <= 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    // the weighted Gram: FP64 MFMA for FP64 panels, FP32 MFMA with FP32
    // accumulation inside each row tile for the mixed form's FP32 panels
#define G(NB)                                                                                              \
    case NB:                                                                                               \
        if (prec == SSA_PREC_F64) launch_gram<double, NB>(SSA_PANEL_OUT, 0, D, tile_items, n_tiles, bulk, small, st); \
        else launch_gram<float, NB>(SSA_PANEL_OUT, 0, D, tile_items, n_tiles, bulk, small, st, true);    \
        break;
    switch (nb) {
        G(1) G(2) G(3) G(4) G(5)
    default: return -1;
    }
#undef G
    int e = (int)hipGetLastError();
    if (e) return e;
    if (n_series <= 0) return 0;
    hipLaunchKernelGGL(k_wnorm, dim3(n_series), dim3(256), 0, (hipStream_t)stream, D, series_items, small);
    return (int)hipGetLastError();
}

int ssa_dev_forecast(int prec, int h, int limit_gsize, int vmode, const ssa_desc *D, const ssa_item *items,
                     int n_items, const int *gdesc, const int *glist, void *bulk, double *small, void *stream)
{
    if (n_items <= 0) return 0;
    if (prec == SSA_PREC_F64)
        hipLaunchKernelGGL(k_forecast<double>, dim3(n_items), dim3(256), 0, (hipStream_t)stream,
                           (ssa_desc *)D, items, (const int4 *)gdesc, glist, (double *)bulk, small, h, limit_gsize,
                           vmode);
    else
        hipLaunchKernelGGL(k_forecast<float>, dim3(n_items), dim3(256), 0, (hipStream_t)stream,
                           (ssa_desc *)D, items, (const int4 *)gdesc, glist, (float *)bulk, small, h, limit_gsize,
                           vmode);
    return (int)hipGetLastError();
}

}  // extern "C"
