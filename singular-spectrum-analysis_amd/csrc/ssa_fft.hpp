// ssa_fft.hpp -- register-resident Stockham FFT for one workgroup (gfx950).
//
// A length-N complex transform is held as E = N/NT values per thread in the
// "natural" layout: thread t holds positions t + e*NT, e < E.  Each pass is a
// Stockham autosort step (Govindaraju et al., "High performance discrete
// Fourier transforms on graphics processors", SC'08): butterfly j reads
// positions j + r*N/R (which is exactly the natural layout), applies the
// twiddles w^(r*(j mod Ns)), an in-register DFT of size R, and writes
// (j/Ns)*Ns*R + (j mod Ns) + r*Ns.  Between passes the values are exchanged
// through LDS (re and im planes, optionally one plane at a time so the
// exchange needs half the LDS); the last pass lands in the natural layout
// without an exchange, so the result stays in registers for the caller's
// pointwise stage and the next transform.
//
// SIGN = -1 is the forward DFT sum_t z_t exp(-2 pi i k t/N), SIGN = +1 the
// unscaled inverse.  Twiddles come from tables built on the host in long
// double.  FP64 transforms read every twiddle from per-pass tables (exactly
// rounded, one load each); FP32 transforms read w, w^2, w^4, w^8 from a
// table exp(2 pi i m / 2^16) and form the others with at most three products
// (a few ulp), which costs fewer memory instructions than 15 loads.
#pragma once
#include <hip/hip_runtime.h>

namespace ssa {

template <typename T> struct C2 { T x, y; };

template <typename T> __device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> __device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// multiply by S*i
template <typename T, int S> __device__ __forceinline__ C2<T> muli(C2<T> a)
{
    return S > 0 ? C2<T>{-a.y, a.x} : C2<T>{a.y, -a.x};
}

// exp(S * 2 pi i * m / 16) for m = 0..7, exact-to-rounding constants
template <typename T, int S> __device__ __forceinline__ C2<T> w16(int m)
{
    const T c1 = (T)0.92387953251128675613, s1 = (T)0.38268343236508977173, h = (T)0.70710678118654752440;
    T c = 0, s = 0;
    switch (m) {
    case 0: c = 1; s = 0; break;
    case 1: c = c1; s = s1; break;
    case 2: c = h; s = h; break;
    case 3: c = s1; s = c1; break;
    case 4: c = 0; s = 1; break;
    case 5: c = -s1; s = c1; break;
    case 6: c = -h; s = h; break;
    default: c = -c1; s = s1; break;
    }
    return {c, (T)S * s};
}

// exp(S * 2 pi i * m / 32) for m = 0..15 (even m from w16)
template <typename T, int S> __device__ __forceinline__ C2<T> w32(int m)
{
    if (m % 2 == 0) return w16<T, S>(m / 2);
    const T a1 = (T)0.98078528040323044913, b1 = (T)0.19509032201612826785;   // cos, sin of pi/16
    const T a3 = (T)0.83146961230254523708, b3 = (T)0.55557023301960222474;   // cos, sin of 3 pi/16
    T c = 0, s = 0;
    switch (m) {
    case 1: c = a1; s = b1; break;
    case 3: c = a3; s = b3; break;
    case 5: c = b3; s = a3; break;
    case 7: c = b1; s = a1; break;
    case 9: c = -b1; s = a1; break;
    case 11: c = -b3; s = a3; break;
    case 13: c = -a3; s = b3; break;
    default: c = -a1; s = b1; break;
    }
    return {c, (T)S * s};
}

// in-register DFTs, natural order in and out
template <typename T, int S> __device__ __forceinline__ void dft2(C2<T> *v)
{
    C2<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename T, int S> __device__ __forceinline__ void dft4(C2<T> &a0, C2<T> &a1, C2<T> &a2, C2<T> &a3)
{
    C2<T> s02 = cadd(a0, a2), d02 = csub(a0, a2), s13 = cadd(a1, a3), d13 = muli<T, S>(csub(a1, a3));
    a0 = cadd(s02, s13);
    a2 = csub(s02, s13);
    a1 = cadd(d02, d13);
    a3 = csub(d02, d13);
}

template <typename T, int S> __device__ __forceinline__ void dft8(C2<T> *v)
{
    C2<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C2<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<T, S>(e0, e1, e2, e3);
    dft4<T, S>(o0, o1, o2, o3);
    const T h = (T)0.70710678118654752440;
    C2<T> t1 = cmul(o1, C2<T>{h, (T)S * h});
    C2<T> t2 = muli<T, S>(o2);
    C2<T> t3 = cmul(o3, C2<T>{-h, (T)S * h});
    v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
    v[1] = cadd(e1, t1); v[5] = csub(e1, t1);
    v[2] = cadd(e2, t2); v[6] = csub(e2, t2);
    v[3] = cadd(e3, t3); v[7] = csub(e3, t3);
}

template <typename T, int S> __device__ __forceinline__ void dft16(C2<T> *v)
{
    C2<T> e[8], o[8];
#pragma unroll
    for (int i = 0; i < 8; i++) { e[i] = v[2 * i]; o[i] = v[2 * i + 1]; }
    dft8<T, S>(e);
    dft8<T, S>(o);
#pragma unroll
    for (int m = 0; m < 8; m++) {
        // m = 4 is S*i: a swap and a negation, the same values as the product
        C2<T> t = (m == 0) ? o[0] : (m == 4) ? muli<T, S>(o[4]) : cmul(o[m], w16<T, S>(m));
        v[m] = cadd(e[m], t);
        v[m + 8] = csub(e[m], t);
    }
}

template <typename T, int S> __device__ __forceinline__ void dft32(C2<T> *v)
{
    C2<T> e[16], o[16];
#pragma unroll
    for (int i = 0; i < 16; i++) { e[i] = v[2 * i]; o[i] = v[2 * i + 1]; }
    dft16<T, S>(e);
    dft16<T, S>(o);
#pragma unroll
    for (int m = 0; m < 16; m++) {
        C2<T> t = (m == 0) ? o[0] : (m == 8) ? muli<T, S>(o[8]) : cmul(o[m], w32<T, S>(m));
        v[m] = cadd(e[m], t);
        v[m + 16] = csub(e[m], t);
    }
}

template <typename T, int S, int R> __device__ __forceinline__ void dftR(C2<T> *v)
{
    if constexpr (R == 2) dft2<T, S>(v);
    else if constexpr (R == 4) dft4<T, S>(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 8) dft8<T, S>(v);
    else if constexpr (R == 16) dft16<T, S>(v);
    else if constexpr (R == 32) dft32<T, S>(v);
}

// LDS index of element p in the exchange that follows a pass of stride
// NS = 2^NSLOG and radix R = 2^RLOG.  The scatter writes groups of NS
// consecutive positions R*NS apart; the gather reads consecutive positions.
// XOR-ing the group bits above max(LW, NSLOG + RLOG) into bits [NSLOG, LW)
// spreads both over all 2^LW banks an element width covers (32 four-byte
// banks: LW = 5 for 4-byte values, 4 for 8-byte values).  The map is a
// bijection inside each block of 2^LW elements, so no padding is needed.
template <typename T, int NSLOG, int RLOG> __device__ __forceinline__ int lswz(int p)
{
    constexpr int LW = sizeof(T) == 4 ? 5 : 4;
    constexpr int S = (NSLOG + RLOG > LW) ? NSLOG + RLOG : LW;
    if constexpr (NSLOG >= LW) return p;
    else return p ^ (((p >> S) << NSLOG) & ((1 << LW) - 1));
}

// E = 8 values per thread (radix-8 passes), so a transform plus an
// accumulator plus the butterfly temporaries fit the 128-VGPR budget of
// 4 waves per SIMD; 1024 threads cover N = 8192.
constexpr int TW_BASE_LOG = 16;   // FP32 base table exp(2 pi i m / 2^16)

// Per-pass twiddle tables (FP64): a pass of stride NS = 2^NSLOG and radix
// R = 2^RLOG reads entry tw_offset(NSLOG, RLOG) + (r-1)*NS + kk =
// exp(+2 pi i r kk / (NS R)) for butterfly kk < NS and r = 1..R-1 (r-major,
// so consecutive lanes read consecutive entries).  Tables for RLOG = 1..4
// and NSLOG = 0..TW_NSLOG_MAX, RLOG-major.
constexpr int TW_NSLOG_MAX = 12;
__host__ __device__ constexpr long long tw_offset(int nslog, int rlog)
{
    long long off = 0;
    for (int r = 1; r < rlog; r++) off += ((1LL << (TW_NSLOG_MAX + 1)) - 1) * ((1LL << r) - 1);
    return off + ((1LL << nslog) - 1) * ((1LL << rlog) - 1);
}
constexpr long long TW_ENTRIES = tw_offset(0, 5);

template <typename T, int LOGN, bool WIDE = false> struct FftGeom {
    static constexpr int N = 1 << LOGN;
    // WIDE, FP32 at N = 8192: 32 values per thread, a radix-32 pass then two
    // radix-16 passes (two LDS exchanges instead of three), at half the waves;
    // WIDE, FP64 at 1024 <= N <= 4096: radix-16 passes at 2 waves per SIMD
    static constexpr int E = (WIDE && sizeof(T) == 4 && LOGN == 13) ? 32 : (sizeof(T) == 4 && LOGN >= 10) || (WIDE && LOGN >= 10 && LOGN <= 12) ? 16 : N >= 16 ? 8 : N / 2;
    static constexpr int NT = N / E;
    static constexpr int LOGE = (E == 32) ? 5 : (E == 16) ? 4 : (E == 8) ? 3 : (E == 4) ? 2 : 1;
    static constexpr int NEXT = (E == 32) ? 4 : LOGE;                     // radix log of passes 1..
    static constexpr int FIRST = (E == 32) ? LOGN - 8 : (LOGN % LOGE) ? (LOGN % LOGE) : LOGE;   // radix log of pass 0
    // waves per SIMD the transform kernels are built for (register budget)
    static constexpr int OCC = (E == 32 || (E == 16 && sizeof(T) == 8)) ? 2 : (E == 8 && sizeof(T) == 4) ? 8 : 4;
    static constexpr int LDS = N;                                        // elements per plane
    // exchange one plane at a time when both planes would keep a second
    // workgroup off the CU (160 KB LDS)
    static constexpr bool SPLIT = 2 * LDS * sizeof(T) > 80000;
};

template <typename T, int LOGN, bool WIDE = false> using FftRegs = C2<T>[FftGeom<T, LOGN, WIDE>::E];

// FP32 twiddle base entries of a pass of stride NS = 2^NSLOG and radix
// R = 2^RLOG: w^(2^b), b < RLOG, of butterfly m of this thread, sign applied.
template <typename T, int SIGN, int NSLOG, int RLOG, int BPT, int NT>
__device__ __forceinline__ void tw_fetch(C2<T> (&wp)[BPT][4], const C2<T> *tw, int tid)
{
    constexpr int R = 1 << RLOG, NS = 1 << NSLOG;
#pragma unroll
    for (int m = 0; m < BPT; m++) {
        const int kk = (tid + m * NT) & (NS - 1);
        const int step = kk << (TW_BASE_LOG - NSLOG - RLOG);
#pragma unroll
        for (int b = 0; (1 << b) < R; b++) {
            wp[m][b] = tw[(1 << b) * step];
            if constexpr (SIGN < 0) wp[m][b].y = -wp[m][b].y;
        }
    }
}

// One pass.  wp: for FP32 passes with NS > 1, this pass's twiddle base entries,
// fetched by the previous pass before its LDS exchange so that their latency
// overlaps the exchange (unused otherwise).
template <typename T, int LOGN, int SIGN, bool SPLIT, int NSLOG, int RLOG, bool WIDE = false>
__device__ __forceinline__ void fft_pass(FftRegs<T, LOGN, WIDE> &c, T *lre, T *lim, const C2<T> *tw, int tid,
                                         const C2<T> (&wp)[FftGeom<T, LOGN, WIDE>::E / (1 << RLOG)][4])
{
    using G = FftGeom<T, LOGN, WIDE>;
    constexpr int N = G::N, NT = G::NT, E = G::E;
    constexpr int R = 1 << RLOG, NS = 1 << NSLOG, BPT = E / R;
#pragma unroll
    for (int m = 0; m < BPT; m++) {
        const int j = tid + m * NT;
        C2<T> v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = c[m + r * BPT];
        if constexpr (NS > 1 && sizeof(T) == 8) {
            static_assert(NSLOG <= TW_NSLOG_MAX, "twiddle table too short");
            const int kk = j & (NS - 1);
            const C2<T> *tp = tw + tw_offset(NSLOG, RLOG) + kk;
#pragma unroll
            for (int r = 1; r < R; r++) {
                C2<T> w = tp[(r - 1) * NS];
                if constexpr (SIGN < 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        } else if constexpr (NS > 1) {
            // w^r for r < R from the base-table entries w^1, w^2, w^4, w^8
#pragma unroll
            for (int r = 1; r < R; r++) {
                C2<T> w = C2<T>{1, 0};
                bool first = true;
#pragma unroll
                for (int b = 0; (1 << b) < R; b++)
                    if (r & (1 << b)) {
                        w = first ? wp[m][b] : cmul(w, wp[m][b]);
                        first = false;
                    }
                v[r] = cmul(v[r], w);
            }
        }
        dftR<T, SIGN, R>(v);
#pragma unroll
        for (int r = 0; r < R; r++) c[m + r * BPT] = v[r];
    }
    if constexpr (NSLOG + RLOG < LOGN) {
        constexpr int NEXT = G::NEXT;
        C2<T> wn[E >> NEXT][4];
        if constexpr (sizeof(T) == 4) tw_fetch<T, SIGN, NSLOG + RLOG, NEXT, (E >> NEXT), NT>(wn, tw, tid);
        // scatter to Stockham destinations, gather natural layout
        constexpr int PLANES = SPLIT ? 2 : 1;
#pragma unroll
        for (int ph = 0; ph < PLANES; ph++) {
#pragma unroll
            for (int m = 0; m < BPT; m++) {
                const int j = tid + m * NT;
                const int base = ((j >> NSLOG) << (NSLOG + RLOG)) + (j & (NS - 1));
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const int p = lswz<T, NSLOG, RLOG>(base + r * NS);
                    if (SPLIT) {
                        if (ph == 0) lre[p] = c[m + r * BPT].x;
                        else lre[p] = c[m + r * BPT].y;
                    } else {
                        lre[p] = c[m + r * BPT].x;
                        lim[p] = c[m + r * BPT].y;
                    }
                }
            }
            __syncthreads();
#pragma unroll
            for (int e = 0; e < E; e++) {
                const int p = lswz<T, NSLOG, RLOG>(tid + e * NT);
                if (SPLIT) {
                    if (ph == 0) c[e].x = lre[p];
                    else c[e].y = lre[p];
                } else {
                    c[e].x = lre[p];
                    c[e].y = lim[p];
                }
            }
            __syncthreads();
        }
        fft_pass<T, LOGN, SIGN, SPLIT, NSLOG + RLOG, NEXT, WIDE>(c, lre, lim, tw, tid, wn);
    }
}

// Full transform of the natural-layout registers c (in place).
template <typename T, int LOGN, int SIGN, bool SPLIT, bool WIDE = false>
__device__ __forceinline__ void fft(FftRegs<T, LOGN, WIDE> &c, T *lre, T *lim, const C2<T> *tw)
{
    // The twiddles and LDS addresses depend only on the thread index, so the
    // compiler would hoist those of every pass out of the callers' loops and
    // keep them all live (register spills).  An opaque copy of the thread
    // index makes each transform recompute its own.  (The table pointer stays
    // the kernel argument: an opaque copy of it would lose its global address
    // space, and flat loads also count against the LDS wait counter.)
    int tid = threadIdx.x;
    const C2<T> *t = tw;
    asm volatile("" : "+v"(tid));
    using G = FftGeom<T, LOGN, WIDE>;
    const C2<T> none[G::E >> G::FIRST][4] = {};   // the first pass (stride 1) has no twiddles
    fft_pass<T, LOGN, SIGN, SPLIT, 0, G::FIRST, WIDE>(c, lre, lim, t, tid, none);
}

}  // namespace ssa
