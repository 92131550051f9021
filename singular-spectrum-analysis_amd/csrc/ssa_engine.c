/*
 * ssa_engine.c -- host C engine of the device-resident SSA path.
 *
 * Sequence of one decomposition (every series of the engine in lockstep),
 * restating MKL/ssa_opt.h:1685-1763 with q subspace iterations:
 *
 *   Y = H Omega                     k_matvec (Omega drawn on the fly or injected)
 *   Q = orth(Y)                     CholeskyQR2: gram, chol, gram(Y R1^-1), chol, apply
 *   repeat q: Z = orth(H^T Q); Q = orth(H Z)
 *   B = H^T Q                       k_matvec
 *   R_B = chol-QR2 factor of B      gram, chol, gram, chol (no Q_B needed)
 *   R_B = Ur S W^T                  k_svd (one-sided Jacobi, sign fix)
 *   U = Q W_k, V = B W_k S^-1       k_apply
 *
 * The bulk arena offsets of a descriptor are element offsets from the base
 * pointer passed to the kernels (the arena of the signals); the panel arena
 * is a separate allocation addressed through the same base.
 */
#include "ssa_engine.h"

#include <hip/hip_runtime_api.h>
#include <limits.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NLG 14                 /* lgn buckets 0..13 (6..13 used) */
#define NBW 6                  /* Gram / apply lists by column blocks nb = 1..5 */
#define OLAP_TILE_STEP 1024   /* overlap-add samples per work item in the batched step */
#define RESULTS_DIRECT_MAX (1 << 20)   /* result bytes up to which the device writes them to the host */

typedef struct { long long off; int n; } span;

struct ssa_engine {
    int prec, count;
    size_t esz;
    ssa_desc *hd;
    int *bmax, *kmax;
    ssa_desc *dd;
    void *arena_a;             /* signals + spectra (base pointer of all offsets) */
    void *arena_b;             /* panels, staging, outputs */
    long long a_elems, b_elems;
    long long a_cap, b_cap, small_cap;   /* allocated elements (rebind reuses them) */
    int dd_cap, host_cap;                /* allocated descriptors (device, host) */
    int q_hint;
    double *small;
    long long small_elems;
    void *tw;
    hipStream_t st;
    int prepared, max_groups, max_h;
    /* work lists */
    ssa_item *ditems;
    long long ditems_cap;
    ssa_item *hitems;
    long long hitems_n, hitems_cap;
    span spec[NLG], mvf[NLG], mva[NLG], tY[NBW], tZ[NBW], ser, ser_nb[NBW];   /* tY/tZ/ser_nb by nb */
    span stgS[2][NLG], stgA[2][NLG];   /* staged block matvec (by direction): spectra, sums */
    span rec_step[NLG], olap_step;
    int lists_dirty, desc_dirty;
    /* >= 0: the test-matrix counter of a one-series engine whose descriptor
     * is otherwise unchanged since its last upload.  An eager decomposition
     * passes it to the test-matrix kernel instead of uploading the
     * descriptor again (a host-to-device copy per call, whose latency varies
     * with the copy engine's state); a batched step uploads it. */
    long long omega_call;
    int dyn_base;              /* first item of the per-call region */
    /* groups (int4 descriptors followed by component lists) */
    int *dgroups;
    long long dgroups_cap;
    /* host copies of what the dynamic item region and the group buffer last
     * received: a call that would upload the same content skips the copy and
     * its synchronisation (repeated reconstructions of one group) */
    ssa_item *dyn_last;
    long long dyn_last_n, dyn_last_cap, dyn_last_from;
    int *grp_last;
    long long grp_last_n, grp_last_cap;
    int step_groups_ready;
    /* pinned host staging */
    void *hpin;
    size_t hpin_bytes;
    /* page-locked, coherent host buffer that the device writes a small
     * decomposition's results into (ssa_engine_results), and its device address */
    void *hres, *hres_dev;
    size_t hres_bytes;
    /* pinned staging of the descriptor upload, which is not waited for: the
     * event marks when the copy has read it, before the next upload refills it */
    void *dpin;
    size_t dpin_bytes;
    hipEvent_t dpin_ev;
    int dpin_busy;
    /* graph of the batched step */
    hipGraphExec_t gexec;
    int graph_q;
    int has_graph;
    /* per-launch profiler (HIP events on the engine stream) */
    int prof_on, prof_n, prof_cap;
    int step_h;                /* forecast horizon of the batched step (0: none) */
    int cholqr2_all;           /* CholeskyQR2 for every orthonormalisation (default: last one only) */
    int final2;                /* CholeskyQR2 for the last orthonormalisation of Y and for B */
    int mv_wide_min;           /* block-matvec launches of at least this many items take the wide geometry */
    int mv_stage;              /* the staged block matvec is used (SSA_AMD_MV_STAGED=0: the fused form only) */
    double dead_tol;           /* Cholesky pivot / original diagonal at or below which a column is dependent */
    /* component spectra cache (ssa_engine_cache_build): its own allocation,
     * addressed from arena_a like the panel arena */
    void *cache;
    long long cache_cap;       /* allocated T elements */
    int cache_valid;           /* built for the current decomposition of every series */
    hipEvent_t *prof_ev;
    int *prof_phase, *prof_lgn;
};

/* profiler phases */
enum { PH_SPECTRA = 0, PH_MV_FWD, PH_MV_ADJ, PH_GRAM, PH_CHOL, PH_APPLY, PH_SVD, PH_UV, PH_RECON, PH_OLAP, PH_FORECAST,
       PH_OMEGA, PH_N };

static void prof_begin(struct ssa_engine *e, int phase, int lgn)
{
    if (!e->prof_on || e->prof_n >= e->prof_cap) return;
    e->prof_phase[e->prof_n] = phase;
    e->prof_lgn[e->prof_n] = lgn;
    hipEventRecord(e->prof_ev[2 * e->prof_n], e->st);
}

static void prof_end(struct ssa_engine *e)
{
    if (!e->prof_on || e->prof_n >= e->prof_cap) return;
    hipEventRecord(e->prof_ev[2 * e->prof_n + 1], e->st);
    e->prof_n++;
}

/* ------------------------------------------------------------------------ */
/* planner                                                                   */
/* ------------------------------------------------------------------------ */

static int ceil_div(int a, int b) { return (a + b - 1) / b; }

/* Measured device time of one transform of length 2^lgn inside a full block
 * matvec launch (ns, whole-device throughput), per precision, set by
 * ssa_engine_autotune; all zero: the n (log2 n + 3) model. */
static double g_tcost[2][NLG];
static int g_tuned[2];
static int g_force_lgn = -1;   /* autotune: every series planned at this length */

static double plan_cost(int prec, int n, int lgn, int M, int AL, int SK, int b, int k, int q)
{
    /* transforms per column pair and block matvec, as the fused form does
     * them: n_out (n_in + 1), or 1 + n_out when one input chunk is shared.
     * The staged form does n_in + n_out, but writes the input spectra and
     * reads them back once per output chunk; priced by its transform count
     * alone it moved series to smaller lengths with more chunks, and ran
     * slower, so a plan is priced as the fused form and the staged form only
     * takes work off the plans chosen this way. */
    int nLc = M * AL;
    double fwd = (SK == 1) ? 1.0 + nLc : (double)nLc * (SK + 1);
    double adj = (nLc == 1) ? 1.0 + SK : (double)SK * (nLc + 1);
    double pairs = ceil_div(b > 0 ? b : 2, 2);
    double transforms = (q + 1) * pairs * (fwd + adj) + (double)AL * SK * (k + 1);
    const double per = g_tuned[prec] ? g_tcost[prec][lgn] : (double)n * (lgn + 3);
    return transforms * per;
}

int ssa_plan_series(int prec, int N, int L, int M, int b, int k, int q, int lgn_force, ssa_plan *out)
{
    if (prec != SSA_PREC_F64) prec = SSA_PREC_F32;
    if (g_force_lgn >= 6) lgn_force = g_force_lgn;
    int K = N - L + 1;
    ssa_plan best;
    best.cost = -1;
    for (int lgn = 6; lgn <= 13; lgn++) {
        if (lgn_force >= 6 && lgn_force <= 13 && lgn != lgn_force) continue;
        int n = 1 << lgn;
        int cand[3][2];
        int nc = 0;
        if (L + 1 <= n) { cand[nc][0] = L; cand[nc][1] = n - L + 1; nc++; }
        if (K + 1 <= n) { cand[nc][0] = n - K + 1; cand[nc][1] = K; nc++; }
        cand[nc][0] = n / 2; cand[nc][1] = n / 2; nc++;
        for (int c = 0; c < nc; c++) {
            int CL = cand[c][0] < L ? cand[c][0] : L;
            int CK = cand[c][1] < K ? cand[c][1] : K;
            if (CL < 1 || CK < 1 || CL + CK - 1 > n) continue;
            int AL = ceil_div(L, CL), SK = ceil_div(K, CK);
            double cost = plan_cost(prec, n, lgn, M, AL, SK, b, k, q);
            if (best.cost < 0 || cost < best.cost) {
                best.n = n; best.lgn = lgn; best.CL = CL; best.CK = CK; best.AL = AL; best.SK = SK;
                best.cost = cost;
            }
        }
    }
    if (best.cost < 0) return -1;
    *out = best;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* helpers                                                                   */
/* ------------------------------------------------------------------------ */

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
    fprintf(stderr, "ssa_engine: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
    return -1; } } while (0)
#define DEVCHK(x) do { int e_ = (x); if (e_ != 0) { \
    fprintf(stderr, "ssa_engine: %s failed: %d (%s:%d)\n", #x, e_, __FILE__, __LINE__); return -1; } } while (0)

static long long align64(long long v) { return (v + 63) & ~63LL; }

/* Y and Z panel columns start on 128-byte lines: the panel kernels' row
 * segments then never straddle two lines */
static int panel_ld(int rows) { return (rows + SSA_PANEL_ALIGN - 1) / SSA_PANEL_ALIGN * SSA_PANEL_ALIGN; }


/* the coherent page-locked buffer the device writes small results into */
static int hres_reserve(ssa_engine *e, size_t bytes)
{
    if (e->hres_bytes >= bytes) return 0;
    if (e->hres) HIPCHK(hipHostFree(e->hres));
    e->hres = e->hres_dev = NULL;
    e->hres_bytes = 0;
    const size_t nb = bytes < (64 << 10) ? (64 << 10) : bytes;
    HIPCHK(hipHostMalloc(&e->hres, nb, hipHostMallocCoherent));
    e->hres_bytes = nb;
    HIPCHK(hipHostGetDevicePointer(&e->hres_dev, e->hres, 0));
    return 0;
}

static int pin_reserve(ssa_engine *e, size_t bytes)
{
    if (e->hpin_bytes >= bytes) return 0;
    if (e->hpin) hipHostFree(e->hpin);
    e->hpin = NULL;
    size_t nb = bytes < (1 << 20) ? (1 << 20) : bytes;
    HIPCHK(hipHostMalloc(&e->hpin, nb, hipHostMallocDefault));
    e->hpin_bytes = nb;
    return 0;
}

/* The captured step holds kernel arguments (item lists, arenas, list
 * lengths, transform geometry): every change of those drops it. */
static void graphs_drop(ssa_engine *e)
{
    if (e->has_graph) { hipGraphExecDestroy(e->gexec); e->has_graph = 0; }
}

static int items_push(ssa_engine *e, int s, int x, int y, int z)
{
    if (e->hitems_n == e->hitems_cap) {
        long long nc = e->hitems_cap ? 2 * e->hitems_cap : 4096;
        ssa_item *p = (ssa_item *)realloc(e->hitems, sizeof(ssa_item) * nc);
        if (!p) return -1;
        e->hitems = p;
        e->hitems_cap = nc;
    }
    ssa_item it = {s, x, y, z};
    e->hitems[e->hitems_n++] = it;
    return 0;
}

static int items_upload(ssa_engine *e, long long from)
{
    long long n = e->hitems_n;
    if (n > e->ditems_cap) {
        graphs_drop(e);   /* the lists move */
        if (e->ditems) hipFree(e->ditems);
        e->ditems = NULL;
        long long cap = n + n / 2 + 1024;
        HIPCHK(hipMalloc((void **)&e->ditems, sizeof(ssa_item) * cap));
        e->ditems_cap = cap;
        from = 0;
    }
    if (from == 0) e->dyn_last_n = -1;   /* the whole list moves: forget the dynamic region */
    if (n > from) {
        const long long m = n - from;
        if (from > 0 && from == e->dyn_last_from && m == e->dyn_last_n &&
            !memcmp(e->dyn_last, e->hitems + from, sizeof(ssa_item) * m))
            return 0;
        DEVCHK(pin_reserve(e, sizeof(ssa_item) * m));
        memcpy(e->hpin, e->hitems + from, sizeof(ssa_item) * m);
        HIPCHK(hipMemcpyAsync(e->ditems + from, e->hpin, sizeof(ssa_item) * m, hipMemcpyHostToDevice, e->st));
        HIPCHK(hipStreamSynchronize(e->st));
        if (from > 0) {
            e->dyn_last_n = -1;
            if (m > e->dyn_last_cap) {
                ssa_item *p = (ssa_item *)realloc(e->dyn_last, sizeof(ssa_item) * m);
                if (!p) return 0;       /* the cache is an optimisation only */
                e->dyn_last = p;
                e->dyn_last_cap = m;
            }
            memcpy(e->dyn_last, e->hitems + from, sizeof(ssa_item) * m);
            e->dyn_last_n = m;
            e->dyn_last_from = from;
        }
    }
    return 0;
}

static const ssa_item *it_ptr(const ssa_engine *e, span sp) { return e->ditems + sp.off; }

/* matvec work items: (s, pair, o0, o1).  When a single input chunk feeds
 * several output chunks, an item covers a range of G output chunks and its
 * input transform serves them all (1 + G transforms instead of 2G).  G is as
 * large as keeps about MV_MIN_ITEMS items in the launch: the whole range for a
 * large batch, short ranges for one series with many output chunks (MSSA),
 * where a few long items would leave most CUs idle. */
enum { MV_MIN_ITEMS = 1024 };
/* default item count from which a block-matvec launch takes the wide geometry */
enum { MV_WIDE_MIN_ITEMS = 2048 };
/* component-loop splits of the batched step's reconstruction items */
enum { RECON_SPLIT = 3 };
/* context reconstructions: split a group's list up to this many ways while
 * the launch holds at most RECON_CTX_ITEMS pieces */
enum { RECON_SPLIT_CTX = 8, RECON_CTX_ITEMS = 64 };
/* A block matvec direction with more than one input chunk and at least
 * MV_STAGE_MIN_OUT output chunks takes the staged form: the input chunks are
 * transformed once into a scratch of the input-chunk spectra of every column
 * pair, then every output chunk sums them. */
#define MV_STAGE_MIN_OUT 2
static int mv_staged(const ssa_desc *d, int dir)
{
    const int nL = d->M * d->AL;
    const int n_in = dir == 0 ? d->SK : nL, n_out = dir == 0 ? nL : d->SK;
    return n_in > 1 && n_out >= MV_STAGE_MIN_OUT;
}

/* T elements of the staged scratch of a series of block width b (either direction) */
static long long mvs_elems(const ssa_desc *d, int b)
{
    if (!mv_staged(d, 0) && !mv_staged(d, 1)) return 0;
    const int n_in = d->SK > d->M * d->AL ? d->SK : d->M * d->AL;
    return 2LL * d->n * ((b + 1) / 2) * n_in;
}

/* the staged lists of a direction and bucket: (s, pair, input chunk) and (s, pair, output chunk) */
static int build_staged(ssa_engine *e, int dir, int lgn)
{
    span *sp = &e->stgS[dir][lgn], *sa = &e->stgA[dir][lgn];
    sp->off = e->hitems_n;
    for (int s = 0; s < e->count; s++) {
        const ssa_desc *d = &e->hd[s];
        if (d->lgn != lgn || d->b <= 0 || !(e->mv_stage && mv_staged(d, dir))) continue;
        const int n_in = dir == 0 ? d->SK : d->M * d->AL;
        for (int p = 0; p < (d->b + 1) / 2; p++)
            for (int c = 0; c < n_in; c++)
                if (items_push(e, s, p, c, 0)) return -1;
    }
    sp->n = (int)(e->hitems_n - sp->off);
    sa->off = e->hitems_n;
    for (int s = 0; s < e->count; s++) {
        const ssa_desc *d = &e->hd[s];
        if (d->lgn != lgn || d->b <= 0 || !(e->mv_stage && mv_staged(d, dir))) continue;
        const int n_out = dir == 0 ? d->M * d->AL : d->SK;
        for (int p = 0; p < (d->b + 1) / 2; p++)
            for (int o = 0; o < n_out; o++)
                if (items_push(e, s, p, o, o + 1)) return -1;
    }
    sa->n = (int)(e->hitems_n - sa->off);
    return 0;
}

static int build_matvec(ssa_engine *e, int dir, int lgn, span *sp)
{
    long long units = 0;   /* (pair, output chunk) units of the launch */
    for (int s = 0; s < e->count; s++) {
        const ssa_desc *d = &e->hd[s];
        if (d->lgn != lgn || d->b <= 0 || (e->mv_stage && mv_staged(d, dir))) continue;
        units += (long long)((d->b + 1) / 2) * (dir == 0 ? d->M * d->AL : d->SK);
    }
    sp->off = e->hitems_n;
    for (int s = 0; s < e->count; s++) {
        const ssa_desc *d = &e->hd[s];
        if (d->lgn != lgn || d->b <= 0 || (e->mv_stage && mv_staged(d, dir))) continue;
        int nLc = d->M * d->AL;
        int n_in = dir == 0 ? d->SK : nLc, n_out = dir == 0 ? nLc : d->SK;
        long long g = n_in == 1 ? units / MV_MIN_ITEMS : 1;
        int G = (int)(g < 1 ? 1 : g > n_out ? n_out : g);
        for (int p = 0; p < (d->b + 1) / 2; p++)
            for (int o = 0; o < n_out; o += G)
                if (items_push(e, s, p, o, o + G < n_out ? o + G : n_out)) return -1;
    }
    sp->n = (int)(e->hitems_n - sp->off);
    return 0;
}

static int rebuild_lists(ssa_engine *e)
{
    graphs_drop(e);   /* list lengths and offsets may change */
    e->hitems_n = 0;
    for (int lg = 0; lg < NLG; lg++) {
        e->spec[lg].off = e->hitems_n;
        for (int s = 0; s < e->count; s++) {
            const ssa_desc *d = &e->hd[s];
            if (d->lgn != lg) continue;
            for (int o = 0; o < d->M * d->AL; o++)
                for (int sk = 0; sk < d->SK; sk++) if (items_push(e, s, o, sk, 0)) return -1;
        }
        e->spec[lg].n = (int)(e->hitems_n - e->spec[lg].off);
    }
    for (int lg = 0; lg < NLG; lg++) {
        if (build_matvec(e, 0, lg, &e->mvf[lg])) return -1;
        if (build_matvec(e, 1, lg, &e->mva[lg])) return -1;
        if (build_staged(e, 0, lg) || build_staged(e, 1, lg)) return -1;
    }
    /* Gram / apply row tiles, one list per column-block count of the series */
    for (int nb = 0; nb < NBW; nb++) {
        e->tY[nb].off = e->hitems_n;
        for (int s = 0; s < e->count; s++) {
            if (e->hd[s].b <= 0 || SSA_NB(e->hd[s].b) != nb) continue;
            for (int t = 0; t < e->hd[s].tilesY; t++) if (items_push(e, s, t, 0, 0)) return -1;
        }
        e->tY[nb].n = (int)(e->hitems_n - e->tY[nb].off);
        e->tZ[nb].off = e->hitems_n;
        for (int s = 0; s < e->count; s++) {
            if (e->hd[s].b <= 0 || SSA_NB(e->hd[s].b) != nb) continue;
            for (int t = 0; t < e->hd[s].tilesZ; t++) if (items_push(e, s, t, 0, 0)) return -1;
        }
        e->tZ[nb].n = (int)(e->hitems_n - e->tZ[nb].off);
    }
    e->ser.off = e->hitems_n;
    for (int s = 0; s < e->count; s++) if (items_push(e, s, 0, 0, 0)) return -1;
    e->ser.n = (int)(e->hitems_n - e->ser.off);
    for (int nb = 0; nb < NBW; nb++) {
        e->ser_nb[nb].off = e->hitems_n;
        for (int s = 0; s < e->count; s++)
            if (e->hd[s].b > 0 && SSA_NB(e->hd[s].b) == nb && items_push(e, s, 0, 0, 0)) return -1;
        e->ser_nb[nb].n = (int)(e->hitems_n - e->ser_nb[nb].off);
    }
    /* batched step: group 0 covering components 0..k-1, its component loop
     * split RECON_SPLIT ways (as many as the staging slots allow) */
    for (int lg = 0; lg < NLG; lg++) {
        e->rec_step[lg].off = e->hitems_n;
        for (int s = 0; s < e->count; s++) {
            const ssa_desc *d = &e->hd[s];
            if (d->lgn != lg || d->k <= 0) continue;
            const long long slots = d->stage_cap / ((long long)d->AL * d->SK * d->n);
            const int S = (int)(slots < RECON_SPLIT ? (slots < 1 ? 1 : slots) : RECON_SPLIT);
            for (int j = 0; j < S; j++)
                for (int a = 0; a < d->AL; a++)
                    for (int sk = 0; sk < d->SK; sk++) if (items_push(e, s, j, a, sk)) return -1;
        }
        e->rec_step[lg].n = (int)(e->hitems_n - e->rec_step[lg].off);
    }
    e->olap_step.off = e->hitems_n;
    for (int s = 0; s < e->count; s++) {
        const ssa_desc *d = &e->hd[s];
        if (d->k <= 0) continue;
        for (int t = 0; t < ceil_div(d->N, OLAP_TILE_STEP); t++) if (items_push(e, s, 0, t, 0)) return -1;
    }
    e->olap_step.n = (int)(e->hitems_n - e->olap_step.off);
    e->dyn_base = (int)e->hitems_n;
    DEVCHK(items_upload(e, 0));
    e->lists_dirty = 0;
    return 0;
}

/* Upload the descriptors when they changed.  The copy is ordered on the
 * engine stream before every kernel that reads them, so the host does not
 * wait for it (one host round trip less per call: every decomposition
 * changes its series' test-matrix counter); its staging buffer is refilled
 * only once the event recorded after the copy has completed. */
static int sync_desc(ssa_engine *e)
{
    if (e->lists_dirty) DEVCHK(rebuild_lists(e));
    if (e->desc_dirty) {
        const size_t bytes = sizeof(ssa_desc) * (size_t)e->count;
        if (!e->dpin_ev) HIPCHK(hipEventCreateWithFlags(&e->dpin_ev, hipEventDisableTiming));
        if (e->dpin_busy) {
            HIPCHK(hipEventSynchronize(e->dpin_ev));
            e->dpin_busy = 0;
        }
        if (e->dpin_bytes < bytes) {
            if (e->dpin) hipHostFree(e->dpin);
            e->dpin = NULL;
            e->dpin_bytes = 0;
            HIPCHK(hipHostMalloc(&e->dpin, bytes, hipHostMallocDefault));
            e->dpin_bytes = bytes;
        }
        memcpy(e->dpin, e->hd, bytes);
        HIPCHK(hipMemcpyAsync(e->dd, e->dpin, bytes, hipMemcpyHostToDevice, e->st));
        HIPCHK(hipEventRecord(e->dpin_ev, e->st));
        e->dpin_busy = 1;
        e->desc_dirty = 0;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* lifecycle                                                                 */
/* ------------------------------------------------------------------------ */

/* descriptors, chunk plans and the signal/spectra arena for a set of series;
 * device buffers are reused when large enough */
static int engine_layout(ssa_engine *e, int count, const int *N, const int *L, const int *M, const int *bmax,
                         const int *kmax)
{
    if (count < 1) return -1;
    if (count > e->host_cap) {
        free(e->hd); free(e->bmax); free(e->kmax);
        e->hd = (ssa_desc *)calloc(count, sizeof(ssa_desc));
        e->bmax = (int *)calloc(count, sizeof(int));
        e->kmax = (int *)calloc(count, sizeof(int));
        if (!e->hd || !e->bmax || !e->kmax) { e->host_cap = 0; return -1; }
        e->host_cap = count;
    } else {
        memset(e->hd, 0, sizeof(ssa_desc) * count);
    }
    e->count = count;
    long long off = 0;
    for (int s = 0; s < count; s++) {
        ssa_desc *d = &e->hd[s];
        int m = M ? M[s] : 1;
        if (N[s] < 4 || L[s] < 2 || L[s] > N[s] - 1 || m < 1) return -1;
        d->N = N[s]; d->L = L[s]; d->K = N[s] - L[s] + 1; d->M = m; d->ML = m * L[s];
        e->bmax[s] = bmax ? bmax[s] : 0;
        e->kmax[s] = kmax ? kmax[s] : 0;
        if (e->bmax[s] > SSA_BMAX) return -1;
        ssa_plan pl;
        int bb = e->bmax[s] > 0 ? e->bmax[s] : 60, kk = e->kmax[s] > 0 ? e->kmax[s] : 50;
        if (ssa_plan_series(e->prec, d->N, d->L, m, bb, kk, e->q_hint, -1, &pl)) return -1;
        d->n = pl.n; d->lgn = pl.lgn; d->CL = pl.CL; d->CK = pl.CK; d->AL = pl.AL; d->SK = pl.SK;
        d->tilesY = ceil_div(d->ML, SSA_GRAM_ROWS);
        d->tilesZ = ceil_div(d->K, SSA_GRAM_ROWS);
        d->ldY = panel_ld(d->ML);
        d->ldZ = panel_ld(d->K);
        d->seed = 42;
        /* every series' signal one after the other, so that one copy uploads them all */
        d->off_x = off; off = align64(off + (long long)m * d->N);
    }
    for (int s = 0; s < count; s++) {
        ssa_desc *d = &e->hd[s];
        d->off_spec = off; off = align64(off + (long long)d->M * d->AL * d->SK * 2 * d->n);
    }
    e->a_elems = off;
    if (off > e->a_cap || !e->arena_a) {
        if (e->arena_a) { hipFree(e->arena_a); e->arena_a = NULL; e->a_cap = 0; }
        const long long cap = off > 64 ? off + off / 4 : 64;
        HIPCHK(hipMalloc(&e->arena_a, e->esz * (size_t)cap));
        e->a_cap = cap;
    }
    if (count > e->dd_cap) {
        if (e->dd) { hipFree(e->dd); e->dd = NULL; e->dd_cap = 0; }
        HIPCHK(hipMalloc((void **)&e->dd, sizeof(ssa_desc) * count));
        e->dd_cap = count;
    }
    e->prepared = 0;
    e->cache_valid = 0;
    e->max_groups = 0;
    e->max_h = 0;
    e->step_h = 0;
    e->lists_dirty = 1;
    e->desc_dirty = 1;
    e->omega_call = -1;
    e->step_groups_ready = 0;
    graphs_drop(e);
    return 0;
}

ssa_engine *ssa_engine_create(int prec, int count, const int *N, const int *L, const int *M, const int *bmax,
                              const int *kmax, int q_hint, int *err)
{
    int dummy;
    if (!err) err = &dummy;
    *err = -1;
    if (count < 1 || (prec != SSA_PREC_F64 && prec != SSA_PREC_F32)) return NULL;
    ssa_engine *e = (ssa_engine *)calloc(1, sizeof(ssa_engine));
    if (!e) return NULL;
    e->prec = prec;
    e->esz = prec == SSA_PREC_F64 ? 8 : 4;
    e->q_hint = q_hint;
    e->dyn_last_n = e->grp_last_n = -1;
    e->omega_call = -1;
    if (hipStreamCreate(&e->st) != hipSuccess) goto fail;
    if (ssa_dev_twiddles(prec, &e->tw, e->st) != 0) goto fail;
    if (engine_layout(e, count, N, L, M, bmax, kmax)) goto fail;
    {
        const char *v = getenv("SSA_AMD_CHOLQR2_ALL");
        e->cholqr2_all = (v && *v && *v != '0') ? 1 : 0;
        /* FP32 panels store Q in FP32, which bounds its orthogonality at
         * about eps32 whatever the factorisation; one FP64 CholeskyQR
         * (defect eps64 * cond^2) is then enough for the last
         * orthonormalisation and for the R factor of B.  FP64 panels keep
         * CholeskyQR2 there. */
        const char *f = getenv("SSA_AMD_FINAL_CHOLQR2");
        e->final2 = (prec == SSA_PREC_F64) || (f && *f && *f != '0');
        /* The wide matvec geometry (fewer LDS exchanges, half the waves) pays
         * when a launch fills the device; a launch of few items is bound by
         * each item's latency, which the narrow geometry's extra waves hide.
         * SSA_AMD_MV_WIDE_MIN moves the threshold (1: always wide). */
        /* A panel column whose residual, after the earlier columns are
         * projected out, is at the rounding floor of the panel's storage is
         * dropped (its Q column is zero).  FP64 panels: pivot / diagonal
         * 1e-13.  FP32 panels carry transform rounding of about 2.5e-7 of a
         * column's norm, so a rank-deficient panel (a noise-free signal of
         * rank below k + p) leaves pivots near 1e-13: kept, such a column is
         * normalised rounding noise, R^-1 scales it by ~1e6 and the FP32
         * apply's cancellation leans it into the signal span (measured: U
         * orthogonal to 1e-3, singular values off by 3e-4).  The FP32 floor
         * is 1e-10 (a residual of 1e-5, about 40x the rounding). */
        e->dead_tol = prec == SSA_PREC_F64 ? 1e-13 : 1e-10;
        const char *w = getenv("SSA_AMD_MV_WIDE_MIN");
        e->mv_wide_min = (w && *w) ? atoi(w) : MV_WIDE_MIN_ITEMS;
        if (e->mv_wide_min < 1) e->mv_wide_min = 1;
        /* the staged form sums the same products in the same order as the
         * fused one; SSA_AMD_MV_STAGED=0 keeps the fused form (comparison) */
        const char *ms = getenv("SSA_AMD_MV_STAGED");
        e->mv_stage = !(ms && *ms == '0');
    }
    *err = 0;
    return e;
fail:
    ssa_engine_destroy(e);
    return NULL;
}

int ssa_engine_rebind(ssa_engine *e, int count, const int *N, const int *L, const int *M, const int *bmax,
                      const int *kmax)
{
    if (!e || !N || !L) return -1;
    HIPCHK(hipStreamSynchronize(e->st));
    return engine_layout(e, count, N, L, M, bmax, kmax);
}

void ssa_engine_destroy(ssa_engine *e)
{
    if (!e) return;
    if (e->st) hipStreamSynchronize(e->st);
    if (e->has_graph) hipGraphExecDestroy(e->gexec);
    if (e->arena_a) hipFree(e->arena_a);
    if (e->arena_b) hipFree(e->arena_b);
    if (e->small) hipFree(e->small);
    if (e->cache) hipFree(e->cache);
    if (e->dd) hipFree(e->dd);
    if (e->ditems) hipFree(e->ditems);
    if (e->dgroups) hipFree(e->dgroups);
    free(e->dyn_last);
    free(e->grp_last);
    if (e->hpin) hipHostFree(e->hpin);
    if (e->hres) hipHostFree(e->hres);
    if (e->dpin) hipHostFree(e->dpin);
    if (e->dpin_ev) hipEventDestroy(e->dpin_ev);
    for (int i = 0; i < 2 * e->prof_cap; i++) hipEventDestroy(e->prof_ev[i]);
    free(e->prof_ev); free(e->prof_phase); free(e->prof_lgn);
    if (e->st) hipStreamDestroy(e->st);
    free(e->hd); free(e->bmax); free(e->kmax); free(e->hitems);
    free(e);
}

int ssa_engine_count(const ssa_engine *e) { return e->count; }
int ssa_engine_precision(const ssa_engine *e) { return e->prec; }
const ssa_desc *ssa_engine_desc(const ssa_engine *e, int s) { return &e->hd[s]; }
void *ssa_engine_stream(ssa_engine *e) { return (void *)e->st; }

long long ssa_engine_bytes(const ssa_engine *e)
{
    return (long long)e->esz * (e->a_elems + e->b_elems) + 8LL * e->small_elems;
}

int ssa_engine_sync(ssa_engine *e)
{
    HIPCHK(hipStreamSynchronize(e->st));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* signals                                                                   */
/* ------------------------------------------------------------------------ */

int ssa_engine_upload_signal(ssa_engine *e, int s, const double *x)
{
    if (s < 0 || s >= e->count || !x) return -1;
    const ssa_desc *d = &e->hd[s];
    size_t n = (size_t)d->M * d->N;
    DEVCHK(pin_reserve(e, n * e->esz));
    if (e->prec == SSA_PREC_F64) memcpy(e->hpin, x, n * 8);
    else for (size_t i = 0; i < n; i++) ((float *)e->hpin)[i] = (float)x[i];
    HIPCHK(hipMemcpyAsync((char *)e->arena_a + e->esz * d->off_x, e->hpin, n * e->esz, hipMemcpyHostToDevice, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    return 0;
}

int ssa_engine_upload_signals_f32(ssa_engine *e, const float *cat, const long long *offsets)
{
    /* the signals sit one after the other at the start of the arena
     * (engine_layout): staged in that layout, they go up in one copy */
    const ssa_desc *last = &e->hd[e->count - 1];
    const size_t total = (size_t)(last->off_x + (long long)last->M * last->N);
    DEVCHK(pin_reserve(e, total * e->esz));
    for (int s = 0; s < e->count; s++) {
        const ssa_desc *d = &e->hd[s];
        const size_t n = (size_t)d->M * d->N;
        char *dst = (char *)e->hpin + (size_t)d->off_x * e->esz;
        if (e->prec == SSA_PREC_F32) memcpy(dst, cat + offsets[s], n * 4);
        else for (size_t i = 0; i < n; i++) ((double *)dst)[i] = cat[offsets[s] + i];
    }
    HIPCHK(hipMemcpyAsync(e->arena_a, e->hpin, total * e->esz, hipMemcpyHostToDevice, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    return 0;
}

static int launch_spectra(ssa_engine *e)
{
    for (int lg = 6; lg < NLG; lg++) {
        if (!e->spec[lg].n) continue;
        prof_begin(e, PH_SPECTRA, lg);
        DEVCHK(ssa_dev_spectra(e->prec, lg, e->dd, it_ptr(e, e->spec[lg]), e->spec[lg].n, e->arena_a, e->tw, e->st));
        prof_end(e);
    }
    return 0;
}

int ssa_engine_spectra(ssa_engine *e)
{
    DEVCHK(sync_desc(e));
    DEVCHK(launch_spectra(e));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* panels                                                                    */
/* ------------------------------------------------------------------------ */

/* Device bytes one series of (N, L, b, k) takes in an engine of precision
 * prec prepared for max_groups groups and horizon max_h: the signal arena
 * (x and segment spectra), the panel arena (Y, Z, U, V, staging, outputs) and
 * the FP64 small arena, with the planner's chunking and the layout of
 * ssa_engine_create / ssa_engine_prepare (host arithmetic only). */
long long ssa_engine_series_bytes(int prec, int N, int L, int b, int k, int max_groups, int max_h, int q)
{
    if (N < 4 || L < 2 || L > N - 1 || b < 2 || k < 1) return -1;
    ssa_plan pl;
    if (ssa_plan_series(prec, N, L, 1, b, k, q, -1, &pl)) return -1;
    const long long esz = prec == SSA_PREC_F64 ? 8 : 4;
    const long long K = N - L + 1, ML = L;
    const int gmax = max_groups > k ? max_groups : k;
    long long a = align64(N) + align64((long long)pl.AL * pl.SK * 2 * pl.n);
    long long bulk = align64((long long)panel_ld((int)ML) * b) + align64((long long)panel_ld((int)K) * b) +
                     align64(ML * k) + align64(K * k);
    bulk += align64((long long)gmax * pl.AL * pl.SK * pl.n);
    bulk += align64((long long)gmax * N + max_h + 64);
    {
        ssa_desc d1;
        memset(&d1, 0, sizeof(d1));
        d1.M = 1; d1.n = pl.n; d1.AL = pl.AL; d1.SK = pl.SK;
        bulk += align64(mvs_elems(&d1, b));   /* staged block-matvec scratch */
    }
    bulk += align64((long long)N + max_h);   /* gathered step output */
    int tilesY = ceil_div((int)ML, SSA_GRAM_ROWS), tilesZ = ceil_div((int)K, SSA_GRAM_ROWS);
    int tiles = tilesY > tilesZ ? tilesY : tilesZ, tout = ceil_div(N, SSA_GRAM_ROWS);
    if (tout > tiles) tiles = tout;
    long long part = (long long)tiles * (b * b + b), need = 2LL * L + max_h + 256;
    if (part < need) part = need;
    const long long small = align64(part) + align64((long long)SSA_NSLOTS * SSA_BMAX * SSA_BMAX);
    return esz * (a + bulk) + 8 * small;
}

int ssa_engine_prepare(ssa_engine *e, int max_groups, int max_h)
{
    if (max_groups < 1) max_groups = 1;
    if (max_h < 0) max_h = 0;
    if (e->prepared && max_groups <= e->max_groups && max_h <= e->max_h) return 0;
    if (max_groups < e->max_groups) max_groups = e->max_groups;
    if (max_h < e->max_h) max_h = e->max_h;
    HIPCHK(hipStreamSynchronize(e->st));
    long long off = 0, soff = 0;
    for (int s = 0; s < e->count; s++) {
        ssa_desc *d = &e->hd[s];
        int bm = e->bmax[s], km = e->kmax[s];
        if (bm < 2 || km < 1) return -1;
        int gmax = max_groups > km ? max_groups : km;   /* W-correlation needs k pieces */
        d->off_Y = off; off = align64(off + (long long)d->ldY * bm);
        d->off_Z = off; off = align64(off + (long long)d->ldZ * bm);
        d->off_U = off; off = align64(off + (long long)d->ML * km);
        d->off_V = off; off = align64(off + (long long)d->K * km);
        d->stage_cap = (long long)gmax * d->AL * d->SK * d->n;
        d->off_stage = off; off = align64(off + d->stage_cap);
        d->out_cap = (long long)gmax * d->N + max_h + 64;
        d->off_out = off; off = align64(off + d->out_cap);
        d->off_mvs = off; off = align64(off + mvs_elems(d, bm));
        int tiles = d->tilesY > d->tilesZ ? d->tilesY : d->tilesZ;
        int tout = ceil_div(d->N, SSA_GRAM_ROWS);
        if (tout > tiles) tiles = tout;
        long long part = (long long)tiles * (bm * bm + bm);
        long long need = 2LL * d->L + max_h + 256;
        if (part < need) part = need;
        d->off_part = soff; soff = align64(soff + part);
        d->off_small = soff; soff = align64(soff + (long long)SSA_NSLOTS * SSA_BMAX * SSA_BMAX);
    }
    /* the step outputs gathered one after the other (one copy to the host) */
    for (int s = 0; s < e->count; s++) {
        ssa_desc *d = &e->hd[s];
        d->off_gather = off; off = align64(off + d->N + max_h);
    }
    if (off > e->b_cap || !e->arena_b) {
        if (e->arena_b) { hipFree(e->arena_b); e->arena_b = NULL; e->b_cap = 0; }
        const long long cap = off > 64 ? off + off / 4 : 64;
        HIPCHK(hipMalloc(&e->arena_b, e->esz * (size_t)cap));
        e->b_cap = cap;
    }
    if (soff > e->small_cap || !e->small) {
        if (e->small) { hipFree(e->small); e->small = NULL; e->small_cap = 0; }
        const long long cap = soff > 64 ? soff + soff / 4 : 64;
        HIPCHK(hipMalloc((void **)&e->small, 8 * (size_t)cap));
        e->small_cap = cap;
    }
    e->b_elems = off;
    e->small_elems = soff;
    /* rebase panel offsets on arena_a */
    long long delta = ((char *)e->arena_b - (char *)e->arena_a);
    if (delta % (long long)e->esz) return -1;
    delta /= (long long)e->esz;
    for (int s = 0; s < e->count; s++) {
        ssa_desc *d = &e->hd[s];
        d->off_Y += delta; d->off_Z += delta; d->off_U += delta; d->off_V += delta;
        d->off_stage += delta; d->off_out += delta; d->off_gather += delta; d->off_mvs += delta;
        if (d->b == 0) { d->b = e->bmax[s]; d->k = e->kmax[s]; }
        if (d->gsize <= 0 || d->gsize > d->k) d->gsize = d->k;
    }
    e->prepared = 1;
    e->cache_valid = 0;   /* arenas may have moved */
    e->max_groups = max_groups;
    e->max_h = max_h;
    e->lists_dirty = 1;
    e->desc_dirty = 1;
    e->step_groups_ready = 0;
    graphs_drop(e);
    return 0;
}

int ssa_engine_set_caps(ssa_engine *e, int s, int bmax, int kmax)
{
    if (s < 0 || s >= e->count || bmax < 2 || bmax > SSA_BMAX || kmax < 1 || kmax > bmax) return -1;
    if (e->bmax[s] != bmax || e->kmax[s] != kmax) {
        e->bmax[s] = bmax;
        e->kmax[s] = kmax;
        e->hd[s].b = bmax;
        e->hd[s].k = kmax;
        e->hd[s].gsize = kmax;
        e->prepared = 0;
        e->max_groups = 0;
        e->max_h = 0;
        e->lists_dirty = 1;
        e->desc_dirty = 1;
    }
    return 0;
}

int ssa_engine_read_spectrum(ssa_engine *e, int s, double *out, int ncomplex)
{
    if (s < 0 || s >= e->count || !out) return -1;
    const ssa_desc *d = &e->hd[s];
    if (ncomplex > d->n) ncomplex = d->n;
    DEVCHK(pin_reserve(e, 2 * (size_t)ncomplex * e->esz));
    HIPCHK(hipMemcpyAsync(e->hpin, (char *)e->arena_a + e->esz * d->off_spec, 2 * (size_t)ncomplex * e->esz,
                          hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    /* stored as X / n; report the unscaled transform */
    for (int i = 0; i < 2 * ncomplex; i++)
        out[i] = (e->prec == SSA_PREC_F64 ? ((double *)e->hpin)[i] : ((float *)e->hpin)[i]) * d->n;
    return 0;
}

int ssa_engine_set_bk(ssa_engine *e, int s, int b, int k)
{
    if (s < 0 || s >= e->count || b < 1 || k < 1 || b > e->bmax[s] || k > e->kmax[s] || k > b) return -1;
    if (e->hd[s].b != b || e->hd[s].k != k) {
        e->hd[s].b = b;
        e->hd[s].k = k;
        e->hd[s].gsize = k;
        e->lists_dirty = 1;
        e->desc_dirty = 1;
        e->step_groups_ready = 0;
        graphs_drop(e);
    }
    return 0;
}

int ssa_engine_set_step(ssa_engine *e, int s, int gsize)
{
    if (s < 0 || s >= e->count || gsize < 1 || gsize > e->hd[s].k) return -1;
    if (e->hd[s].gsize != gsize) {
        e->hd[s].gsize = gsize;
        e->desc_dirty = 1;
    }
    return 0;
}

int ssa_engine_set_step_h(ssa_engine *e, int h)
{
    if (h < 0 || (h > 0 && (!e->prepared || h > e->max_h))) return -1;
    if (h != e->step_h) {
        e->step_h = h;
        graphs_drop(e);
    }
    return 0;
}

int ssa_engine_query(ssa_engine *e)
{
    hipError_t r = hipStreamQuery(e->st);
    if (r == hipSuccess) return 1;
    if (r == hipErrorNotReady) return 0;
    return -1;
}

int ssa_engine_status(ssa_engine *e, int s)
{
    if (s < 0 || s >= e->count) return -1;
    int st = 0;
    HIPCHK(hipMemcpyAsync(&st, (char *)(e->dd + s) + offsetof(ssa_desc, status), sizeof(int), hipMemcpyDeviceToHost,
                          e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    return st;
}

void ssa_engine_set_wide_min(ssa_engine *e, int n)
{
    if (!e) return;
    e->mv_wide_min = n < 1 ? 1 : n;
    graphs_drop(e);
}

int ssa_engine_set_rng(ssa_engine *e, int s, unsigned long long seed, unsigned long long call)
{
    if (s < 0 || s >= e->count) return -1;
    e->hd[s].call = call;
    if (e->count == 1 && !e->desc_dirty && e->hd[s].seed == seed && call <= (unsigned long long)LLONG_MAX) {
        e->omega_call = (long long)call;
        return 0;
    }
    e->hd[s].seed = seed;
    e->desc_dirty = 1;
    e->omega_call = -1;
    return 0;
}

int ssa_engine_inject_omega(ssa_engine *e, int s, const double *Omega)
{
    if (!e->prepared || s < 0 || s >= e->count || !Omega) return -1;
    const ssa_desc *d = &e->hd[s];
    size_t n = (size_t)d->K * d->b;
    DEVCHK(pin_reserve(e, n * e->esz));
    if (e->prec == SSA_PREC_F64) memcpy(e->hpin, Omega, n * 8);
    else for (size_t i = 0; i < n; i++) ((float *)e->hpin)[i] = (float)Omega[i];
    /* K x b, column-major, into the Z panel (leading dimension ldZ) */
    HIPCHK(hipMemcpy2DAsync((char *)e->arena_a + e->esz * d->off_Z, (size_t)d->ldZ * e->esz, e->hpin,
                            (size_t)d->K * e->esz, (size_t)d->K * e->esz, (size_t)d->b, hipMemcpyHostToDevice, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* decomposition                                                             */
/* ------------------------------------------------------------------------ */

static int bmax_all(const ssa_engine *e)
{
    int m = 0;
    for (int s = 0; s < e->count; s++) if (e->hd[s].b > m) m = e->hd[s].b;
    return m;
}

static int matvec_all(ssa_engine *e, int dir, int in_panel, int out_panel)
{
    const int ph = dir == 0 ? PH_MV_FWD : PH_MV_ADJ;
    for (int lg = 6; lg < NLG; lg++) {
        span sp = dir == 0 ? e->mvf[lg] : e->mva[lg];
        if (!sp.n) continue;
        prof_begin(e, ph, lg);
        DEVCHK(ssa_dev_matvec(e->prec, lg, dir, in_panel, out_panel, sp.n >= e->mv_wide_min, e->dd,
                              it_ptr(e, sp), sp.n, e->arena_a,
                              e->tw, e->st));
        prof_end(e);
    }
    for (int lg = 6; lg < NLG; lg++) {
        const span ss = e->stgS[dir][lg], sa = e->stgA[dir][lg];
        if (!ss.n) continue;
        /* one transform geometry for both launches of a bucket */
        const int wide = (ss.n > sa.n ? ss.n : sa.n) >= e->mv_wide_min;
        prof_begin(e, ph, lg);
        DEVCHK(ssa_dev_mv_spec(e->prec, lg, dir, in_panel, wide, e->dd, it_ptr(e, ss), ss.n, e->arena_a, e->tw,
                               e->st));
        prof_end(e);
        prof_begin(e, ph, lg);
        DEVCHK(ssa_dev_mv_acc(e->prec, lg, dir, out_panel, wide, e->dd, it_ptr(e, sa), sa.n, e->arena_a, e->tw,
                              e->st));
        prof_end(e);
    }
    return 0;
}

/* the Philox test matrix of every series into its Z panel */
static int omega_all(ssa_engine *e)
{
    for (int nb = 1; nb < NBW; nb++) {
        if (!e->tZ[nb].n) continue;
        prof_begin(e, PH_OMEGA, nb);
        DEVCHK(ssa_dev_omega(e->prec, e->dd, it_ptr(e, e->tZ[nb]), e->tZ[nb].n, e->arena_a, e->omega_call, e->st));
        prof_end(e);
    }
    return 0;
}

/* partial Grams of a panel, one launch per column-block count */
static int gram_all(ssa_engine *e, int panel, int use_m)
{
    const span *t = panel == SSA_PANEL_Y ? e->tY : e->tZ;
    for (int nb = 1; nb < NBW; nb++) {
        if (!t[nb].n) continue;
        prof_begin(e, PH_GRAM, panel);
        DEVCHK(ssa_dev_gram(e->prec, panel, use_m, nb, e->dd, it_ptr(e, t[nb]), t[nb].n, e->arena_a, e->small,
                            e->st));
        prof_end(e);
    }
    return 0;
}

/* dst = src * (small matrix in slot) over the row tiles of src */
static int apply_all(ssa_engine *e, int phase, int src, int dst, int slot, int ncols_k)
{
    const span *t = src == SSA_PANEL_Y ? e->tY : e->tZ;
    for (int nb = 1; nb < NBW; nb++) {
        if (!t[nb].n) continue;
        prof_begin(e, phase, src);
        DEVCHK(ssa_dev_apply(e->prec, src, dst, slot, ncols_k, nb, e->dd, it_ptr(e, t[nb]), t[nb].n, e->arena_a,
                             e->small, e->st));
        prof_end(e);
    }
    return 0;
}

/* Cholesky of every series' summed Gram in one launch (sized for the widest
 * series): each series is one latency-bound wave, so separate launches per
 * width would each pay the slowest series' latency */
static int chol_all(ssa_engine *e, int panel, int pass, int bm)
{
    prof_begin(e, PH_CHOL, panel);
    DEVCHK(ssa_dev_chol(panel, pass, bm, e->dead_tol, e->dd, it_ptr(e, e->ser), e->ser.n, e->small, e->st));
    prof_end(e);
    return 0;
}

/* the products of a CholeskyQR2 that a caller needs */
enum { NEED_R = 1, NEED_RINV = 2 };

static int trmm_all(ssa_engine *e, int sa, int sb, int sc)
{
    for (int nb = 1; nb < NBW; nb++) {
        if (!e->ser_nb[nb].n) continue;
        prof_begin(e, PH_CHOL, nb);
        DEVCHK(ssa_dev_trmm(nb, sa, sb, sc, e->dd, it_ptr(e, e->ser_nb[nb]), e->ser_nb[nb].n, e->small, e->st));
        prof_end(e);
    }
    return 0;
}

static int cholqr_factor(ssa_engine *e, int panel, int bm, int need)
{
    DEVCHK(gram_all(e, panel, 0));
    DEVCHK(chol_all(e, panel, 1, bm));
    DEVCHK(gram_all(e, panel, 1));
    DEVCHK(chol_all(e, panel, 2, bm));
    if (need & NEED_R) DEVCHK(trmm_all(e, SSA_SLOT_R2, SSA_SLOT_R1, SSA_SLOT_R));
    if (need & NEED_RINV) DEVCHK(trmm_all(e, SSA_SLOT_R1INV, SSA_SLOT_R2INV, SSA_SLOT_RINV));
    return 0;
}

/* Orthonormalise a panel in place.  The last orthonormalisation of Y (the Q
 * that forms B = H^T Q and U = Q W) is CholeskyQR2; the ones inside the
 * power iterations only have to keep the span well conditioned, and a
 * single CholeskyQR in FP64 leaves an orthogonality defect of about
 * eps64 * cond^2 there, which the next orthonormalisation absorbs. */
static int orth(ssa_engine *e, int panel, int bm, int two_pass)
{
    int slot;
    if (two_pass) {
        DEVCHK(cholqr_factor(e, panel, bm, NEED_RINV));
        slot = SSA_SLOT_RINV;
    } else {
        DEVCHK(gram_all(e, panel, 0));
        DEVCHK(chol_all(e, panel, 1, bm));
        slot = SSA_SLOT_R1INV;
    }
    return apply_all(e, PH_APPLY, panel, panel, slot, 0);
}

static int launch_decompose(ssa_engine *e, int q, int injected)
{
    int bm = bmax_all(e);
    const int all2 = e->cholqr2_all, fin2 = e->final2 || all2;
    if (!injected) DEVCHK(omega_all(e));
    DEVCHK(matvec_all(e, 0, SSA_PANEL_Z, SSA_PANEL_Y));
    DEVCHK(orth(e, SSA_PANEL_Y, bm, all2 || (fin2 && q == 0)));
    for (int it = 0; it < q; it++) {
        DEVCHK(matvec_all(e, 1, SSA_PANEL_Y, SSA_PANEL_Z));
        DEVCHK(orth(e, SSA_PANEL_Z, bm, all2));
        DEVCHK(matvec_all(e, 0, SSA_PANEL_Z, SSA_PANEL_Y));
        DEVCHK(orth(e, SSA_PANEL_Y, bm, all2 || (fin2 && it == q - 1)));
    }
    DEVCHK(matvec_all(e, 1, SSA_PANEL_Y, SSA_PANEL_Z));
    /* R factor of B; the R^-1 of the final Q (for colsum(Q)) stays in its slot */
    int r_slot, rinv_slot;
    if (fin2) {
        DEVCHK(cholqr_factor(e, SSA_PANEL_Z, bm, NEED_R));
        r_slot = SSA_SLOT_R;
        rinv_slot = SSA_SLOT_RINV;
    } else {
        DEVCHK(gram_all(e, SSA_PANEL_Z, 0));
        DEVCHK(chol_all(e, SSA_PANEL_Z, 2, bm));   /* into the R2 slots: R1^-1 of Y is still needed */
        r_slot = SSA_SLOT_R2;
        rinv_slot = SSA_SLOT_R1INV;
    }
    /* Jacobi stops after a sweep whose largest column cosine was below this
     * (the residual is then of its square): 1e-16 plain, 1e-12 mixed */
    const double early = e->prec == SSA_PREC_F64 ? 1e-8 : 1e-6;
    for (int nb = 1; nb < NBW; nb++) {
        if (!e->ser_nb[nb].n) continue;
        prof_begin(e, PH_SVD, nb);
        DEVCHK(ssa_dev_svd(16 * nb < SSA_BMAX ? 16 * nb : SSA_BMAX, r_slot, rinv_slot, early, e->dd,
                           it_ptr(e, e->ser_nb[nb]), e->ser_nb[nb].n, e->small, e->st));
        prof_end(e);
    }
    DEVCHK(apply_all(e, PH_UV, SSA_PANEL_Y, SSA_PANEL_U, SSA_SLOT_WK, 1));
    DEVCHK(apply_all(e, PH_UV, SSA_PANEL_Z, SSA_PANEL_V, SSA_SLOT_WKS, 1));
    return 0;
}

int ssa_engine_decompose(ssa_engine *e, int q, int injected)
{
    if (!e->prepared || q < 0) return -1;
    e->cache_valid = 0;   /* the cached spectra belong to the previous factors */
    DEVCHK(sync_desc(e));
    DEVCHK(launch_decompose(e, q, injected));
    return 0;
}

double *ssa_engine_host_alloc(size_t n)
{
    void *p = NULL;
    if (hipHostMalloc(&p, n ? n * sizeof(double) : sizeof(double), hipHostMallocDefault) != hipSuccess) return NULL;
    memset(p, 0, n * sizeof(double));
    return (double *)p;
}

void ssa_engine_host_free(double *p)
{
    if (p) hipHostFree(p);
}

int ssa_engine_results(ssa_engine *e, int s, double *sigma, double *lambda, double *totvar, double *U, double *V,
                       double *sweeps)
{
    if (!e->prepared || s < 0 || s >= e->count) return -1;
    const ssa_desc *d = &e->hd[s];
    size_t misc_n = 4 * SSA_BMAX + 3;
    size_t u_n = U ? (size_t)d->ML * d->k : 0, v_n = V ? (size_t)d->K * d->k : 0;
    const int f64 = e->prec == SSA_PREC_F64;
    const double *dmisc = e->small + d->off_small + (long long)SSA_SLOT_MISC * SSA_BMAX * SSA_BMAX;
    const char *dU = (const char *)e->arena_a + e->esz * d->off_U, *dV = (const char *)e->arena_a + e->esz * d->off_V;
    const size_t all = misc_n * 8 + (u_n + v_n) * e->esz;
    /* Small results are written by one kernel into coherent page-locked
     * memory: no copy command.  Copies of U and V cost about 10 us each on
     * an idle copy path, and about 50 us each once the process has freed
     * large device buffers (measured at N = 500, k = 15); the copy of the
     * small FP64 block alone did not slow down.  Up to 1 MiB the kernel's
     * writes over the host link were as fast as the copies or faster (N =
     * 1000, k = 50: 1.05 -> 1.04 ms plain); 8 MiB of FP64 U and V (N =
     * 20000) took 0.28 ms longer than copying them. */
    const int direct = all <= RESULTS_DIRECT_MAX;
    /* large U and V are page-locked (ssa_engine_host_alloc): FP64 results are copied into them directly */
    const int du = !direct && f64 && u_n, dv = !direct && f64 && v_n;
    const size_t su = du ? 0 : u_n, sv = dv ? 0 : v_n;   /* elements staged */
    double *misc;
    if (direct) {
        DEVCHK(hres_reserve(e, all));
        DEVCHK(ssa_dev_pack_results(e->prec, dmisc, (int)misc_n, dU, (long long)u_n, dV, (long long)v_n,
                                    e->hres_dev, e->st));
        misc = (double *)e->hres;
    } else {
        DEVCHK(pin_reserve(e, misc_n * 8 + (su + sv) * e->esz));
        misc = (double *)e->hpin;
        char *su_p = (char *)e->hpin + misc_n * 8;
        HIPCHK(hipMemcpyAsync(misc, dmisc, misc_n * 8, hipMemcpyDeviceToHost, e->st));
        if (u_n) HIPCHK(hipMemcpyAsync(du ? (void *)U : (void *)su_p, dU, u_n * e->esz, hipMemcpyDeviceToHost, e->st));
        if (v_n) HIPCHK(hipMemcpyAsync(dv ? (void *)V : (void *)(su_p + su * e->esz), dV, v_n * e->esz,
                                       hipMemcpyDeviceToHost, e->st));
    }
    char *pu = (char *)misc + misc_n * 8;
    char *pv = pu + su * e->esz;
    HIPCHK(hipStreamSynchronize(e->st));
    if (sigma) memcpy(sigma, misc + SSA_MISC_SIGMA, sizeof(double) * d->k);
    if (lambda) memcpy(lambda, misc + SSA_MISC_LAMBDA, sizeof(double) * d->k);
    if (totvar) *totvar = misc[SSA_MISC_TOTVAR];
    if (sweeps) *sweeps = misc[SSA_MISC_SWEEPS];
    if (su) {
        if (f64) memcpy(U, pu, su * 8);
        else for (size_t i = 0; i < su; i++) U[i] = ((float *)pu)[i];
    }
    if (sv) {
        if (f64) memcpy(V, pv, sv * 8);
        else for (size_t i = 0; i < sv; i++) V[i] = ((float *)pv)[i];
    }
    return d->k;
}

/* ------------------------------------------------------------------------ */
/* reconstruction, W-correlation, forecast                                   */
/* ------------------------------------------------------------------------ */

static int groups_upload(ssa_engine *e, const int *gdesc4, int ng, const int *glist, int nl)
{
    long long need = 4LL * ng + nl + 16;
    if (need > e->dgroups_cap) {
        if (e->dgroups) hipFree(e->dgroups);
        e->dgroups = NULL;
        long long cap = need + 1024;
        HIPCHK(hipMalloc((void **)&e->dgroups, sizeof(int) * cap));
        e->dgroups_cap = cap;
        e->grp_last_n = -1;
    }
    const long long m = 4LL * ng + nl;
    if (m == e->grp_last_n && !memcmp(e->grp_last, gdesc4, sizeof(int) * 4 * ng) &&
        !memcmp(e->grp_last + 4 * ng, glist, sizeof(int) * nl))
        return 0;
    DEVCHK(pin_reserve(e, sizeof(int) * need));
    int *h = (int *)e->hpin;
    memcpy(h, gdesc4, sizeof(int) * 4 * ng);
    memcpy(h + 4 * ng, glist, sizeof(int) * nl);
    HIPCHK(hipMemcpyAsync(e->dgroups, h, sizeof(int) * m, hipMemcpyHostToDevice, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    e->grp_last_n = -1;
    if (m > e->grp_last_cap) {
        int *p = (int *)realloc(e->grp_last, sizeof(int) * m);
        if (!p) return 0;               /* the cache is an optimisation only */
        e->grp_last = p;
        e->grp_last_cap = m;
    }
    memcpy(e->grp_last, h, sizeof(int) * m);
    e->grp_last_n = m;
    return 0;
}

/* launch the reconstruction of ng groups of series s into its output slots */
static int launch_recon(ssa_engine *e, int s, int ng, int scale_sigma, int weight, int use_cache)
{
    const ssa_desc *d = &e->hd[s];
    const int cached = use_cache && e->cache_valid;
    const int n_pieces = ng * d->AL * d->SK;
    /* A launch of few pieces splits each group's component list S ways (the
     * overlap-add sums the S pieces), so that more workgroups share a short
     * reconstruction; as many as the staging slots hold for ng groups. */
    int S = 1;
    if (!cached) {
        const long long slots = d->stage_cap / ((long long)d->AL * d->SK * d->n);
        while (2 * S <= RECON_SPLIT_CTX && (long long)ng * 2 * S <= slots && n_pieces * 2 * S <= RECON_CTX_ITEMS)
            S *= 2;
    }
    e->hitems_n = e->dyn_base;
    span rec, ol, edge;
    rec.off = e->hitems_n;
    if (cached) {
        /* two pieces per item (one inverse transform) */
        for (int p = 0; p < n_pieces; p += 2) DEVCHK(items_push(e, s, p, 0, 0));
    } else {
        for (int g = 0; g < ng * S; g++)
            for (int a = 0; a < d->AL; a++)
                for (int sk = 0; sk < d->SK; sk++) DEVCHK(items_push(e, s, g, a, sk));
    }
    rec.n = (int)(e->hitems_n - rec.off);
    ol.off = e->hitems_n;
    for (int g = 0; g < ng; g++)
        for (int t = 0; t < ceil_div(d->N, SSA_OLAP_TILE); t++) DEVCHK(items_push(e, s, g, t, 0));
    ol.n = (int)(e->hitems_n - ol.off);
    edge.off = e->hitems_n;
    for (int g = 0; g < ng; g++) DEVCHK(items_push(e, s, g, 0, 0));
    edge.n = ng;
    DEVCHK(items_upload(e, e->dyn_base));
    if (cached)
        DEVCHK(ssa_dev_recon_cached(e->prec, d->lgn, scale_sigma, e->dd, it_ptr(e, rec), rec.n, n_pieces, e->dgroups,
                                    e->dgroups + 4 * ng, e->arena_a, e->small, e->tw, e->st));
    else
        DEVCHK(ssa_dev_recon(e->prec, d->lgn, scale_sigma, 0, S, e->dd, it_ptr(e, rec), rec.n, e->dgroups,
                             e->dgroups + 4 * ng, e->arena_a, e->small, e->tw, e->st));
    DEVCHK(ssa_dev_olap(e->prec, weight, S, SSA_OLAP_TILE, e->dd, it_ptr(e, ol), ol.n, e->dgroups, e->arena_a, e->st));
    DEVCHK(ssa_dev_edge(e->prec, scale_sigma, weight, 0, e->dd, it_ptr(e, edge), edge.n, e->dgroups,
                        e->dgroups + 4 * ng, e->arena_a, e->small, e->st));
    return 0;
}

static int fetch_out(ssa_engine *e, int s, long long count, double *out)
{
    const ssa_desc *d = &e->hd[s];
    const char *src = (const char *)e->arena_a + e->esz * d->off_out;
    const void *h;
    if ((size_t)count * e->esz <= RESULTS_DIRECT_MAX) {
        /* written into host memory by the device, as small decomposition results are */
        DEVCHK(hres_reserve(e, (size_t)count * e->esz));
        DEVCHK(ssa_dev_pack_results(e->prec, NULL, 0, src, count, NULL, 0, e->hres_dev, e->st));
        h = e->hres;
    } else {
        DEVCHK(pin_reserve(e, count * e->esz));
        HIPCHK(hipMemcpyAsync(e->hpin, src, count * e->esz, hipMemcpyDeviceToHost, e->st));
        h = e->hpin;
    }
    HIPCHK(hipStreamSynchronize(e->st));
    if (e->prec == SSA_PREC_F64) memcpy(out, h, count * 8);
    else for (long long i = 0; i < count; i++) out[i] = ((const float *)h)[i];
    return 0;
}

int ssa_engine_reconstruct(ssa_engine *e, int s, int ch, const int *glist, const int *gcount, int n_groups,
                           int scale_sigma, int weight, double *out)
{
    return ssa_engine_reconstruct_ex(e, s, ch, glist, gcount, n_groups, scale_sigma, weight, 1, out);
}

int ssa_engine_reconstruct_ex(ssa_engine *e, int s, int ch, const int *glist, const int *gcount, int n_groups,
                              int scale_sigma, int weight, int use_cache, double *out)
{
    if (!e->prepared || s < 0 || s >= e->count || n_groups < 1) return -1;
    const ssa_desc *d = &e->hd[s];
    if (ch < 0 || ch >= d->M) return -1;
    int gmax = e->max_groups > e->kmax[s] ? e->max_groups : e->kmax[s];
    /* the staging area holds gmax groups: without an output buffer the groups
     * must fit (the caller reads them on the device), otherwise go in chunks */
    if (!out && n_groups > gmax) return -1;
    DEVCHK(sync_desc(e));
    int *gd = (int *)malloc(sizeof(int) * 4 * (size_t)gmax);
    if (!gd) return -1;
    int rc = 0, first = 0;
    for (int g0 = 0; rc == 0 && g0 < n_groups; g0 += gmax) {
        int ng = n_groups - g0 < gmax ? n_groups - g0 : gmax;
        int start = 0;
        for (int g = 0; g < ng; g++) {
            gd[4 * g] = start; gd[4 * g + 1] = gcount[g0 + g]; gd[4 * g + 2] = ch; gd[4 * g + 3] = 0;
            start += gcount[g0 + g];
        }
        rc = groups_upload(e, gd, ng, glist + first, start);
        if (rc == 0) rc = launch_recon(e, s, ng, scale_sigma, weight, use_cache);
        if (rc == 0 && out) rc = fetch_out(e, s, (long long)ng * d->N, out + (size_t)g0 * d->N);
        first += start;
    }
    free(gd);
    return rc;
}

int ssa_engine_wcorr(ssa_engine *e, int s, int use_cache, double *W)
{
    if (!e->prepared || s < 0 || s >= e->count || !W) return -1;
    const ssa_desc *d = &e->hd[s];
    int k = d->k;
    int *gl = (int *)malloc(sizeof(int) * k), *gc = (int *)malloc(sizeof(int) * k);
    if (!gl || !gc) { free(gl); free(gc); return -1; }
    for (int i = 0; i < k; i++) { gl[i] = i; gc[i] = 1; }
    int rc = ssa_engine_reconstruct_ex(e, s, 0, gl, gc, k, 0, 1, use_cache, NULL);
    free(gl); free(gc);
    DEVCHK(rc);
    /* Gram tiles over the N x k weighted component panel */
    e->hitems_n = e->dyn_base;
    span t, one;
    t.off = e->hitems_n;
    for (int i = 0; i < ceil_div(d->N, SSA_GRAM_ROWS); i++) DEVCHK(items_push(e, s, i, 0, 0));
    t.n = (int)(e->hitems_n - t.off);
    one.off = e->hitems_n;
    DEVCHK(items_push(e, s, 0, 0, 0));
    one.n = 1;
    DEVCHK(items_upload(e, e->dyn_base));
    DEVCHK(ssa_dev_wcorr(e->prec, SSA_NB(k), e->dd, it_ptr(e, t), t.n, it_ptr(e, one), 1, e->arena_a, e->small, e->st));
    DEVCHK(pin_reserve(e, sizeof(double) * k * k));
    HIPCHK(hipMemcpyAsync(e->hpin, e->small + d->off_small + (long long)SSA_SLOT_W * SSA_BMAX * SSA_BMAX,
                          sizeof(double) * k * k, hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    memcpy(W, e->hpin, sizeof(double) * k * k);
    return 0;
}

int ssa_engine_forecast(ssa_engine *e, int s, const int *group, int n_group, int h, double *out, int full,
                        int vmode, const double *base, int base_len)
{
    if (!e->prepared || s < 0 || s >= e->count || n_group < 1 || h < 1 || h > e->max_h) return -1;
    const ssa_desc *d = &e->hd[s];
    if (n_group > SSA_BMAX) return -1;
    for (int g = 0; g < n_group; g++) if (group[g] < 0 || group[g] >= d->k) return -1;  /* MKL/ssa_opt.h:2782-2787 */
    const int w = d->L - 1;
    if (base && (base_len < w || full)) return -1;
    int gc = n_group;
    if (!base) {
        DEVCHK(ssa_engine_reconstruct(e, s, 0, group, &gc, 1, 1, 0, NULL));
    } else {
        /* the caller's base signal stands in for the reconstruction: its last
         * L-1 values go where the recurrence reads them (off_out + N - (L-1)) */
        int gd[4] = {0, n_group, 0, 0};
        DEVCHK(sync_desc(e));
        DEVCHK(groups_upload(e, gd, 1, group, n_group));
        DEVCHK(pin_reserve(e, (size_t)(w > 0 ? w : 1) * e->esz));
        for (int j = 0; j < w; j++) {
            const double v = base[base_len - w + j];
            if (e->prec == SSA_PREC_F64) ((double *)e->hpin)[j] = v;
            else ((float *)e->hpin)[j] = (float)v;
        }
        if (w > 0) {
            HIPCHK(hipMemcpyAsync((char *)e->arena_a + e->esz * (d->off_out + d->N - w), e->hpin, (size_t)w * e->esz,
                                  hipMemcpyHostToDevice, e->st));
            HIPCHK(hipStreamSynchronize(e->st));   /* the staging buffer is reused below */
        }
    }
    e->hitems_n = e->dyn_base;
    span one;
    one.off = e->hitems_n;
    DEVCHK(items_push(e, s, 0, 0, 0));
    one.n = 1;
    DEVCHK(items_upload(e, e->dyn_base));
    DEVCHK(ssa_dev_forecast(e->prec, h, 0, vmode, e->dd, it_ptr(e, one), 1, e->dgroups, e->dgroups + 4, e->arena_a,
                            e->small, e->st));
    /* status of the series: 2 = verticality >= 1 */
    int status = 0;
    HIPCHK(hipMemcpyAsync(&status, (char *)(e->dd + s) + offsetof(ssa_desc, status), sizeof(int),
                          hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    if (status) {
        int zero = 0;
        HIPCHK(hipMemcpyAsync((char *)(e->dd + s) + offsetof(ssa_desc, status), &zero, sizeof(int),
                              hipMemcpyHostToDevice, e->st));
        HIPCHK(hipStreamSynchronize(e->st));
        return -1;
    }
    double *tmp = (double *)malloc(sizeof(double) * ((size_t)d->N + h));
    if (!tmp) return -1;
    int rc = fetch_out(e, s, (long long)d->N + h, tmp);
    if (rc == 0) {
        if (full) memcpy(out, tmp, sizeof(double) * ((size_t)d->N + h));
        else memcpy(out, tmp + d->N, sizeof(double) * h);
    }
    free(tmp);
    return rc;
}

int ssa_engine_matvec(ssa_engine *e, int s, int dir, const double *in, double *out)
{
    if (!e->prepared || s < 0 || s >= e->count) return -1;
    ssa_desc *d = &e->hd[s];
    int b = d->b, k = d->k;
    /* column 0 of Z (dir 0) or Y (dir 1) holds the input; column 0 of the other panel the output */
    int lin = dir == 0 ? d->K : d->ML, lout = dir == 0 ? d->ML : d->K;
    long long oin = dir == 0 ? d->off_Z : d->off_Y, oout = dir == 0 ? d->off_Y : d->off_Z;
    DEVCHK(pin_reserve(e, (size_t)(lin > lout ? lin : lout) * e->esz));
    if (e->prec == SSA_PREC_F64) memcpy(e->hpin, in, (size_t)lin * 8);
    else for (int i = 0; i < lin; i++) ((float *)e->hpin)[i] = (float)in[i];
    HIPCHK(hipMemcpyAsync((char *)e->arena_a + e->esz * oin, e->hpin, (size_t)lin * e->esz, hipMemcpyHostToDevice, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    /* temporarily a block of width 1 */
    d->b = 1; d->k = 1;
    e->desc_dirty = 1;
    e->lists_dirty = 1;
    int rc = sync_desc(e);
    if (rc == 0) {
        e->hitems_n = e->dyn_base;
        span sp;
        sp.off = e->hitems_n;
        int n_out = dir == 0 ? d->M * d->AL : d->SK;
        for (int o = 0; o < n_out; o++) items_push(e, s, 0, o, o + 1);
        sp.n = (int)(e->hitems_n - sp.off);
        rc = items_upload(e, e->dyn_base);
        if (rc == 0)
            rc = ssa_dev_matvec(e->prec, d->lgn, dir, dir == 0 ? SSA_PANEL_Z : SSA_PANEL_Y,
                                dir == 0 ? SSA_PANEL_Y : SSA_PANEL_Z, sp.n >= e->mv_wide_min, e->dd, it_ptr(e, sp),
                                sp.n, e->arena_a, e->tw, e->st);
    }
    if (rc == 0 && hipMemcpyAsync(e->hpin, (char *)e->arena_a + e->esz * oout, (size_t)lout * e->esz,
                                  hipMemcpyDeviceToHost, e->st) != hipSuccess) rc = -1;
    if (rc == 0 && hipStreamSynchronize(e->st) != hipSuccess) rc = -1;
    if (rc == 0) {
        if (e->prec == SSA_PREC_F64) memcpy(out, e->hpin, (size_t)lout * 8);
        else for (int i = 0; i < lout; i++) out[i] = ((float *)e->hpin)[i];
    }
    d->b = b; d->k = k;
    e->desc_dirty = 1;
    e->lists_dirty = 1;
    return rc;
}

/* ------------------------------------------------------------------------ */
/* component spectra cache                                                   */
/* ------------------------------------------------------------------------ */

/* T elements of series s's cache: k components x (M AL + SK) chunks x (n/2 + 1) complex */
static long long cache_elems(const ssa_desc *d)
{
    return 2LL * d->k * (d->M * d->AL + d->SK) * (d->n / 2 + 1);
}

int ssa_engine_cache_build(ssa_engine *e)
{
    if (!e->prepared) return -1;
    long long total = 0;
    for (int s = 0; s < e->count; s++) total += align64(cache_elems(&e->hd[s]));
    if (total > e->cache_cap || !e->cache) {
        if (e->cache) { hipFree(e->cache); e->cache = NULL; e->cache_cap = 0; }
        HIPCHK(hipMalloc(&e->cache, e->esz * (size_t)(total > 64 ? total : 64)));
        e->cache_cap = total > 64 ? total : 64;
    }
    long long delta = (char *)e->cache - (char *)e->arena_a;
    if (delta % (long long)e->esz) return -1;
    delta /= (long long)e->esz;
    long long off = 0;
    for (int s = 0; s < e->count; s++) {
        e->hd[s].off_cache = delta + off;
        off += align64(cache_elems(&e->hd[s]));
    }
    e->desc_dirty = 1;
    DEVCHK(sync_desc(e));
    /* items (s, component pair, chunk), one launch per transform length */
    e->hitems_n = e->dyn_base;
    span per[NLG];
    for (int lg = 0; lg < NLG; lg++) {
        per[lg].off = e->hitems_n;
        for (int s = 0; s < e->count; s++) {
            const ssa_desc *d = &e->hd[s];
            if (d->lgn != lg) continue;
            for (int p = 0; p < (d->k + 1) / 2; p++)
                for (int c = 0; c < d->M * d->AL + d->SK; c++) DEVCHK(items_push(e, s, p, c, 0));
        }
        per[lg].n = (int)(e->hitems_n - per[lg].off);
    }
    DEVCHK(items_upload(e, e->dyn_base));
    for (int lg = 6; lg < NLG; lg++)
        if (per[lg].n)
            DEVCHK(ssa_dev_cache_build(e->prec, lg, e->dd, it_ptr(e, per[lg]), per[lg].n, e->arena_a, e->tw, e->st));
    /* later reconstructions are queued behind the build on the same stream */
    e->cache_valid = 1;
    return 0;
}

void ssa_engine_cache_free(ssa_engine *e)
{
    if (!e) return;
    e->cache_valid = 0;
    if (e->cache) {
        hipStreamSynchronize(e->st);
        hipFree(e->cache);
        e->cache = NULL;
        e->cache_cap = 0;
    }
}

int ssa_engine_cache_valid(const ssa_engine *e) { return e && e->cache_valid; }

/* ------------------------------------------------------------------------ */
/* batched step (spectra + decomposition + reconstruction of all k comps)    */
/* ------------------------------------------------------------------------ */

static int prepare_step_groups(ssa_engine *e)
{
    if (e->step_groups_ready) return 0;
    /* gid 0 of every series: components 0..k-1 (gdesc shared: glist = 0..BMAX-1) */
    int gd[4] = {0, 0, 0, 0};
    int gl[SSA_BMAX];
    for (int i = 0; i < SSA_BMAX; i++) gl[i] = i;
    /* per-series k differs: the kernels skip indices >= k, so count = BMAX covers all */
    gd[1] = SSA_BMAX;
    DEVCHK(groups_upload(e, gd, 1, gl, SSA_BMAX));
    e->step_groups_ready = 1;
    return 0;
}

static int launch_step(ssa_engine *e, int q)
{
    DEVCHK(launch_spectra(e));
    DEVCHK(launch_decompose(e, q, 0));
    for (int lg = 6; lg < NLG; lg++) {
        if (!e->rec_step[lg].n) continue;
        prof_begin(e, PH_RECON, lg);
        DEVCHK(ssa_dev_recon(e->prec, lg, 1, 1, RECON_SPLIT, e->dd, it_ptr(e, e->rec_step[lg]), e->rec_step[lg].n,
                             e->dgroups,
                             e->dgroups + 4, e->arena_a, e->small, e->tw, e->st));
        prof_end(e);
    }
    prof_begin(e, PH_OLAP, 0);
    DEVCHK(ssa_dev_olap(e->prec, 0, RECON_SPLIT, OLAP_TILE_STEP, e->dd, it_ptr(e, e->olap_step), e->olap_step.n, e->dgroups,
                        e->arena_a, e->st));
    DEVCHK(ssa_dev_edge(e->prec, 1, 0, 1, e->dd, it_ptr(e, e->ser), e->ser.n, e->dgroups, e->dgroups + 4,
                        e->arena_a, e->small, e->st));
    prof_end(e);
    if (e->step_h > 0) {
        prof_begin(e, PH_FORECAST, 0);
        DEVCHK(ssa_dev_forecast(e->prec, e->step_h, 1, 0, e->dd, it_ptr(e, e->ser), e->ser.n, e->dgroups, e->dgroups + 4,
                                e->arena_a, e->small, e->st));
        prof_end(e);
    }
    return 0;
}

int ssa_engine_step(ssa_engine *e, int q, int use_graph)
{
    if (!e->prepared || q < 0) return -1;
    /* the step (and its graph) reads the counter from the descriptor */
    if (e->omega_call >= 0) {
        e->omega_call = -1;
        e->desc_dirty = 1;
    }
    DEVCHK(sync_desc(e));
    DEVCHK(prepare_step_groups(e));
    if (!use_graph) return launch_step(e, q);
    if (!e->has_graph || e->graph_q != q) {
        if (e->has_graph) { hipGraphExecDestroy(e->gexec); e->has_graph = 0; }
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
        int rc = launch_step(e, q);
        hipError_t ce = hipStreamEndCapture(e->st, &g);
        if (rc != 0 || ce != hipSuccess) return -1;
        HIPCHK(hipGraphInstantiate(&e->gexec, g, NULL, NULL, 0));
        hipGraphDestroy(g);
        e->has_graph = 1;
        e->graph_q = q;
    }
    HIPCHK(hipGraphLaunch(e->gexec, e->st));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* autotuner: measured transform cost per FFT length                         */
/* ------------------------------------------------------------------------ */

/* For every length 2^lgn, lgn = 6..13: a throw-away engine of series with
 * N = n, L = n/4 (one chunk each way, so a column pair costs exactly two
 * transforms per block matvec), enough of them to fill the device; the
 * block matvec is timed with HIP events (best of 3) and divided by the
 * transforms it performs.  The planner then compares candidate chunkings by
 * measured device time instead of the n (log2 n + 3) model. */
int ssa_engine_autotune(int prec, double *ns_out)
{
    if (prec != SSA_PREC_F64 && prec != SSA_PREC_F32) return -1;
    double t[NLG] = {0};
    for (int lg = 6; lg <= 13; lg++) {
        const int n = 1 << lg, L = n / 4, K = n - L + 1;
        const int b = L < 60 ? L : 60, pairs = (b + 1) / 2;
        const int count = (int)((48000 + 2 * pairs - 1) / (2 * pairs));
        int *N = (int *)malloc(sizeof(int) * count), *Lv = (int *)malloc(sizeof(int) * count);
        int *bv = (int *)malloc(sizeof(int) * count), *kv = (int *)malloc(sizeof(int) * count);
        if (!N || !Lv || !bv || !kv) { free(N); free(Lv); free(bv); free(kv); return -1; }
        for (int i = 0; i < count; i++) { N[i] = n; Lv[i] = L; bv[i] = b; kv[i] = b / 2 > 0 ? b / 2 : 1; }
        (void)K;
        int err = 0;
        g_force_lgn = lg;
        ssa_engine *e = ssa_engine_create(prec, count, N, Lv, NULL, bv, kv, 2, &err);
        g_force_lgn = -1;
        free(N); free(Lv); free(bv); free(kv);
        if (!e) return -1;
        int rc = ssa_engine_prepare(e, 1, 0);
        if (rc == 0) rc = sync_desc(e);
        /* defined (zero) signals, spectra and panels: timing only */
        if (rc == 0 && (hipMemsetAsync(e->arena_a, 0, e->esz * (size_t)e->a_elems, e->st) != hipSuccess ||
                        hipMemsetAsync(e->arena_b, 0, e->esz * (size_t)e->b_elems, e->st) != hipSuccess))
            rc = -1;
        hipEvent_t e0 = NULL, e1 = NULL;
        if (rc == 0 && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)) rc = -1;
        float best = -1.0f;
        for (int rep = 0; rc == 0 && rep < 4; rep++) {
            if (hipEventRecord(e0, e->st) != hipSuccess) { rc = -1; break; }
            rc = ssa_dev_matvec(e->prec, lg, 0, SSA_PANEL_Z, SSA_PANEL_Y, e->mvf[lg].n >= e->mv_wide_min, e->dd,
                                it_ptr(e, e->mvf[lg]), e->mvf[lg].n,
                                e->arena_a, e->tw, e->st);
            if (rc || hipEventRecord(e1, e->st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = -1; break; }
            float ms = 0;
            hipEventElapsedTime(&ms, e0, e1);
            if (rep > 0 && (best < 0 || ms < best)) best = ms;   /* the first launch warms up */
        }
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        ssa_engine_destroy(e);
        if (rc || best <= 0) return -1;
        t[lg] = 1e6 * (double)best / (2.0 * pairs * count);
    }
    for (int lg = 0; lg < NLG; lg++) g_tcost[prec][lg] = t[lg];
    g_tuned[prec] = 1;
    if (ns_out) for (int lg = 6; lg <= 13; lg++) ns_out[lg - 6] = t[lg];
    return 0;
}

/* ------------------------------------------------------------------------ */
/* several engines as parallel branches of one step                          */
/* ------------------------------------------------------------------------ */

struct ssa_multi {
    ssa_engine **es;
    int G;
    hipEvent_t fork;
    hipEvent_t *join;
    hipGraphExec_t gexec;
    int has_graph, graph_q;
};

ssa_multi *ssa_multi_create(ssa_engine *const *es, int G)
{
    if (G < 1 || !es) return NULL;
    ssa_multi *m = (ssa_multi *)calloc(1, sizeof(ssa_multi));
    if (!m) return NULL;
    m->es = (ssa_engine **)malloc(sizeof(ssa_engine *) * G);
    m->join = (hipEvent_t *)calloc(G, sizeof(hipEvent_t));
    if (!m->es || !m->join) { free(m->es); free(m->join); free(m); return NULL; }
    memcpy(m->es, es, sizeof(ssa_engine *) * G);
    m->G = G;
    int ok = hipEventCreateWithFlags(&m->fork, hipEventDisableTiming) == hipSuccess;
    for (int g = 0; ok && g < G; g++) ok = hipEventCreateWithFlags(&m->join[g], hipEventDisableTiming) == hipSuccess;
    if (!ok) { ssa_multi_destroy(m); return NULL; }
    return m;
}

void ssa_multi_destroy(ssa_multi *m)
{
    if (!m) return;
    if (m->has_graph) hipGraphExecDestroy(m->gexec);
    if (m->fork) hipEventDestroy(m->fork);
    for (int g = 0; g < m->G; g++) if (m->join[g]) hipEventDestroy(m->join[g]);
    free(m->join);
    free(m->es);
    free(m);
}

/* every engine's step on its own stream, forked from and joined into the
 * first engine's stream */
static int multi_launch(ssa_multi *m, int q)
{
    HIPCHK(hipEventRecord(m->fork, m->es[0]->st));
    for (int g = 1; g < m->G; g++) HIPCHK(hipStreamWaitEvent(m->es[g]->st, m->fork, 0));
    for (int g = 0; g < m->G; g++) DEVCHK(launch_step(m->es[g], q));
    for (int g = 1; g < m->G; g++) {
        HIPCHK(hipEventRecord(m->join[g], m->es[g]->st));
        HIPCHK(hipStreamWaitEvent(m->es[0]->st, m->join[g], 0));
    }
    return 0;
}

int ssa_multi_step(ssa_multi *m, int q, int use_graph)
{
    if (!m || q < 0) return -1;
    for (int g = 0; g < m->G; g++) {
        ssa_engine *e = m->es[g];
        if (!e->prepared) return -1;
        DEVCHK(sync_desc(e));
        DEVCHK(prepare_step_groups(e));
    }
    /* a graph launched on the first stream is not ordered after the other
     * branches' streams: their descriptor uploads (not waited for on the
     * host) must complete first */
    for (int g = 1; g < m->G; g++)
        if (m->es[g]->dpin_busy) HIPCHK(hipStreamWaitEvent(m->es[0]->st, m->es[g]->dpin_ev, 0));
    if (!use_graph) return multi_launch(m, q);
    if (!m->has_graph || m->graph_q != q) {
        if (m->has_graph) { hipGraphExecDestroy(m->gexec); m->has_graph = 0; }
        hipGraph_t gr;
        HIPCHK(hipStreamBeginCapture(m->es[0]->st, hipStreamCaptureModeThreadLocal));
        int rc = multi_launch(m, q);
        hipError_t ce = hipStreamEndCapture(m->es[0]->st, &gr);
        if (rc != 0 || ce != hipSuccess) return -1;
        HIPCHK(hipGraphInstantiate(&m->gexec, gr, NULL, NULL, 0));
        hipGraphDestroy(gr);
        m->has_graph = 1;
        m->graph_q = q;
    }
    HIPCHK(hipGraphLaunch(m->gexec, m->es[0]->st));
    return 0;
}

/* the whole step is joined into the first engine's stream */
int ssa_multi_sync(ssa_multi *m) { return m ? ssa_engine_sync(m->es[0]) : -1; }

int ssa_multi_query(ssa_multi *m) { return m ? ssa_engine_query(m->es[0]) : -1; }

int ssa_engine_profile_step(ssa_engine *e, int q, int max_rec, int *phase, int *lgn, float *ms)
{
    if (!e->prepared || q < 0 || max_rec < 1) return -1;
    DEVCHK(sync_desc(e));
    DEVCHK(prepare_step_groups(e));
    if (e->prof_cap < max_rec) {
        for (int i = 0; i < 2 * e->prof_cap; i++) hipEventDestroy(e->prof_ev[i]);
        free(e->prof_ev); free(e->prof_phase); free(e->prof_lgn);
        e->prof_ev = (hipEvent_t *)calloc(2 * (size_t)max_rec, sizeof(hipEvent_t));
        e->prof_phase = (int *)calloc(max_rec, sizeof(int));
        e->prof_lgn = (int *)calloc(max_rec, sizeof(int));
        if (!e->prof_ev || !e->prof_phase || !e->prof_lgn) { e->prof_cap = 0; return -1; }
        for (int i = 0; i < 2 * max_rec; i++) HIPCHK(hipEventCreate(&e->prof_ev[i]));
        e->prof_cap = max_rec;
    }
    e->prof_on = 1;
    e->prof_n = 0;
    int rc = launch_step(e, q);
    e->prof_on = 0;
    HIPCHK(hipStreamSynchronize(e->st));
    if (rc) return -1;
    for (int i = 0; i < e->prof_n; i++) {
        phase[i] = e->prof_phase[i];
        lgn[i] = e->prof_lgn[i];
        HIPCHK(hipEventElapsedTime(&ms[i], e->prof_ev[2 * i], e->prof_ev[2 * i + 1]));
    }
    return e->prof_n;
}

/* every series' step output (N + step h values each, FP64) into out at
 * offsets[s], and the device status words, with one synchronisation */
int ssa_engine_step_outputs_all(ssa_engine *e, double *out, const long long *offsets, int *status)
{
    if (!e->prepared || !out || !offsets) return -1;
    DEVCHK(sync_desc(e));
    /* every series' output into its gather slot (one after the other in the
     * panel arena), then one copy of the whole gather region */
    DEVCHK(ssa_dev_gather(e->prec, e->step_h, e->dd, it_ptr(e, e->ser), e->ser.n, e->arena_a, e->st));
    const ssa_desc *first = &e->hd[0], *last = &e->hd[e->count - 1];
    const long long span_n = last->off_gather + last->N + e->step_h - first->off_gather;
    const size_t dbytes = status ? sizeof(ssa_desc) * (size_t)e->count : 0;
    const size_t obytes = ((size_t)span_n * e->esz + 255) & ~(size_t)255;
    DEVCHK(pin_reserve(e, obytes + dbytes));
    HIPCHK(hipMemcpyAsync(e->hpin, (char *)e->arena_a + e->esz * first->off_gather, (size_t)span_n * e->esz,
                          hipMemcpyDeviceToHost, e->st));
    ssa_desc *hd = (ssa_desc *)((char *)e->hpin + obytes);
    if (status) HIPCHK(hipMemcpyAsync(hd, e->dd, dbytes, hipMemcpyDeviceToHost, e->st));
    HIPCHK(hipStreamSynchronize(e->st));
    for (int s = 0; s < e->count; s++) {
        const long long n = e->hd[s].N + e->step_h, at = e->hd[s].off_gather - first->off_gather;
        double *o = out + offsets[s];
        if (e->prec == SSA_PREC_F64) memcpy(o, (const double *)e->hpin + at, n * 8);
        else {
            const float *src = (const float *)e->hpin + at;
            for (long long i = 0; i < n; i++) o[i] = src[i];
        }
        if (status) status[s] = hd[s].status;
    }
    return 0;
}

int ssa_engine_step_outputs(ssa_engine *e, int s, double *out)
{
    if (!e->prepared || s < 0 || s >= e->count) return -1;
    return fetch_out(e, s, (long long)e->hd[s].N + e->step_h, out);
}
