// ssa_noalloc.hip -- the reconstruction of ssa_opt_reconstruct (MKL/ssa_opt.h:2325-2400)
// for callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).
//
// The caller holds U (L x k), V (K x k) and sigma (k) in device memory and
// passes a device output of N values.  Everything else the kernels need -- the
// series descriptor, the work lists, the group and the staged convolution
// pieces -- lives in `scratch`, a static device array of SSA_NOALLOC_SCRATCH_BYTES
// per GPU.  A call allocates nothing: a setup kernel writes the descriptor and
// lists into scratch from its by-value arguments (no host-to-device copy), then
// the reconstruction, overlap-add and edge kernels of the context path run on
// the caller's stream.  The call returns once the work is queued.
//
// Every call on a GPU shares that GPU's scratch: calls that overlap on one GPU
// (two streams, two threads) race and must be ordered by the caller.  Nothing
// in scratch has to start at zero: the setup kernel writes every descriptor,
// list and group word the later kernels read, and every staged piece is written
// in full before the overlap-add reads it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "ssa_device.h"
#include "ssa_engine.h"

#define SSA_NOALLOC_SCRATCH_BYTES (32LL << 20)
#define SSA_NOALLOC_MAX_GROUP 256
#define SSA_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) double scratch[SSA_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

struct GroupArg {
    int idx[SSA_NOALLOC_MAX_GROUP];
};

// Descriptor, items and group into scratch.  Items: AL*SK reconstruction pieces
// (s 0, group 0, a, sk), then n_olap overlap-add tiles (s 0, group 0, tile),
// then one edge item (s 0, group 0).
__global__ __launch_bounds__(256) void k_noalloc_setup(ssa_desc d, ssa_desc *dd, ssa_item *items, int AL, int SK,
                                                      int n_olap, int4 *gdesc, int *glist, GroupArg g, int n_group)
{
    const int tid = threadIdx.x;
    if (tid == 0) {
        *dd = d;
        *gdesc = make_int4(0, n_group, 0, 0);
    }
    for (int i = tid; i < n_group; i += 256) glist[i] = g.idx[i];
    const int n_rec = AL * SK;
    for (int i = tid; i < n_rec + n_olap + 1; i += 256) {
        ssa_item it = {0, 0, 0, 0};
        if (i < n_rec) {
            it.y = i / SK;
            it.z = i - it.y * SK;
        } else if (i < n_rec + n_olap) {
            it.y = i - n_rec;
        }
        items[i] = it;
    }
}

struct DeviceState {
    double *base;      // hipGetSymbolAddress(scratch) on this device
    const void *tw;    // FP64 twiddle tables on this device
};

DeviceState g_state[SSA_NOALLOC_MAX_DEVICES];
std::mutex g_lock;

// once per device, with the device current: the scratch address (never the
// host-side name of the array) and the twiddle tables (filled on the caller's
// stream the first time)
int device_state(DeviceState *out, void *stream)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= SSA_NOALLOC_MAX_DEVICES) return -1;
    std::lock_guard<std::mutex> guard(g_lock);
    DeviceState &s = g_state[dev];
    if (!s.base) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) {
            fprintf(stderr, "ssa_amd_reconstruct_noalloc: hipGetSymbolAddress(scratch) failed: %s\n",
                    hipGetErrorString(e));
            return (int)e;
        }
        void *tw = nullptr;
        const int rc = ssa_dev_twiddles(SSA_PREC_F64, &tw, stream);
        if (rc != 0) return rc;
        s.tw = tw;
        s.base = (double *)p;
    }
    *out = s;
    return 0;
}

long long align32(long long v) { return (v + 31) & ~31LL; }   // doubles (256 bytes)

}  // namespace

extern "C" int ssa_amd_reconstruct_noalloc(const double *U, const double *V, const double *sigma, int N, int L,
                                           int k, const int *group, int n_group, double *output, void *stream)
{
    if (!U || !V || !sigma || !output || !group || N < 4 || L < 2 || L > N - 1 || k < 1) return -1;
    if (n_group < 1) return -1;   // MKL/ssa_opt.h:2327-2328
    if (n_group > SSA_NOALLOC_MAX_GROUP) {
        fprintf(stderr, "ssa_amd_reconstruct_noalloc: a group of %d exceeds the maximum of %d\n", n_group,
                SSA_NOALLOC_MAX_GROUP);
        return -2;
    }
    const uintptr_t mis = ((uintptr_t)U | (uintptr_t)V | (uintptr_t)sigma | (uintptr_t)output) & 7u;
    if (mis) return -1;
    // transform length and chunking: the planner's choice for a reconstruction
    // alone (q = -1 leaves only the reconstruction term in its cost)
    ssa_plan pl;
    if (ssa_plan_series(SSA_PREC_F64, N, L, 1, 2, k, -1, -1, &pl) != 0) return -1;
    const int K = N - L + 1;
    const int n_rec = pl.AL * pl.SK, n_olap = (N + SSA_OLAP_TILE - 1) / SSA_OLAP_TILE;
    // scratch layout (doubles): descriptor | items | group | staging
    const long long o_desc = 0;
    const long long o_items = align32(o_desc + (long long)(sizeof(ssa_desc) + 7) / 8);
    const long long o_group = align32(o_items + ((long long)(n_rec + n_olap + 1) * (long long)sizeof(ssa_item) + 7) / 8);
    const long long o_stage = align32(o_group + (16 + 4LL * SSA_NOALLOC_MAX_GROUP + 7) / 8);
    const long long stage = (long long)n_rec * pl.n;
    const long long need = o_stage + stage;
    if (need * (long long)sizeof(double) > SSA_NOALLOC_SCRATCH_BYTES) {
        fprintf(stderr, "ssa_amd_reconstruct_noalloc: N = %d, L = %d needs %lld bytes of scratch; the maximum is %lld\n",
                N, L, need * (long long)sizeof(double), (long long)SSA_NOALLOC_SCRATCH_BYTES);
        return -2;
    }
    DeviceState st;
    int rc = device_state(&st, stream);
    if (rc != 0) return rc;
    double *base = st.base;
    // element offsets from the scratch base: the caller's buffers are addressed
    // through the same base (as the engine addresses its second arena)
    ssa_desc d;
    memset(&d, 0, sizeof(d));
    d.N = N; d.L = L; d.K = K; d.M = 1; d.ML = L;
    d.b = k; d.k = k; d.gsize = k;
    d.n = pl.n; d.lgn = pl.lgn; d.CL = pl.CL; d.CK = pl.CK; d.AL = pl.AL; d.SK = pl.SK;
    d.off_U = (long long)(U - base);
    d.off_V = (long long)(V - base);
    d.off_out = (long long)(output - base);
    d.off_stage = o_stage;
    d.stage_cap = stage;
    d.out_cap = N;
    // k_recon and k_edge read sigma at small + off_small + SLOT_MISC * BMAX^2 + MISC_SIGMA
    d.off_small = (long long)(sigma - base) - (long long)SSA_SLOT_MISC * SSA_BMAX * SSA_BMAX - SSA_MISC_SIGMA;
    GroupArg g;
    memset(&g, 0, sizeof(g));
    memcpy(g.idx, group, sizeof(int) * (size_t)n_group);
    ssa_desc *dd = (ssa_desc *)(base + o_desc);
    ssa_item *items = (ssa_item *)(base + o_items);
    int *gdesc = (int *)(base + o_group);
    int *glist = gdesc + 4;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(256), 0, s, d, dd, items, pl.AL, pl.SK, n_olap,
                       (int4 *)gdesc, glist, g, n_group);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return (int)e;
    rc = ssa_dev_recon(SSA_PREC_F64, pl.lgn, 1, 0, 1, dd, items, n_rec, gdesc, glist, base, base, st.tw, stream);
    if (rc == 0) rc = ssa_dev_olap(SSA_PREC_F64, 0, 1, SSA_OLAP_TILE, dd, items + n_rec, n_olap, gdesc, base, stream);
    if (rc == 0) rc = ssa_dev_edge(SSA_PREC_F64, 1, 0, 0, dd, items + n_rec + n_olap, 1, gdesc, glist, base, base,
                                   stream);
    return rc;
}
