/*
 * mpic_launch.h -- kernels as (argument struct, body) pairs, so that one
 * launch can serve many runs.
 *
 * Every per-step kernel of the library is written as the body of a
 * KernelDef<A> specialisation: a __device__ function of its argument struct
 * A and of a block index and grid size passed in (VBlock), instead of the
 * blockIdx / gridDim builtins.  Two __global__ wrappers run a body:
 *
 *   k_one<A>    one run: the launch's own grid (what the stage calls use);
 *   k_multi<A>  many runs in one launch: run r owns blocks [tab[r],
 *               tab[r+1]) of a 1-D grid, each mapped back to a block of that
 *               run's own grid, and reads its arguments from a device array.
 *
 * launch_def<A>() launches a body on a stream, or, while a Recorder is set
 * on the calling thread, appends the launch to it as data (kind, grid,
 * argument bytes).  A run's step graphs are the recorded launches as
 * kernel nodes; the ensemble code records one step of every run and fuses
 * the launches of equal kind across runs into k_multi nodes: one launch
 * per stage per step for the whole ensemble.
 */
#ifndef MPIC_LAUNCH_H
#define MPIC_LAUNCH_H

#include <string.h>

#include <type_traits>
#include <vector>

#include "mpic_internal.h"

struct VBlock {
    unsigned x, y, z;
};

/* Specialised per argument struct: kBlock (threads per workgroup) and
 * body(const A &, VBlock block, VBlock grid); optionally kWaves, the
 * wavefronts per SIMD the compiler must keep room for (its register
 * budget). */
template <class A> struct KernelDef;

template <class A, class = void> struct WavesOf {
    static constexpr int v = 0;
};
template <class A> struct WavesOf<A, std::void_t<decltype(KernelDef<A>::kWaves)>> {
    static constexpr int v = KernelDef<A>::kWaves;
};

/* Read-only data of a launch in the constant address space: its loads are
 * invariant, so the bodies read argument fields in place, when they are
 * used, instead of holding whole structs in registers (or copying them to
 * scratch) for the kernel's life. */
#define CAS __attribute__((address_space(4)))

/* A one-run launch's argument struct is its first kernel argument, at the
 * start of the kernel-argument segment: the body reads it there.  With the
 * struct passed by value the compiler keeps every field it uses live in
 * SGPRs; the push kernels' structs then spilled (46 SGPRs in the FP32 push,
 * 247 VGPRs for the FP64 push instead of 164). */
template <class A>
__device__ static inline void run_one()
{
    const CAS A *ap = (const CAS A *)__builtin_amdgcn_kernarg_segment_ptr();
    const VBlock b = {blockIdx.x, blockIdx.y, blockIdx.z}, g = {gridDim.x, gridDim.y, gridDim.z};
    KernelDef<A>::body(*(const A *)ap, b, g);
}

/* tab: [n + 1] first block of every run, then [3 n] the runs' own grids. */
template <class A>
__device__ static inline void run_multi(const CAS A *__restrict__ args,
                                        const CAS unsigned *__restrict__ tab, int n)
{
    const unsigned blk = blockIdx.x;
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (tab[mid] <= blk) lo = mid;
        else hi = mid;
    }
    const CAS unsigned *gr = tab + (n + 1) + 3 * lo;
    const VBlock g = {gr[0], gr[1], gr[2]};
    const unsigned L = blk - tab[lo];
    const VBlock b = {L % g.x, (L / g.x) % g.y, L / (g.x * g.y)};
    KernelDef<A>::body(*(const A *)(args + lo), b, g);
}

template <class A> __global__ void __launch_bounds__(KernelDef<A>::kBlock) k_one(A)
{
    run_one<A>();
}
template <class A>
__global__ void __launch_bounds__(KernelDef<A>::kBlock)
    k_multi(const A *__restrict__ args, const unsigned *__restrict__ tab, int n)
{
    run_multi((const CAS A *)args, (const CAS unsigned *)tab, n);
}
template <class A>
__global__ void __launch_bounds__(KernelDef<A>::kBlock)
    __attribute__((amdgpu_waves_per_eu(WavesOf<A>::v))) k_one_w(A)
{
    run_one<A>();
}
template <class A>
__global__ void __launch_bounds__(KernelDef<A>::kBlock)
    __attribute__((amdgpu_waves_per_eu(WavesOf<A>::v)))
    k_multi_w(const A *__restrict__ args, const unsigned *__restrict__ tab, int n)
{
    run_multi((const CAS A *)args, (const CAS unsigned *)tab, n);
}

template <class A> static inline void start_one(hipStream_t s, const A &a, dim3 grid)
{
    if constexpr (WavesOf<A>::v > 0)
        k_one_w<A><<<grid, KernelDef<A>::kBlock, 0, s>>>(a);
    else
        k_one<A><<<grid, KernelDef<A>::kBlock, 0, s>>>(a);
}

/* Two bodies with no order between them in one launch (same block size):
 * blocks [0, na) run A over its own grid ga, the others B over gb (both
 * flattened, x fastest).  Every tiny launch costs a few microseconds of
 * stream time, so independent neighbours of a step share one. */
template <class A, class B> struct PairArgs {
    A        a;
    B        b;
    VBlock   ga, gb;
    unsigned na;
};

__device__ static inline VBlock unflat(unsigned L, const VBlock &g)
{
    const VBlock b = {L % g.x, (L / g.x) % g.y, L / (g.x * g.y)};
    return b;
}

template <class A, class B> struct KernelDef<PairArgs<A, B>> {
    static_assert(KernelDef<A>::kBlock == KernelDef<B>::kBlock, "a pair shares one block size");
    static constexpr int kBlock = KernelDef<A>::kBlock;
    __device__ __forceinline__ static void body(const PairArgs<A, B> &p, VBlock blk, VBlock)
    {
        const unsigned L = blk.x;            /* pairs are launched over a 1-D grid */
        if (L < p.na)
            KernelDef<A>::body(p.a, unflat(L, p.ga), p.ga);
        else
            KernelDef<B>::body(p.b, unflat(L - p.na, p.gb), p.gb);
    }
};

/* One launch as data. */
struct LaunchRec {
    const void *kind;                 /* identifies the body (KernelDef<A>::tag) */
    unsigned    grid[3];
    std::vector<unsigned char> args;
    /* k_multi<A> over n runs as a kernel node of graph gr, after node dep
     * (none: nullptr): dargs = device array of n argument structs, dtab =
     * the block table, blocks = tab[n] */
    int (*multi_node)(hipGraph_t gr, hipGraphNode_t dep, hipGraphNode_t *node, const void *dargs,
                      const unsigned *dtab, int n, unsigned blocks);
    /* k_one<A> of this record as a kernel node */
    int (*one_node)(hipGraph_t gr, hipGraphNode_t dep, hipGraphNode_t *node, const void *args,
                    const unsigned grid[3]);
};

struct Recorder {
    std::vector<LaunchRec> recs;
};

/* Set while a run's step is recorded instead of launched (calling thread). */
extern thread_local Recorder *g_recorder;

template <class A> struct KernelTag {
    static const char tag;
};
template <class A> const char KernelTag<A>::tag = 0;

/* Graphs are built node by node from recorded launches, not by stream
 * capture: building one then needs no stream, and device calls made by
 * other host threads meanwhile cannot invalidate it.  The node copies the
 * kernel arguments. */
static inline int add_kernel_node(hipGraph_t gr, hipGraphNode_t dep, hipGraphNode_t *node,
                                  const void *func, dim3 grid, unsigned block, void **params)
{
    hipKernelNodeParams p;
    memset(&p, 0, sizeof(p));
    p.func = (void *)func;
    p.gridDim = grid;
    p.blockDim = dim3(block);
    p.sharedMemBytes = 0;
    p.kernelParams = params;
    p.extra = nullptr;
    MPIC_CHECK(hipGraphAddKernelNode(node, gr, dep ? &dep : nullptr, dep ? 1 : 0, &p));
    return MPIC_OK;
}

template <class A>
static int multi_node(hipGraph_t gr, hipGraphNode_t dep, hipGraphNode_t *node, const void *dargs,
                      const unsigned *dtab, int n, unsigned blocks)
{
    const A *pa = (const A *)dargs;
    const unsigned *pt = dtab;
    int pn = n;
    void *params[3] = {(void *)&pa, (void *)&pt, (void *)&pn};
    const void *f;
    if constexpr (WavesOf<A>::v > 0)
        f = (const void *)&k_multi_w<A>;
    else
        f = (const void *)&k_multi<A>;
    return add_kernel_node(gr, dep, node, f, dim3(blocks), KernelDef<A>::kBlock, params);
}

template <class A>
static int one_node(hipGraph_t gr, hipGraphNode_t dep, hipGraphNode_t *node, const void *args,
                    const unsigned grid[3])
{
    A a;
    memcpy(&a, args, sizeof(A));
    void *params[1] = {(void *)&a};
    const void *f;
    if constexpr (WavesOf<A>::v > 0)
        f = (const void *)&k_one_w<A>;
    else
        f = (const void *)&k_one<A>;
    return add_kernel_node(gr, dep, node, f, dim3(grid[0], grid[1], grid[2]), KernelDef<A>::kBlock,
                           params);
}

/* Launches (or records) body A over a gx x gy x gz grid; an empty grid is a
 * no-op. */
template <class A>
static int launch_def(hipStream_t s, const A &a, unsigned gx, unsigned gy = 1, unsigned gz = 1)
{
    if (gx == 0 || gy == 0 || gz == 0) return MPIC_OK;
    if (g_recorder) {
        LaunchRec r;
        r.kind = &KernelTag<A>::tag;
        r.grid[0] = gx;
        r.grid[1] = gy;
        r.grid[2] = gz;
        r.args.resize(sizeof(A));
        memcpy(r.args.data(), &a, sizeof(A));
        r.multi_node = &multi_node<A>;
        r.one_node = &one_node<A>;
        g_recorder->recs.push_back(std::move(r));
        return MPIC_OK;
    }
    start_one(s, a, dim3(gx, gy, gz));
    MPIC_CHECK(hipGetLastError());
    return MPIC_OK;
}

/* A over grid ga and B over grid gb in one launch (one of them alone when
 * the other's grid is empty). */
template <class A, class B>
static int launch_pair(hipStream_t s, const A &a, dim3 ga, const B &b, dim3 gb)
{
    const unsigned na = ga.x * ga.y * ga.z, nb = gb.x * gb.y * gb.z;
    if (na == 0) return launch_def(s, b, gb.x, gb.y, gb.z);
    if (nb == 0) return launch_def(s, a, ga.x, ga.y, ga.z);
    PairArgs<A, B> p;
    p.a = a;
    p.b = b;
    p.ga = {ga.x, ga.y, ga.z};
    p.gb = {gb.x, gb.y, gb.z};
    p.na = na;
    return launch_def(s, p, na + nb);
}

#endif
