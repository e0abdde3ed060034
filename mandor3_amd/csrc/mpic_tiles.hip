/*
 * mpic_tiles.hip -- per-step counting sort of the markers by tile-major
 * cell key.
 *
 * Per step: the push kernel writes each marker's new key and bumps
 * count[key] once per run of equal keys in a wavefront; the scan turns the
 * counts into per-key offsets, builds the list of non-empty tiles and resets
 * the tile queues; the rank pass gives every live marker its slot
 * (cursor[key] reserved once per run) and records the marker's position in
 * the next order's permutation, dropping deleted markers.  The next push
 * reads its markers through that permutation and writes them to the other
 * buffer in order, so no pass moves markers only to sort them (the untiled
 * FP64 push, which keeps no order, is compacted by a scatter instead).
 *
 * Roofline (HBM-bound): scan ~ 3 x 4 B per cell key; rank 4 B read and 4 B
 * written per marker; scatter 4 + 60 + 60 B per FP64 marker.
 */
#include "mpic_launch.h"
#include "mpic_tiles.h"

#define SB MPIC_SCAN_BLOCK       /* tiles per scan block */

/* Keys of the current array, with per-run counting.  Only the first
 * nmark[parity] markers are live; the buffer past them holds records of
 * earlier orders (or of loaded markers that the walls deleted), which must
 * not be counted. */
__global__ void k_keys64(Markers64 m, long n, Grid g, TileView tv, int parity)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long live = tv.nmark[parity];
    int key = -1;
    if (p < n && p < live && m.qm[p] != 0.0f) {
        const double r[3] = {m.x[p], m.y[p], m.z[p]};
        key = cell_key_of(tv, g, r);
    }
    if (p < n) tv.key[p] = key;
    wave_run_count(tv.count, key);
}

int launch_keys64(hipStream_t s, const Markers64 &m, long n, const Grid &g, const TileView &tv,
                  int parity)
{
    if (n <= 0) return MPIC_OK;
    k_keys64<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(m, n, g, tv, parity);
    MPIC_CHECK(hipGetLastError());
    return MPIC_OK;
}

/* Block-local exclusive scans of (markers per tile, tile non-empty, work
 * items of the tile), three ints per tile and per block. */
#define NSCAN 3
struct ScanAArgs {
    TileView tv;
    int     *blocksum, *tileloc;
};

__device__ static inline void scan_a(const TileView &tv, int *blocksum, int *tileloc,
                                     const VBlock &bk)
{
    __shared__ int wsum[NSCAN][SB / 64];
    const int t = bk.x * SB + threadIdx.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int c = 0;
    if (t < tv.ntiles) {
        const int4 *q = (const int4 *)(tv.count + (long)t * TILE_CELLS);
        for (int v = 0; v < TILE_CELLS / 4; ++v) {
            const int4 x = q[v];
            c += x.x + x.y + x.z + x.w;
        }
    }
    const int own[NSCAN] = {c, c > 0, tile_items(c, tv.chm)};
    int inc[NSCAN];
    for (int a = 0; a < NSCAN; ++a) {        /* wavefront inclusive scan */
        int x = own[a];
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(x, d, 64);
            if (lane >= d) x += y;
        }
        inc[a] = x;
        if (lane == 63) wsum[a][w] = x;
    }
    __syncthreads();
    for (int a = 0; a < NSCAN; ++a)
        for (int q = 0; q < w; ++q) inc[a] += wsum[a][q];
    if (t < tv.ntiles)
        for (int a = 0; a < NSCAN; ++a) tileloc[NSCAN * t + a] = inc[a] - own[a];
    if (threadIdx.x == SB - 1)
        for (int a = 0; a < NSCAN; ++a) blocksum[NSCAN * bk.x + a] = inc[a];
}

template <> struct KernelDef<ScanAArgs> {
    static constexpr int kBlock = SB;
    __device__ __forceinline__ static void body(const ScanAArgs &a, VBlock b, VBlock)
    {
        scan_a(a.tv, a.blocksum, a.tileloc, b);
    }
};

/* Exclusive scan of the block sums (one block, loops for any count). */
struct ScanBArgs {
    int *blocksum;
    int  nb;
};

__device__ static inline void scan_b(int *blocksum, int nb)
{
    __shared__ int sv[NSCAN][SB];
    __shared__ int carry[NSCAN];
    if (threadIdx.x < NSCAN) carry[threadIdx.x] = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nb; b0 += SB) {
        const int b = b0 + threadIdx.x;
        int own[NSCAN];
        for (int a = 0; a < NSCAN; ++a) {
            own[a] = (b < nb) ? blocksum[NSCAN * b + a] : 0;
            sv[a][threadIdx.x] = own[a];
        }
        __syncthreads();
        for (int d = 1; d < SB; d <<= 1) {
            int x[NSCAN] = {0, 0, 0};
            if ((int)threadIdx.x >= d)
                for (int a = 0; a < NSCAN; ++a) x[a] = sv[a][threadIdx.x - d];
            __syncthreads();
            for (int a = 0; a < NSCAN; ++a) sv[a][threadIdx.x] += x[a];
            __syncthreads();
        }
        if (b < nb)
            for (int a = 0; a < NSCAN; ++a)
                blocksum[NSCAN * b + a] = carry[a] + sv[a][threadIdx.x] - own[a];
        __syncthreads();
        if (threadIdx.x == SB - 1)
            for (int a = 0; a < NSCAN; ++a) carry[a] += sv[a][SB - 1];
        __syncthreads();
    }
    if (threadIdx.x < NSCAN) blocksum[NSCAN * nb + threadIdx.x] = carry[threadIdx.x];
}

template <> struct KernelDef<ScanBArgs> {
    static constexpr int kBlock = SB;
    __device__ __forceinline__ static void body(const ScanBArgs &a, VBlock, VBlock) { scan_b(a.blocksum, a.nb); }
};

/* Tile offsets, the active-tile list and, for every non-empty tile, the
 * per-key offsets and cursors with its counters reset.  One wavefront per
 * SC_TPW consecutive tiles: lane l < SC_TPW keeps the books of tile l (its
 * total comes from the tile scan), then the wavefront takes the non-empty
 * tiles one after the other, one lane per cell key (coalesced).  Empty
 * tiles (their counters are all 0) cost two words: most tiles of a sparse
 * plasma are empty, and fewer wavefronts than tiles keep the launch short. */
#define SC_TPW 8
struct ScanCArgs {
    TileView   tv;
    const int *blocksum, *tileloc;
    int        nb, parity_next;
};

__device__ static inline void scan_c(const TileView &tv, const int *blocksum, const int *tileloc,
                                     int nb, int parity_next, const VBlock &bk)
{
    const int lane = threadIdx.x & 63;
    const int w = (int)(((long)bk.x * 256 + threadIdx.x) >> 6);
    const int t0 = w * SC_TPW;
    if (t0 >= tv.ntiles) return;                     /* whole waves exit together */
    int base = 0, tot = 0;
    if (lane < SC_TPW && t0 + lane < tv.ntiles) {
        const int t = t0 + lane, tb = t / SB, tn = t + 1;
        base = blocksum[NSCAN * tb] + tileloc[NSCAN * t];
        const int next = tn < tv.ntiles ? blocksum[NSCAN * (tn / SB)] + tileloc[NSCAN * tn]
                                        : blocksum[NSCAN * nb];
        tot = next - base;
        tv.toff[t] = base;
        const int a = blocksum[NSCAN * tb + 1] + tileloc[NSCAN * t + 1];
        if (tot > 0) {
            const int is = blocksum[NSCAN * tb + 2] + tileloc[NSCAN * t + 2];
            tv.active[a] = t;
            tv.istart[a] = is;
            const int ni = tile_items(tot, tv.chm);
            for (int u = 0; u < ni; ++u) tv.item[is + u] = a;
        }
        tv.slot[t] = tot > 0 ? a : -1;
    }
    if (w == 0 && lane == 0) {
        tv.off[tv.ncellp] = blocksum[NSCAN * nb];
        tv.toff[tv.ntiles] = blocksum[NSCAN * nb];
        *tv.nactive = blocksum[NSCAN * nb + 1];
        tv.istart[blocksum[NSCAN * nb + 1]] = blocksum[NSCAN * nb + 2];
        *tv.nitems = blocksum[NSCAN * nb + 2];
        tv.nmark[parity_next] = blocksum[NSCAN * nb];
        for (int x = 0; x < MPIC_XCDS; ++x) tv.queue[x] = 0;
    }
    for (int it = 0; it < SC_TPW; ++it) {
        if (__shfl(tot, it, 64) == 0) continue;      /* wave-uniform */
        const int tb0 = __shfl(base, it, 64);
        const long q = (long)(t0 + it) * TILE_CELLS + lane;
        const int c = tv.count[q];
        int incl = c;
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d, 64);
            if (lane >= d) incl += v;
        }
        const int o = tb0 + incl - c;
        tv.off[q] = o;
        tv.cursor[q] = o;
        tv.count[q] = 0;
    }
}

template <> struct KernelDef<ScanCArgs> {
    static constexpr int kBlock = 256;
    __device__ __forceinline__ static void body(const ScanCArgs &a, VBlock b, VBlock)
    {
        scan_c(a.tv, a.blocksum, a.tileloc, a.nb, a.parity_next, b);
    }
};

int launch_tile_scan(hipStream_t s, const TileView &tv, int *blocksum, int parity_next)
{
    const int nb = (tv.ntiles + SB - 1) / SB;
    int *tileloc = blocksum + NSCAN * (nb + 1);
    ScanAArgs a = {tv, blocksum, tileloc};
    int rc = launch_def(s, a, (unsigned)nb);
    if (rc != MPIC_OK) return rc;
    ScanBArgs b = {blocksum, nb};
    rc = launch_def(s, b, 1);
    if (rc != MPIC_OK) return rc;
    const long waves = (tv.ntiles + SC_TPW - 1) / SC_TPW;
    ScanCArgs c = {tv, blocksum, tileloc, nb, parity_next};
    return launch_def(s, c, (unsigned)((waves * 64 + 255) / 256));
}

/* The next order: every position p < nmark[parity] with a key (the push's
 * output, or the markers keyed by k_keys32 / k_keys64) takes slot
 * cursor[key]++ among its key's, one atomic per run of equal keys in a
 * wavefront, and perm[slot] = p.  4 B read and 4 B written per marker
 * instead of moving its records (36 B FP32, 60 B FP64).  A
 * workgroup ranks RANK_BATCHES consecutive batches of 256 positions with
 * all their loads, then all their atomics, in flight together: one
 * batch's chain of dependent accesses would leave the pass latency-bound. */
#define RANK_BATCHES 4
struct RankArgs {
    TileView tv;
    long     nbound;
    int      parity;
};

__device__ static inline void rank_pass(const TileView &tv, long nbound, int parity, long b0)
{
    const long n = tv.nmark[parity];
    int key[RANK_BATCHES], start[RANK_BATCHES], base[RANK_BATCHES];
#pragma unroll
    for (int u = 0; u < RANK_BATCHES; ++u) {
        const long p = b0 + 256 * u + threadIdx.x;
        key[u] = (p < n && p < nbound) ? tv.key[p] : -1;
    }
#pragma unroll
    for (int u = 0; u < RANK_BATCHES; ++u) {
        int len;
        bool head;
        wave_run(key[u], start[u], len, head);
        base[u] = 0;
        if (head && key[u] >= 0) base[u] = atomicAdd(tv.cursor + key[u], len);
    }
#pragma unroll
    for (int u = 0; u < RANK_BATCHES; ++u) {
        const int bs = __shfl(base[u], start[u], 64);
        if (key[u] >= 0)
            tv.perm[bs + ((int)(threadIdx.x & 63) - start[u])] = (int)(b0 + 256 * u + threadIdx.x);
    }
}

template <> struct KernelDef<RankArgs> {
    static constexpr int kBlock = 256;
    __device__ __forceinline__ static void body(const RankArgs &a, VBlock b, VBlock)
    {
        rank_pass(a.tv, a.nbound, a.parity, (long)b.x * (256 * RANK_BATCHES));
    }
};

int launch_rank(hipStream_t s, long nbound, const TileView &tv, int parity)
{
    if (nbound <= 0) return MPIC_OK;
    RankArgs a = {tv, nbound, parity};
    const long per = 256L * RANK_BATCHES;
    return launch_def(s, a, (unsigned)((nbound + per - 1) / per));
}

/* Moves each live marker of src to its slot in dst (the untiled FP64 push's
 * compacting sort). */
struct Scatter64Args {
    Markers64 src, dst;
    long      nbound;
    TileView  tv;
    int       parity;
};

__device__ static inline void scatter64(const Markers64 &src, const Markers64 &dst, long nbound,
                                        const TileView &tv, int parity, long p)
{
    const long n = tv.nmark[parity];
    const int key = (p < n && p < nbound) ? tv.key[p] : -1;
    int start, len;
    bool head;
    wave_run(key, start, len, head);
    int base = 0;
    if (head && key >= 0) base = atomicAdd(tv.cursor + key, len);
    base = __shfl(base, start, 64);
    if (key < 0) return;
    const long d = base + ((int)(threadIdx.x & 63) - start);
    /* every marker is streamed once: nontemporal, so the fields stay cached */
    __builtin_nontemporal_store(__builtin_nontemporal_load(src.x + p), dst.x + d);
    __builtin_nontemporal_store(__builtin_nontemporal_load(src.y + p), dst.y + d);
    __builtin_nontemporal_store(__builtin_nontemporal_load(src.z + p), dst.z + d);
    __builtin_nontemporal_store(__builtin_nontemporal_load(src.ux + p), dst.ux + d);
    __builtin_nontemporal_store(__builtin_nontemporal_load(src.uy + p), dst.uy + d);
    __builtin_nontemporal_store(__builtin_nontemporal_load(src.uz + p), dst.uz + d);
    __builtin_nontemporal_store(__builtin_nontemporal_load(src.w + p), dst.w + d);
    __builtin_nontemporal_store(__builtin_nontemporal_load(src.qm + p), dst.qm + d);
    __builtin_nontemporal_store(__builtin_nontemporal_load(src.id + p), dst.id + d);
}

template <> struct KernelDef<Scatter64Args> {
    static constexpr int kBlock = 256;
    __device__ __forceinline__ static void body(const Scatter64Args &a, VBlock b, VBlock)
    {
        scatter64(a.src, a.dst, a.nbound, a.tv, a.parity, (long)b.x * 256 + threadIdx.x);
    }
};

int launch_scatter64(hipStream_t s, const Markers64 &src, const Markers64 &dst, long nbound,
                     const TileView &tv, int parity)
{
    if (nbound <= 0) return MPIC_OK;
    Scatter64Args a = {src, dst, nbound, tv, parity};
    return launch_def(s, a, (unsigned)((nbound + 255) / 256));
}

/* Per-cell marker counts in natural (i, j, k) order from the offsets. */
__global__ void k_counts_natural(TileView tv, Grid g, int32_t *out)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nc = (long)g.n[0] * g.n[1] * g.n[2];
    if (c >= nc) return;
    const int k = (int)(c % g.n[2]);
    const long r = c / g.n[2];
    const int j = (int)(r % g.n[1]), i = (int)(r / g.n[1]);
    const int key = cell_key_ijk(tv, g, i, j, k), t = key / TILE_CELLS;
    if (tv.toff[t + 1] == tv.toff[t]) {
        out[c] = 0;                       /* empty tile: its key offsets are not kept */
        return;
    }
    const int end = (key % TILE_CELLS == TILE_CELLS - 1) ? tv.toff[t + 1] : tv.off[key + 1];
    out[c] = end - tv.off[key];
}

int launch_counts_natural(hipStream_t s, const TileView &tv, const Grid &g, int32_t *out)
{
    const long nc = (long)g.n[0] * g.n[1] * g.n[2];
    k_counts_natural<<<(unsigned)((nc + 255) / 256), 256, 0, s>>>(tv, g, out);
    MPIC_CHECK(hipGetLastError());
    return MPIC_OK;
}
