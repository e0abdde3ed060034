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

/* The tile scan, in one pass: every workgroup takes SB consecutive tiles,
 * sums their per-key counts, scans (markers, tile non-empty, work items)
 * over its tiles, and finds the sum of all earlier tiles by a decoupled
 * look-back (Merrill and Garland, "Single-pass parallel prefix scan with
 * decoupled look-back", NVIDIA technical report NVR-2016-002): it publishes
 * its own totals at once, then adds up its predecessors' published values
 * back to the first one that has published its inclusive prefix, and
 * publishes its own inclusive prefix.  A workgroup only waits for
 * workgroups of lower index, which the hardware has dispatched before it,
 * so every wait ends; the wait is bounded all the same (SCAN_SPIN polls,
 * then an error flag, and the kernel still exits).
 *
 * Published values are 64-bit words (tag << 32 | value), three for a
 * workgroup's totals and three for its inclusive prefix; the tag is the
 * scan's sequence number (+1), so words of earlier scans never match and
 * the words are never cleared.  The last workgroup advances the sequence
 * number after every workgroup has published (its inclusive prefix needs
 * all of them), i.e. after every workgroup has read it.
 *
 * Layout of the counts: a tile's 64 per-key counts are 16 int4; 16 lanes of
 * a wavefront read one tile (coalesced), a wavefront four tiles per load,
 * and keep them in registers for the per-key offsets written at the end. */
#define NSCAN 3
#define SCAN_SPIN (1 << 20)
struct ScanArgs {
    TileView            tv;
    unsigned long long *flags;        /* [nb][6] published words */
    int                *seq;          /* scan sequence number */
    unsigned           *err;          /* bit 1: a look-back wait ran out */
    int                 nb, parity_next;
};

__device__ static inline unsigned long long scan_word(unsigned tag, int v)
{
    return ((unsigned long long)tag << 32) | (unsigned)v;
}

__device__ static inline void scan_pass(const ScanArgs &A, const VBlock &bk)
{
    const TileView &tv = A.tv;
    __shared__ int wsum[NSCAN][SB / 64];
    __shared__ int spre[NSCAN];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, blk = (int)bk.x;
    const int wt0 = blk * SB + w * 64;               /* first tile of the wavefront */
    const int4 *cnt4 = (const int4 *)tv.count;
    int4 x[16];
    int ex[16];                                      /* this lane's first key, within its tile */
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int t = wt0 + v * 4 + (lane >> 4);
        x[v] = t < tv.ntiles ? cnt4[(long)t * 16 + (lane & 15)] : make_int4(0, 0, 0, 0);
    }
    int c = 0;                                       /* markers of tile wt0 + lane */
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int s = x[v].x + x[v].y + x[v].z + x[v].w;
        int inc = s;
        for (int d = 1; d < 16; d <<= 1) {
            const int y = __shfl_up(inc, d, 16);
            if ((lane & 15) >= d) inc += y;
        }
        ex[v] = inc - s;
        const int tot = __shfl(inc, (lane & 3) * 16 + 15, 64);   /* tile v*4 + (lane & 3) */
        if ((lane >> 2) == v) c = tot;
    }
    const int own[NSCAN] = {c, c > 0, tile_items(c, tv.chm)};
    int inc[NSCAN];
    for (int a = 0; a < NSCAN; ++a) {                /* wavefront inclusive scan */
        int y = own[a];
        for (int d = 1; d < 64; d <<= 1) {
            const int z = __shfl_up(y, d, 64);
            if (lane >= d) y += z;
        }
        inc[a] = y;
        if (lane == 63) wsum[a][w] = y;
    }
    __syncthreads();
    int agg[NSCAN];
    for (int a = 0; a < NSCAN; ++a) {
        agg[a] = 0;
        for (int q = 0; q < SB / 64; ++q) {
            if (q < w) inc[a] += wsum[a][q];
            agg[a] += wsum[a][q];
        }
    }
    if (w == 0) {                                    /* the look-back, one wavefront */
        unsigned long long *F = A.flags;
        const int seq = *A.seq;
        const unsigned tag = (unsigned)(seq % 0x7fffffff) + 1u;
        int pre[NSCAN] = {0, 0, 0};
        if (blk > 0) {
            if (lane < NSCAN)
                __hip_atomic_store(F + 6 * blk + lane, scan_word(tag, agg[lane]), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            int end = blk - 1;                       /* the window's nearest predecessor */
            unsigned polls = 0;
            for (;;) {
                const int j = end - lane;
                bool ready = true, incl = false;
                int v[NSCAN] = {0, 0, 0};
                if (j >= 0) {
                    unsigned long long q[NSCAN];
                    bool ok = true;
                    for (int a = 0; a < NSCAN; ++a) {
                        q[a] = __hip_atomic_load(F + 6 * j + 3 + a, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
                        ok = ok && (unsigned)(q[a] >> 32) == tag;
                    }
                    incl = ok;
                    if (!ok) {
                        ok = true;
                        for (int a = 0; a < NSCAN; ++a) {
                            q[a] = __hip_atomic_load(F + 6 * j + a, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT);
                            ok = ok && (unsigned)(q[a] >> 32) == tag;
                        }
                    }
                    ready = ok;
                    for (int a = 0; a < NSCAN; ++a) v[a] = (int)(unsigned)q[a];
                }
                /* lanes up to the nearest inclusive prefix (or all) are needed */
                const unsigned long long im = __ballot(incl), nr = __ballot(!ready);
                const int L = im ? __ffsll((long long)im) - 1 : 63;
                const unsigned long long need = L == 63 ? ~0ull : ((2ull << L) - 1ull);
                if (nr & need) {
                    if (++polls > SCAN_SPIN) {       /* never expected: give up, flagged */
                        if (lane == 0 && A.err) atomicOr(A.err, 2u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                for (int a = 0; a < NSCAN; ++a) {
                    int y = (lane <= L && j >= 0) ? v[a] : 0;
                    for (int d = 32; d > 0; d >>= 1) y += __shfl_xor(y, d, 64);
                    pre[a] += y;
                }
                if (im) break;
                end -= 64;                           /* block 0 is always inclusive: end stays >= 0 */
            }
        }
        if (lane < NSCAN) {
            __hip_atomic_store(F + 6 * blk + 3 + lane, scan_word(tag, pre[lane] + agg[lane]),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            spre[lane] = pre[lane];
        }
        if (blk == A.nb - 1 && lane == 0) {          /* totals, then the next sequence number */
            const int tot = pre[0] + agg[0], na = pre[1] + agg[1], ni = pre[2] + agg[2];
            tv.off[tv.ncellp] = tot;
            tv.toff[tv.ntiles] = tot;
            *tv.nactive = na;
            tv.istart[na] = ni;
            *tv.nitems = ni;
            tv.nmark[A.parity_next] = tot;
            for (int q = 0; q < MPIC_XCDS; ++q) tv.queue[q] = 0;
            *A.seq = seq + 1;
        }
    }
    __syncthreads();
    /* tile offsets, the active list, the work items */
    const int base = spre[0] + inc[0] - own[0];
    const int t = wt0 + lane;
    if (t < tv.ntiles) {
        tv.toff[t] = base;
        const int a = spre[1] + inc[1] - own[1];
        if (c > 0) {
            const int is = spre[2] + inc[2] - own[2];
            tv.active[a] = t;
            tv.istart[a] = is;
            for (int u = 0; u < own[2]; ++u) tv.item[is + u] = a;
        }
        tv.slot[t] = c > 0 ? a : -1;
    }
    /* per-key offsets and cursors of the non-empty tiles, counters reset
     * (entries of empty tiles are not maintained) */
    int4 *off4 = (int4 *)tv.off, *cur4 = (int4 *)tv.cursor, *cnt4w = (int4 *)tv.count;
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int tt = v * 4 + (lane >> 4);
        const int tb = __shfl(base, tt, 64), tc = __shfl(c, tt, 64);
        if (tc > 0) {
            const long q = (long)(wt0 + tt) * 16 + (lane & 15);
            const int o0 = tb + ex[v];
            const int4 o = make_int4(o0, o0 + x[v].x, o0 + x[v].x + x[v].y,
                                     o0 + x[v].x + x[v].y + x[v].z);
            off4[q] = o;
            cur4[q] = o;
            cnt4w[q] = make_int4(0, 0, 0, 0);
        }
    }
}

template <> struct KernelDef<ScanArgs> {
    static constexpr int kBlock = SB;
    __device__ __forceinline__ static void body(const ScanArgs &a, VBlock b, VBlock)
    {
        scan_pass(a, b);
    }
};

long tile_scan_words(int ntiles)
{
    return 6L * ((ntiles + SB - 1) / SB) + 1;
}

int launch_tile_scan(hipStream_t s, const TileView &tv, unsigned long long *words, unsigned *err,
                     int parity_next)
{
    const int nb = (tv.ntiles + SB - 1) / SB;
    ScanArgs a = {tv, words, (int *)(words + 6L * nb), err, nb, parity_next};
    return launch_def(s, a, (unsigned)nb);
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
