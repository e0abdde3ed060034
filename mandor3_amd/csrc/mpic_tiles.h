/*
 * mpic_tiles.h -- tile-major cell keys, the per-step counting sort and the
 * per-XCD tile queues shared by the FP64 and FP32 push kernels.
 *
 * The grid of cells is cut into 4x4x4-cell tiles; a marker's key is
 *   key = tile * 64 + (li * 4 + lj) * 4 + lk
 * so that sorting by key groups markers by tile and, inside a tile, by
 * cell.  Offsets are int32 (a run holds fewer than 2^31 markers).
 */
#ifndef MPIC_TILES_H
#define MPIC_TILES_H

#include "mpic_internal.h"

#define TILE_CELLS 64
#define MPIC_SCAN_BLOCK 256     /* tiles per workgroup of the tile scan */

struct TileView {
    int  nt[3];              /* tiles per axis */
    int  ntiles;
    long ncellp;             /* ntiles * 64 */
    int *off;                /* [ncellp + 1] first marker of every cell key of a non-empty
                              * tile (entries of empty tiles are not maintained) */
    int *toff;               /* [ntiles + 1] first marker of every tile */
    int *cursor;             /* [ncellp] scatter cursors */
    int *count;              /* [ncellp] markers per key after the push (zeroed by the scan) */
    int *key;                /* [capacity] new key of every marker, -1 = deleted */
    int *perm;               /* [capacity] the current order as indices into the
                              * current buffer (nmark[parity] of them; tiled runs) */
    int *active;             /* [ntiles] tiles holding markers, in tile order */
    int *slot;               /* [ntiles] position of a tile in 'active', -1 if empty */
    int *nactive;            /* [1] */
    int *queue;              /* [MPIC_XCDS] work counters */
    int *nmark;              /* [2] markers in the current / next order */
    /* Work items of the FP32 push: an active tile's markers split evenly
     * into ceil(count / chm) items of at most chm markers (chm = 0: one
     * item per active tile). */
    int  chm;
    int *item;               /* [item capacity] active slot of each work item */
    int *istart;             /* [ntiles + 1] first item of each active slot, [nactive] = items */
    int *nitems;             /* [1] */
};

/* Work items of a tile holding c markers. */
__host__ __device__ static inline int tile_items(int c, int chm)
{
    return c <= 0 ? 0 : (chm > 0 ? (c + chm - 1) / chm : 1);
}

__device__ static inline void tile_coords(const TileView &tv, int t, int &t0, int &t1, int &t2)
{
    t2 = t % tv.nt[2];
    const int r = t / tv.nt[2];
    t1 = r % tv.nt[1];
    t0 = r / tv.nt[1];
}

__device__ static inline int cell_key_ijk(const TileView &tv, const Grid &g, int i, int j, int k)
{
    i = i < 0 ? 0 : (i >= g.n[0] ? g.n[0] - 1 : i);
    j = j < 0 ? 0 : (j >= g.n[1] ? g.n[1] - 1 : j);
    k = k < 0 ? 0 : (k >= g.n[2] ? g.n[2] - 1 : k);
    const int t = ((i >> 2) * tv.nt[1] + (j >> 2)) * tv.nt[2] + (k >> 2);
    return t * TILE_CELLS + (((i & 3) * 4 + (j & 3)) * 4 + (k & 3));
}

/* Key of a position: floor(x/h) as the reference's MF_DBL_TO_INT. */
__device__ static inline int cell_key_of(const TileView &tv, const Grid &g, const double r[3])
{
    const int i = (int)(r[0] / g.h[0] + 128.0) - 128;
    const int j = (int)(r[1] / g.h[1] + 128.0) - 128;
    const int k = (int)(r[2] / g.h[2] + 128.0) - 128;
    return cell_key_ijk(tv, g, i, j, k);
}

/* Next active tile for a workgroup of XCD 'xcd'; the active list is split
 * into one contiguous range per XCD (neighbouring tiles share an L2), and
 * a workgroup whose range is drained steals from the next XCDs.  Returns -1
 * when every range is drained, so every workgroup exits. */
__device__ static inline int tile_next(const TileView &tv, int xcd)
{
    const int na = *tv.nactive;
    for (int d = 0; d < MPIC_XCDS; ++d) {
        const int x = (xcd + d) % MPIC_XCDS;
        const int lo = (int)((long)na * x / MPIC_XCDS), hi = (int)((long)na * (x + 1) / MPIC_XCDS);
        if (lo >= hi) continue;
        if (lo + __hip_atomic_load(tv.queue + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= hi)
            continue;                        /* drained: no contended atomic on the way out */
        const int q = atomicAdd(tv.queue + x, 1);
        if (lo + q < hi) return tv.active[lo + q];
    }
    return -1;
}

/* Wave-level run detection on a wave64: lanes holding equal keys in a row
 * form a run; returns (start lane, length) of the caller's run. */
__device__ static inline void wave_run(int key, int &start, int &len, bool &head)
{
    const int lane = threadIdx.x & 63;
    const int prev = __shfl_up(key, 1, 64);
    head = (lane == 0) || (key != prev);
    const unsigned long long hm = __ballot(head);
    const unsigned long long le = (lane == 63) ? ~0ull : ((1ull << (lane + 1)) - 1ull);
    start = 63 - __clzll(hm & le);
    const unsigned long long gt = hm & ~le;
    const int end = gt ? (__ffsll((long long)gt) - 1) : 64;
    len = end - start;
}

/* count[key] += run length, one atomic per run (all lanes must call). */
__device__ static inline void wave_run_count(int *count, int key)
{
    int start, len;
    bool head;
    wave_run(key, start, len, head);
    if (head && key >= 0) atomicAdd(count + key, len);
}

int launch_keys64(hipStream_t s, const Markers64 &m, long n, const Grid &g, const TileView &tv,
                  int parity);
/* The tile scan's word buffer (zeroed once at allocation) and launcher;
 * err bit 1 is set if a wait of the scan ran out (never expected). */
long tile_scan_words(int ntiles);
int launch_tile_scan(hipStream_t s, const TileView &tv, unsigned long long *words, unsigned *err,
                     int parity_next);
int launch_rank(hipStream_t s, long nbound, const TileView &tv, int parity);
int launch_scatter64(hipStream_t s, const Markers64 &src, const Markers64 &dst, long nbound,
                     const TileView &tv, int parity);
int launch_push64_tiled(hipStream_t s, const Markers64 &m, const Markers64 &dst,
                        const Fields64 &f, const Grid &g, const PushConsts &c,
                        const TileView &tv, double *Wacc, unsigned *err, int nblocks);
int push64_blocks_per_cu(void);
int launch_counts_natural(hipStream_t s, const TileView &tv, const Grid &g, int32_t *out);

/* J block of one tile: nodes [4t-1, 4t+5] per axis, xyz interleaved. */
#define TJN 7
#define TJB (TJN * TJN * TJN * 3)

/* nblocks workgroups of 'waves' (1, 2 or 4) wavefronts each. */
int launch_push32(hipStream_t s, const Markers32 &m, const Markers32 &dst, const Fields32 &f,
                  const Grid &g, const PushConsts &pc, const float *qm, const TileView &tv,
                  float *tj, double *Wacc, int nblocks, int waves);
int launch_jgather32(hipStream_t s, const Fields32 &f, const Grid &g, const TileView &tv,
                     const float *tj, int nblocks);
int launch_keys32(hipStream_t s, const Markers32 &m, long n, const TileView &tv, int parity);
int launch_rho32(hipStream_t s, const Markers32 &m, long n, double *rho, const Grid &g,
                 const TileView &tv, int parity);
int launch_cells32(hipStream_t s, const Markers32 &m, long n, const Grid &g, const TileView &tv,
                   int parity, int32_t *cell);

#endif
