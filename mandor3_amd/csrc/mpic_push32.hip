/*
 * mpic_push32.hip -- FP32 split-precision marker push and current deposit.
 *
 * Marker layout, two 16-byte records per marker (32 B read + 32 B written
 * per push, each record one 128-bit access) and the id:
 *   pos   cell word (int32 bits: (species << 29) | tile-major cell key,
 *         -1 deleted), ox oy oz FP32 offset inside the cell, cell units
 *   mom   ux uy uz FP32 momentum gamma*v/c, w FP32 charge density of the
 *         marker (the reference's rho)
 *   id    int32 reference marker id (carried along by the push)
 * Position = (cell index + offset) * h: the integer part carries the
 * position exactly, so FP32 resolves 6e-8 of a cell anywhere in the box.
 *
 * Per marker: trilinear gather of FP32 E/H (source/misc_PIC.h:69-126),
 * Boris rotation in FP32 (source/core/plasma.c:132-171), straight move in
 * cell units, charge-conserving deposit of the trajectory split at cell
 * faces (the Villasenor-Buneman form of source/core/plasma_currentKernel.c:
 * 20-124 in cell units), walls (source/core/plasma_walls.c:27-104).
 *
 * Deposit: markers are sorted by cell, so the lanes of a wavefront mostly
 * share one start cell.  The segment inside the start cell (every marker
 * has one) is summed across each run of equal cells inside each 16-lane
 * row with a segmented DPP scan (VALU only), then added to the tile's FP32 J
 * block in LDS once per run and row; later segments (cell crossings) go to
 * LDS with atomics.  The block is
 * flushed to the FP64 J in HBM once per tile.  The tile's E/H nodes are
 * staged in LDS, so gathers never leave the CU.
 *
 * Order: the markers of a step are taken in tile-major cell order through a
 * permutation (indices into the current buffer) that the sort's rank pass
 * builds from the push's new keys; the push writes each pushed marker to
 * the other buffer at its position in that order.  So no pass moves the
 * markers only to sort them: the next step reads the other buffer through
 * the next permutation.
 *
 * Roofline: HBM-bound.  Algorithmic bytes per marker-step 32 B read + 32 B
 * written, plus per occupied cell 24 B of FP32 E/H reads and 48 B of FP64 J
 * read-modify-write (the id, the permutation index and the new key are
 * 4 B each on top).
 */
#include "mpic_internal.h"
#include "mpic_launch.h"
#include "mpic_tiles.h"

#define TPB 256
#define TW 64                  /* wavefront width */
#define TILE 4
#define TFN (TILE + 2)         /* E/H nodes per axis staged for a tile's gathers */
#define TFN3 (TFN * TFN * TFN)
#define SPECIES_SHIFT 29
#define KEY_MASK ((1 << SPECIES_SHIFT) - 1)

struct F32View {
    const float *E[3], *H[3];
    double *J[3];
    Grid g;
};

/* 12 J values of one straight segment inside a cell, cell units:
 * start (x1,y1,z1), end (x2,y2,z2) relative to the cell's origin. */
struct Seg12 {
    float v[12];
};

__device__ static inline void seg_values(float x1, float y1, float z1, float x2, float y2,
                                         float z2, float wx, float wy, float wz, Seg12 &s)
{
    const float dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
    const float tx = 0.5f * (x1 + x2), ty = 0.5f * (y1 + y2), tz = 0.5f * (z1 + z2);
    const float yz = dy * dz * (1.0f / 12), xz = dx * dz * (1.0f / 12), xy = dx * dy * (1.0f / 12);
    const float ax = wx * dx, ay = wy * dy, az = wz * dz;
    s.v[0] = ax * ((1 - ty) * (1 - tz) + yz);
    s.v[1] = ax * (ty * (1 - tz) - yz);
    s.v[2] = ax * ((1 - ty) * tz - yz);
    s.v[3] = ax * (ty * tz + yz);
    s.v[4] = ay * ((1 - tx) * (1 - tz) + xz);
    s.v[5] = ay * (tx * (1 - tz) - xz);
    s.v[6] = ay * ((1 - tx) * tz - xz);
    s.v[7] = ay * (tx * tz + xz);
    s.v[8] = az * ((1 - tx) * (1 - ty) + xy);
    s.v[9] = az * (tx * (1 - ty) - xy);
    s.v[10] = az * ((1 - tx) * ty - xy);
    s.v[11] = az * (tx * ty + xy);
}

/* Offset of node (+da,+db,+dc), component comp, from a node of the LDS J block. */
#define LJO(da, db, dc, comp) (3 * (((da) * TJN + (db)) * TJN + (dc)) + (comp))

/* Adds a segment of local cell (a,b,c) (LDS block coordinates) to LDS:
 * one address, the twelve nodes as constant offsets of it. */
__device__ static inline void seg_to_lds(float *lj, int a, int b, int c, const Seg12 &s)
{
    float *p = lj + __mul24(__mul24(__mul24(a, TJN) + b, TJN) + c, 3);
    atomicAdd(p + LJO(1, 0, 0, 0), s.v[0]);
    atomicAdd(p + LJO(1, 1, 0, 0), s.v[1]);
    atomicAdd(p + LJO(1, 0, 1, 0), s.v[2]);
    atomicAdd(p + LJO(1, 1, 1, 0), s.v[3]);
    atomicAdd(p + LJO(0, 1, 0, 1), s.v[4]);
    atomicAdd(p + LJO(1, 1, 0, 1), s.v[5]);
    atomicAdd(p + LJO(0, 1, 1, 1), s.v[6]);
    atomicAdd(p + LJO(1, 1, 1, 1), s.v[7]);
    atomicAdd(p + LJO(0, 0, 1, 2), s.v[8]);
    atomicAdd(p + LJO(1, 0, 1, 2), s.v[9]);
    atomicAdd(p + LJO(0, 1, 1, 2), s.v[10]);
    atomicAdd(p + LJO(1, 1, 1, 2), s.v[11]);
}

/* Same, straight to the FP64 J in HBM (segment outside the LDS block). */
__device__ static inline void seg_to_hbm(const F32View &v, int i, int j, int k, const Seg12 &s)
{
    const Grid &g = v.g;
    unsafeAtomicAdd(v.J[0] + gidx(g, i + 1, j, k), (double)s.v[0]);
    unsafeAtomicAdd(v.J[0] + gidx(g, i + 1, j + 1, k), (double)s.v[1]);
    unsafeAtomicAdd(v.J[0] + gidx(g, i + 1, j, k + 1), (double)s.v[2]);
    unsafeAtomicAdd(v.J[0] + gidx(g, i + 1, j + 1, k + 1), (double)s.v[3]);
    unsafeAtomicAdd(v.J[1] + gidx(g, i, j + 1, k), (double)s.v[4]);
    unsafeAtomicAdd(v.J[1] + gidx(g, i + 1, j + 1, k), (double)s.v[5]);
    unsafeAtomicAdd(v.J[1] + gidx(g, i, j + 1, k + 1), (double)s.v[6]);
    unsafeAtomicAdd(v.J[1] + gidx(g, i + 1, j + 1, k + 1), (double)s.v[7]);
    unsafeAtomicAdd(v.J[2] + gidx(g, i, j, k + 1), (double)s.v[8]);
    unsafeAtomicAdd(v.J[2] + gidx(g, i + 1, j, k + 1), (double)s.v[9]);
    unsafeAtomicAdd(v.J[2] + gidx(g, i, j + 1, k + 1), (double)s.v[10]);
    unsafeAtomicAdd(v.J[2] + gidx(g, i + 1, j + 1, k + 1), (double)s.v[11]);
}

__device__ static inline void seg_add(const F32View &v, float *lj, int o0, int o1, int o2, int i,
                                      int j, int k, const Seg12 &s)
{
    const int a = i - o0, b = j - o1, c = k - o2;
    if (a < 0 || b < 0 || c < 0 || a > TJN - 2 || b > TJN - 2 || c > TJN - 2)
        seg_to_hbm(v, i, j, k, s);
    else
        seg_to_lds(lj, a, b, c, s);
}

struct Push32Consts {
    float tau, dtn[3];       /* tau, tau/h */
    float cj[3];             /* h/tau: J of a segment per unit w and displacement */
    double vol;              /* h1*h2*h3 (energy) */
    float qm[4];             /* q/m per species (species is 2 bits of the cell word) */
    int bc[6];
};

/* Wall of axis a on cell index c / offset o (cell units); returns 0 when
 * an open wall deletes the marker.  Reference order: mins then maxes. */
__device__ static inline int wall_lo(int bc, int &c, float &o, float &u, int N)
{
    if (bc == MPIC_BC_PERIODIC) {
        if (c < 0) c += N;                              /* x <= 0: x += L */
    } else if (bc == MPIC_BC_MIRROR) {
        if (c < 0 || (c == 0 && o == 0.0f)) {           /* x <= 0: x = -x, u = -u */
            const float x = -((float)(c) + o);          /* in (0, 1] */
            if (x >= 1.0f) { c = 1; o = 0.0f; } else { c = 0; o = x; }
            u = -u;
        }
    } else if (bc == MPIC_BC_OPEN) {
        if (c < 1 || (c == 1 && o == 0.0f)) return 0;   /* x <= h: deleted */
    }
    return 1;
}

__device__ static inline int wall_hi(int bc, int &c, float &o, float &u, int N)
{
    if (bc == MPIC_BC_PERIODIC) {
        if (c >= N) c -= N;                             /* x >= L: x -= L */
    } else if (bc == MPIC_BC_MIRROR) {
        if (c >= N) {                                   /* x >= L: x = 2L - x */
            /* reflected position relative to cell N-1: 1 - o - (c - N) */
            float on = 1.0f - o - (float)(c - N);
            if (on >= 1.0f) on = 0.99999994f;           /* x == L stays at the wall */
            if (on < 0.0f) on = 0.0f;
            c = N - 1;
            o = on;
            u = -u;
        }
    } else if (bc == MPIC_BC_OPEN) {
        if (c >= N - 1) return 0;                       /* x >= L - h: deleted */
    }
    return 1;
}

/* Point at fraction a of the move (ox,oy,oz) -> (x2,y2,z2), placed exactly
 * on the crossed face of axis ax. */
__device__ static inline void crossing_point(float ox, float oy, float oz, float x2, float y2,
                                             float z2, int fx, int fy, int fz, float a, int ax,
                                             float &qx, float &qy, float &qz)
{
    qx = (ax == 0) ? (fx > 0 ? 1.0f : 0.0f) : ox + a * (x2 - ox);
    qy = (ax == 1) ? (fy > 0 ? 1.0f : 0.0f) : oy + a * (y2 - oy);
    qz = (ax == 2) ? (fz > 0 ? 1.0f : 0.0f) : oz + a * (z2 - oz);
}

__device__ static inline void cswap(float &a, int &ia, float &b, int &ib)
{
    if (b < a) {
        const float t = a; a = b; b = t;
        const int u = ia; ia = ib; ib = u;
    }
}

/* Hardware reciprocal and reciprocal square root (1 ulp): the split form's
 * error budget is set by FP32 storage, not by these. */
__device__ static inline float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ static inline float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

/* Marker state held in registers between the prefetch and the push. */
struct MarkerRegs {
    int   word;
    float ox, oy, oz, ux, uy, uz, w;
    int   id;
};

/* Element p of a 4-byte array as base + 32-bit byte offset, so the access
 * uses a scalar base and a 32-bit vector offset (used for the field staging,
 * whose offsets from the tile's first node span six planes). */
template <class T>
__device__ static inline T &at(T *base, unsigned p)
{
    return *(T *)((char *)base + (p << 2));
}

/* Index of position min(p, last) of the order: unconditional, so that every
 * path through the push loop issues the same memory operations and the
 * compiler's wait counts stay exact (a skippable load makes it wait for
 * later loads too).  Lanes past the item's end carry a duplicate, masked by
 * the caller. */
__device__ static inline unsigned perm_at(const int *perm, unsigned p, unsigned last)
{
    return (unsigned)perm[p < last ? p : last];
}

/* The records of marker s of the current buffer.  Marker records are read
 * once and written once per step: nontemporal, so that they do not push the
 * field arrays (which every kernel of the step reads again) out of the
 * caches. */
typedef float v4f_nt __attribute__((ext_vector_type(4)));
__device__ static inline void load_marker(const Markers32 &m, unsigned s, MarkerRegs &r)
{
    const v4f_nt a = __builtin_nontemporal_load((const v4f_nt *)(m.pos + s));
    const v4f_nt b = __builtin_nontemporal_load((const v4f_nt *)(m.mom + s));
    r.word = __float_as_int(a.x);
    r.ox = a.y;
    r.oy = a.z;
    r.oz = a.w;
    r.ux = b.x;
    r.uy = b.y;
    r.uz = b.z;
    r.w = b.w;
    r.id = __builtin_nontemporal_load(m.id + s);
}

/* Trilinear gather from a tile's staged field block (TFN^3 nodes, node
 * (a,b,c) relative to the tile's first cell). */
__device__ static inline float lgather(const float *F, int a, float sx, int b, float sy, int c,
                                       float sz)
{
    const float sx1 = 1.0f - sx, sy1 = 1.0f - sy, sz1 = 1.0f - sz;
    const int o = (a * TFN + b) * TFN + c;
    const int dj = TFN, di = TFN * TFN;
    return sz1 * (sy1 * (sx1 * F[o] + sx * F[o + di]) + sy * (sx1 * F[o + dj] + sx * F[o + di + dj])) +
           sz * (sy1 * (sx1 * F[o + 1] + sx * F[o + di + 1]) +
                 sy * (sx1 * F[o + dj + 1] + sx * F[o + di + dj + 1]));
}

/* The tile being pushed: its index, its neighbour strides along i and j
 * (in tiles) and whether a wall lies within reach of its markers (a tile
 * one tile or more away from every face is out of reach: markers move less
 * than a cell per step). */
struct TilePos {
    int  t, s0, s1;
    bool edge;
};

/* Push of one marker of tile-local cell (li,lj,lk) (cell (ci,cj,ck)):
 * gather from the staged fields, Boris, move, deposit of every segment but
 * the first into the tile's LDS J (or HBM beyond it), walls.  Returns the
 * new cell word (-1: deleted), the new offsets and momenta in 'out' and the
 * first segment, which lies inside the start cell. */
__device__ static inline int push32_marker(MarkerRegs &out, const MarkerRegs &r, int li,
                                           int lj_, int lk, int ci, int cj, int ck,
                                           const float *lf, const float *lqm, float *lj, int o0,
                                           int o1, int o2,
                                           const F32View &v, const Push32Consts &c,
                                           const TileView &tv, const TilePos &tp,
                                           Seg12 &first, float &w0,
                                           float &w1, float &w2)
{
    const Grid &g = v.g;
    const int sp = (int)((unsigned)r.word >> SPECIES_SHIFT);
    const float ox = r.ox, oy = r.oy, oz = r.oz;
    float ux = r.ux, uy = r.uy, uz = r.uz;
    /* q/m from the workgroup's LDS copy: an indexed read of the argument
     * block would be a vector memory load, and its wait would also wait for
     * the prefetched markers */
    const float qm = lqm[sp];
    const float w = r.w;
    /* gather: E_x at (i-1/2, j, k), E_y at (i, j-1/2, k), ... */
    const int hx = ox >= 0.5f, hy = oy >= 0.5f, hz = oz >= 0.5f;
    const float sxh = ox + 0.5f - (float)hx, syh = oy + 0.5f - (float)hy, szh = oz + 0.5f - (float)hz;
    const int ah = li + hx, bh = lj_ + hy, ch = lk + hz;
    const float Ex = lgather(lf + 0 * TFN3, ah, sxh, lj_, oy, lk, oz),
                Ey = lgather(lf + 1 * TFN3, li, ox, bh, syh, lk, oz),
                Ez = lgather(lf + 2 * TFN3, li, ox, lj_, oy, ch, szh),
                Hx = lgather(lf + 3 * TFN3, li, ox, bh, syh, ch, szh),
                Hy = lgather(lf + 4 * TFN3, ah, sxh, lj_, oy, ch, szh),
                Hz = lgather(lf + 5 * TFN3, ah, sxh, bh, syh, lk, oz);
    const float tm = 0.5f * c.tau * qm;
    const float px = ux + tm * Ex, py = uy + tm * Ey, pz = uz + tm * Ez;
    const float p2 = px * px + py * py + pz * pz;
    const float ig0 = rsq(1.0f + p2);
    const float gam = (1.0f + p2) * ig0;
    /* (gamma - 1)/p^2 = 1/(gamma + 1): no cancellation for slow markers */
    const float W = w * rcp((gam + 1.0f) * qm);
    w0 += W * px * px;
    w1 += W * py * py;
    w2 += W * pz * pz;
    const float tg = tm * ig0;
    const float ax = tg * Hx, ay = tg * Hy, az = tg * Hz;
    const float uu = px + py * az - pz * ay, vv = py + pz * ax - px * az, ww = pz + px * ay - py * ax;
    const float bxx = ax * ax, byy = ay * ay, bzz = az * az, bxy = ax * ay, byz = ay * az,
                bxz = az * ax;
    const float del = rcp(1.0f + bxx + byy + bzz);
    ux = tm * Ex + ((1 + bxx) * uu + (bxy + az) * vv + (bxz - ay) * ww) * del;
    uy = tm * Ey + ((bxy - az) * uu + (1 + byy) * vv + (byz + ax) * ww) * del;
    uz = tm * Ez + ((bxz + ay) * uu + (byz - ax) * vv + (1 + bzz) * ww) * del;
    const float ig = rsq(1.0f + ux * ux + uy * uy + uz * uz);
    /* move in cell units from (ox,oy,oz) of cell (ci,cj,ck) */
    const float x2 = ox + ux * c.dtn[0] * ig, y2 = oy + uy * c.dtn[1] * ig, z2 = oz + uz * c.dtn[2] * ig;
    const int fx = (int)floorf(x2), fy = (int)floorf(y2), fz = (int)floorf(z2);
    const float wx = w * c.cj[0], wy = w * c.cj[1], wz = w * c.cj[2];
    /* Segments: from the start to the first face crossing (or the end),
     * then one per crossed face.  Crossing fractions ascend; an axis that is
     * not crossed gets 2 and sorts last.  Each axis is crossed at most once
     * per step, so a crossing point lies on face 0 or 1 of the start cell's
     * frame on that axis. */
    const int nc = (fx != 0) + (fy != 0) + (fz != 0);
    float a0 = 2.0f, a1 = 2.0f, a2 = 2.0f;
    int x0 = 0, x1 = 1, x2i = 2;
    float qx = x2, qy = y2, qz = z2;
    if (nc) {
        if (fx) a0 = ((fx > 0 ? 1.0f : 0.0f) - ox) * rcp(x2 - ox);
        if (fy) a1 = ((fy > 0 ? 1.0f : 0.0f) - oy) * rcp(y2 - oy);
        if (fz) a2 = ((fz > 0 ? 1.0f : 0.0f) - oz) * rcp(z2 - oz);
        cswap(a0, x0, a1, x1);
        cswap(a1, x1, a2, x2i);
        cswap(a0, x0, a1, x1);
        crossing_point(ox, oy, oz, x2, y2, z2, fx, fy, fz, a0, x0, qx, qy, qz);
    }
    seg_values(ox, oy, oz, qx, qy, qz, wx, wy, wz, first);
    if (nc) {
        int sx = 0, sy = 0, sz = 0, axs = x0;
#pragma unroll 1
        for (int s = 1; s <= nc; ++s) {
            sx += (axs == 0) ? fx : 0;
            sy += (axs == 1) ? fy : 0;
            sz += (axs == 2) ? fz : 0;
            a0 = a1;
            x0 = x1;
            a1 = a2;
            x1 = x2i;
            float ex = x2, ey = y2, ez = z2;
            if (s < nc) crossing_point(ox, oy, oz, x2, y2, z2, fx, fy, fz, a0, x0, ex, ey, ez);
            Seg12 sg;
            seg_values(qx - sx, qy - sy, qz - sz, ex - sx, ey - sy, ez - sz, wx, wy, wz, sg);
            seg_add(v, lj, o0, o1, o2, ci + sx, cj + sy, ck + sz, sg);
            qx = ex;
            qy = ey;
            qz = ez;
            axs = x0;
        }
    }
    /* new cell and offset, then the walls */
    int ni = ci + fx, nj = cj + fy, nk = ck + fz;
    float nx = x2 - (float)fx, ny = y2 - (float)fy, nz = z2 - (float)fz;
    if (nx >= 1.0f) { nx = 0.0f; ++ni; }
    if (ny >= 1.0f) { ny = 0.0f; ++nj; }
    if (nz >= 1.0f) { nz = 0.0f; ++nk; }
    int key;
    if (tp.edge) {
        const int alive =
            wall_lo(c.bc[0], ni, nx, ux, g.n[0]) && wall_lo(c.bc[1], nj, ny, uy, g.n[1]) &&
            wall_lo(c.bc[2], nk, nz, uz, g.n[2]) && wall_hi(c.bc[3], ni, nx, ux, g.n[0]) &&
            wall_hi(c.bc[4], nj, ny, uy, g.n[1]) && wall_hi(c.bc[5], nk, nz, uz, g.n[2]);
        if (!alive) return -1;
        key = cell_key_ijk(tv, g, ni, nj, nk);
    } else {
        /* no wall within reach: the new cell is within one cell of the
         * tile, so its key follows from the tile's neighbour strides */
        const int ui = li + ni - ci, uj = lj_ + nj - cj, uk = lk + nk - ck;
        const int nt = tp.t + __mul24(ui >> 2, tp.s0) + __mul24(uj >> 2, tp.s1) + (uk >> 2);
        key = nt * TILE_CELLS + (((ui & 3) * TILE + (uj & 3)) * TILE + (uk & 3));
    }
    out.ox = nx;
    out.oy = ny;
    out.oz = nz;
    out.ux = ux;
    out.uy = uy;
    out.uz = uz;
    out.w = w;
    return (sp << SPECIES_SHIFT) | key;
}

/* One step of the row scan for all 12 values: v[q] += v[q] of lane - SH
 * times the lane's 0/1 mask M (v_fmac_f32_dpp; a source outside the
 * 16-lane row reads 0).  x + y*1 and x + y*0 round as x + y and x. */
#define SCAN_STEP(SH, M)                                                                    \
    "v_fmac_f32_dpp %0, %0, %" #M " row_shr:" #SH " row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_fmac_f32_dpp %1, %1, %" #M " row_shr:" #SH " row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_fmac_f32_dpp %2, %2, %" #M " row_shr:" #SH " row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_fmac_f32_dpp %3, %3, %" #M " row_shr:" #SH " row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_fmac_f32_dpp %4, %4, %" #M " row_shr:" #SH " row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_fmac_f32_dpp %5, %5, %" #M " row_shr:" #SH " row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_fmac_f32_dpp %6, %6, %" #M " row_shr:" #SH " row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_fmac_f32_dpp %7, %7, %" #M " row_shr:" #SH " row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_fmac_f32_dpp %8, %8, %" #M " row_shr:" #SH " row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_fmac_f32_dpp %9, %9, %" #M " row_shr:" #SH " row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_fmac_f32_dpp %10, %10, %" #M " row_shr:" #SH " row_mask:0xf bank_mask:0xf bound_ctrl:1\n" \
    "v_fmac_f32_dpp %11, %11, %" #M " row_shr:" #SH " row_mask:0xf bank_mask:0xf bound_ctrl:1\n"

/* Segmented inclusive prefix sum of the 12 first-segment values over runs
 * of equal start cells inside each 16-lane row; 'start' is the first lane of
 * the caller's run.  Afterwards the last lane of each run in a row holds
 * that row's part of the run total.  A step adds only when its source lane
 * belongs to the caller's run, by a 0/1 factor, so each step is one VALU
 * instruction per value (a compiled conditional add is a DPP add plus a
 * select).  All lanes of the wavefront execute it.  A DPP read of a VGPR
 * needs 2 wait states after the VALU write of that VGPR: s_nop 1 covers the
 * writes before the block, and inside it the 12 independent chains put 11
 * instructions between a value's write and its next DPP read. */
__device__ static inline void row_scan(Seg12 &f, int start)
{
    const int lane = threadIdx.x & 63;
    const float m1 = lane - 1 >= start ? 1.0f : 0.0f, m2 = lane - 2 >= start ? 1.0f : 0.0f,
                m4 = lane - 4 >= start ? 1.0f : 0.0f, m8 = lane - 8 >= start ? 1.0f : 0.0f;
    asm volatile("s_nop 1\n" SCAN_STEP(1, 12) SCAN_STEP(2, 13) SCAN_STEP(4, 14) SCAN_STEP(8, 15)
                 : "+v"(f.v[0]), "+v"(f.v[1]), "+v"(f.v[2]), "+v"(f.v[3]), "+v"(f.v[4]),
                   "+v"(f.v[5]), "+v"(f.v[6]), "+v"(f.v[7]), "+v"(f.v[8]), "+v"(f.v[9]),
                   "+v"(f.v[10]), "+v"(f.v[11])
                 : "v"(m1), "v"(m2), "v"(m4), "v"(m8));
}
#undef SCAN_STEP

__device__ static inline void key_to_ijk(const TileView &tv, int key, int &i, int &j, int &k)
{
    int t0, t1, t2;
    tile_coords(tv, key >> 6, t0, t1, t2);
    const int loc = key & 63;
    i = TILE * t0 + (loc >> 4);
    j = TILE * t1 + ((loc >> 2) & 3);
    k = TILE * t2 + (loc & 3);
}

/* Results of one batch, written back at the start of the next batch (after
 * its prefetch is issued) so that no wait for the prefetch also waits for
 * fresh stores. */
struct Pending {
    MarkerRegs r;
    unsigned   p;
    int        word;      /* new cell word, -1 deleted */
    bool       valid;     /* lane pushed a marker */
    int        ckey, clen;/* count run head: key (-1 none), length */
};

/* Position p of the order in the output buffer d: the records, the id and
 * the new key (-1: deleted, which the rank pass leaves out). */
__device__ static inline void write_back(const Markers32 &d, const TileView &tv, const Pending &q)
{
    if (q.valid) {
        const v4f_nt a = {__int_as_float(q.word), q.r.ox, q.r.oy, q.r.oz};
        __builtin_nontemporal_store(a, (v4f_nt *)(d.pos + q.p));
        if (q.word >= 0) {
            const v4f_nt b = {q.r.ux, q.r.uy, q.r.uz, q.r.w};
            __builtin_nontemporal_store(b, (v4f_nt *)(d.mom + q.p));
        }
        __builtin_nontemporal_store(q.r.id, d.id + q.p);
        tv.key[q.p] = q.word < 0 ? -1 : (q.word & KEY_MASK);   /* read by the rank pass next */
    }
    if (q.ckey >= 0) atomicAdd(tv.count + q.ckey, q.clen);
}

/* Tiled push: one workgroup of W wavefronts per work item (an even share
 * of one active tile's (4x4x4 cells) markers, or the whole tile); items are
 * taken with stride = grid size (a shared queue would serialise on its
 * counter).  The tile's E/H nodes are staged in LDS once for the W
 * wavefronts, which push interleaved batches of 64 markers each and add to
 * one FP32 LDS J block; the block is stored, without atomics, as the item's
 * block in 'tj'.  More wavefronts per tile keep every SIMD busy when tiles
 * are few (a 32^3 box has 512) and let the hardware balance tiles of very
 * different occupancy (workgroups finish at different times and new ones
 * start); each wavefront loads its next batch before pushing the current. */
struct Push32Args {
    Markers32    m, d;         /* current buffer (read through tv.perm), output buffer */
    F32View      v;
    Push32Consts c;
    TileView     tv;
    float       *tj;
    double      *Wacc;
};

template <int W>
__device__ __forceinline__ static void push32_tiles(const Markers32 &m, const Markers32 &d,
                                           const F32View &v, const Push32Consts &c,
                                           const TileView &tv, float *tj, double *Wacc,
                                           const VBlock &bk, const VBlock &gd)
{
    constexpr int BS = TW * W;                       /* threads of the workgroup */
    constexpr int NSTG = (TFN3 + BS - 1) / BS;       /* staged nodes per thread and component */
    __shared__ float lj[TJB];
    __shared__ float lf[6 * TFN3];
    __shared__ float lqm[4];
    if (threadIdx.x < 4) lqm[threadIdx.x] = c.qm[threadIdx.x];   /* ordered by the loop's barrier */
    const Grid &g = v.g;
    const int lane = threadIdx.x & 63;
    const int nitems = *tv.nitems;
    float w0 = 0, w1 = 0, w2 = 0;         /* per-thread kinetic energy partials */
    for (int it = (int)bk.x; it < nitems; it += (int)gd.x) {
        const int sl = tv.item[it];
        const int t = tv.active[sl];
        const int u = it - tv.istart[sl], nu = tv.istart[sl + 1] - tv.istart[sl];
        int t0, t1, t2;
        tile_coords(tv, t, t0, t1, t2);
        const int c0 = TILE * t0, c1 = TILE * t1, c2 = TILE * t2;
        const int o0 = c0 - 1, o1 = c1 - 1, o2 = c2 - 1;
        TilePos tp;
        tp.t = t;
        tp.s0 = tv.nt[1] * tv.nt[2];
        tp.s1 = tv.nt[2];
        tp.edge = !(c0 >= TILE && c0 + 2 * TILE <= g.n[0] && c1 >= TILE && c1 + 2 * TILE <= g.n[1] &&
                    c2 >= TILE && c2 + 2 * TILE <= g.n[2]);
        __syncthreads();                  /* previous tile's LDS reads are done */
        for (int q = threadIdx.x; q < TJB; q += BS) lj[q] = 0.0f;
        /* nodes [c, c + TFN) per axis; those past the last ghost layer (a
         * partial tile at the high edge) are never read by a gather and are
         * staged as 0.  Every load of a lane is issued before the first LDS
         * store, so staging costs one memory latency, not one per node. */
        {
            const long base = gidx(g, c0, c1, c2);
            unsigned offs[NSTG];
            bool ok[NSTG];
#pragma unroll
            for (int r = 0; r < NSTG; ++r) {
                const int rr = (int)threadIdx.x + BS * r;
                const int a = rr / (TFN * TFN), b = (rr / TFN) % TFN, cc = rr % TFN;
                ok[r] = rr < TFN3 && c0 + a <= g.n[0] + 2 && c1 + b <= g.n[1] + 2 &&
                        c2 + cc <= g.n[2] + 2;
                offs[r] = ok[r] ? (unsigned)((a * g.GY + b) * g.GZ + cc) : 0u;
            }
            float st[6][NSTG];
#pragma unroll
            for (int comp = 0; comp < 6; ++comp) {
                const float *F = (comp < 3 ? v.E[comp] : v.H[comp - 3]) + base;
#pragma unroll
                for (int r = 0; r < NSTG; ++r) st[comp][r] = at(F, offs[r]);
            }
#pragma unroll
            for (int comp = 0; comp < 6; ++comp)
#pragma unroll
                for (int r = 0; r < NSTG; ++r) {
                    const int rr = (int)threadIdx.x + BS * r;
                    if (rr < TFN3) lf[comp * TFN3 + rr] = ok[r] ? st[comp][r] : 0.0f;
                }
        }
        /* item u of nu: an even share of the tile's markers */
        const unsigned tb = tv.toff[t], cnt = tv.toff[t + 1] - tb;
        const unsigned beg = tb + (unsigned)((unsigned long long)u * cnt / nu),
                       end = tb + (unsigned)((unsigned long long)(u + 1) * cnt / nu);
        /* the order's indices are loaded two batches ahead and the records
         * one batch ahead, so no load waits for an index just loaded */
        MarkerRegs cur;
        load_marker(m, perm_at(tv.perm, beg + threadIdx.x, end - 1), cur);
        unsigned snx = perm_at(tv.perm, beg + BS + threadIdx.x, end - 1);
        /* everything outstanding completes here, so that the loop's waits
         * only ever concern the prefetch of the batch in flight */
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        Pending pend;
        pend.valid = false;
        pend.ckey = -1;
        for (unsigned base = beg; base < end; base += BS) {
            const unsigned p = base + threadIdx.x;
            MarkerRegs nxt;
            load_marker(m, snx, nxt);
            snx = perm_at(tv.perm, p + 2 * BS, end - 1);
            write_back(d, tv, pend);
            int rkey = -1, newword = -1, ci = 0, cj = 0, ck = 0;
            Seg12 first;
            for (int q = 0; q < 12; ++q) first.v[q] = 0.0f;
            const bool in = p < end && cur.word >= 0;
            if (in) {
                rkey = cur.word & KEY_MASK;
                const int loc = rkey & (TILE_CELLS - 1);
                const int li = loc >> 4, lj_ = (loc >> 2) & 3, lk = loc & 3;
                ci = c0 + li;
                cj = c1 + lj_;
                ck = c2 + lk;
                newword = push32_marker(pend.r, cur, li, lj_, lk, ci, cj, ck, lf, lqm, lj, o0, o1, o2,
                                        v, c, tv, tp, first, w0, w1, w2);
            }
            pend.r.id = cur.id;
            pend.valid = in;
            pend.p = p;
            pend.word = newword;
            int start, len;
            bool head;
            wave_run(rkey, start, len, head);
            row_scan(first, start);
            /* last lane of the run inside its row: adds the row's part */
            const int ln = threadIdx.x & 63;
            const bool tail = ln == start + len - 1 || ((ln & 15) == 15 && ln < start + len - 1);
            if (tail && rkey >= 0) seg_add(v, lj, o0, o1, o2, ci, cj, ck, first);
            /* count[new key] += run length, once per run of equal new keys */
            const int nkey = newword < 0 ? -1 : (newword & KEY_MASK);
            wave_run(nkey, start, len, head);
            pend.ckey = (head && nkey >= 0) ? nkey : -1;
            pend.clen = len;
            cur = nxt;
        }
        write_back(d, tv, pend);
        __syncthreads();
        /* the tile's J block, stored whole (coalesced, no atomics) */
        float *dst = tj + (long)it * TJB;
        for (int q = threadIdx.x; q < TJB; q += BS) dst[q] = lj[q];
    }
    double d0 = w0, d1 = w1, d2 = w2;     /* wave sums in FP64 */
    for (int d = 32; d > 0; d >>= 1) {
        d0 += __shfl_down(d0, d, 64);
        d1 += __shfl_down(d1, d, 64);
        d2 += __shfl_down(d2, d, 64);
    }
    if (lane == 0 && (d0 != 0.0 || d1 != 0.0 || d2 != 0.0)) {
        atomicAdd(Wacc + 0, d0 * c.vol);
        atomicAdd(Wacc + 1, d1 * c.vol);
        atomicAdd(Wacc + 2, d2 * c.vol);
    }
}

/* W wavefronts per workgroup (a distinct argument type per W, so that the
 * fused ensemble launches each shape on its own), register budget for 4
 * wavefronts per SIMD. */
template <int W> struct Push32ArgsW : Push32Args {
};

template <int W> struct KernelDef<Push32ArgsW<W>> {
    static constexpr int kBlock = TW * W;
    static constexpr int kWaves = 4;
    __device__ __forceinline__ static void body(const Push32ArgsW<W> &a, VBlock b, VBlock gd)
    {
        push32_tiles<W>(a.m, a.d, a.v, a.c, a.tv, a.tj, a.Wacc, b, gd);
    }
};

/* Range of tiles u (along one axis) whose J block [4u-1, 4u+5] holds node i. */
__device__ static inline void covering(int i, int nt, int &lo, int &hi)
{
    lo = (i - 2 + 1024) / 4 - 256;          /* ceil((i - 5) / 4) = floor((i - 2) / 4), i >= -1022 */
    hi = (i + 1 + 1024) / 4 - 256;          /* floor((i + 1) / 4) */
    lo = lo < 0 ? 0 : lo;
    hi = hi > nt - 1 ? nt - 1 : hi;
}

/* J += sum of the item blocks of the tiles covering each node, in FP64.
 * Every node of an active tile's block is written once, by the first active
 * tile (in tile order) whose block holds it; the sum runs over the covering
 * tiles in tile order and over each tile's items in order, so the gather
 * adds in a fixed order (what a block holds follows the device order of
 * the tile's markers, which the sort's slot reservation by atomics does not
 * fix from run to run).  One thread per node and all three
 * components: a node lies in at most 2 blocks per axis, and the first item
 * of each of the (at most 8) covering tiles is loaded before any is added,
 * so the gather waits for one memory latency per node, not one per block. */
#define TJN3 (TJN * TJN * TJN)
struct JGatherArgs {
    F32View      v;
    TileView     tv;
    const float *tj;
};

__device__ __forceinline__ static void jgather32(const F32View &v, const TileView &tv, const float *tj,
                                        const VBlock &bk, const VBlock &gd)
{
    __shared__ int nfirst[27], nitem[27];    /* items of the 3x3x3 tiles around t (0: none) */
    const Grid &g = v.g;
    const int nact = *tv.nactive;
    for (int sl = (int)bk.x; sl < nact; sl += (int)gd.x) {
        const int t = tv.active[sl];
        int t0, t1, t2;
        tile_coords(tv, t, t0, t1, t2);
        __syncthreads();
        if (threadIdx.x < 27) {
            const int ux = t0 - 1 + threadIdx.x / 9, uy = t1 - 1 + (threadIdx.x / 3) % 3,
                      uz = t2 - 1 + threadIdx.x % 3;
            const bool in = ux >= 0 && uy >= 0 && uz >= 0 && ux < tv.nt[0] && uy < tv.nt[1] &&
                            uz < tv.nt[2];
            const int su = in ? tv.slot[(ux * tv.nt[1] + uy) * tv.nt[2] + uz] : -1;
            nfirst[threadIdx.x] = su >= 0 ? tv.istart[su] : 0;
            nitem[threadIdx.x] = su >= 0 ? tv.istart[su + 1] - tv.istart[su] : 0;
        }
        __syncthreads();
        /* nodes in reverse order: the nodes an interior tile owns (block
         * positions 3..6 along each axis, the others belong to the tiles
         * below) all fall in the first pass of the workgroup, so a tile
         * inside the plasma waits on memory once */
        for (int pass = threadIdx.x; pass < TJN3; pass += TPB) {
            const int node = TJN3 - 1 - pass;
            const int cc = node % TJN, bb = (node / TJN) % TJN, aa = node / (TJN * TJN);
            const int i = TILE * t0 - 1 + aa, j = TILE * t1 - 1 + bb, k = TILE * t2 - 1 + cc;
            if (i < -2 || j < -2 || k < -2 || i > g.n[0] + 2 || j > g.n[1] + 2 || k > g.n[2] + 2)
                continue;
            int xl, xh, yl, yh, zl, zh;
            covering(i, tv.nt[0], xl, xh);
            covering(j, tv.nt[1], yl, yh);
            covering(k, tv.nt[2], zl, zh);
            /* the covering tiles (2 per axis at most), in tile order */
            int nb[8], off[8];
            bool use[8];
            int own = -1;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int ux = xl + (q >> 2), uy = yl + ((q >> 1) & 1), uz = zl + (q & 1);
                const bool inr = ux <= xh && uy <= yh && uz <= zh;
                const int b = inr ? ((ux - t0 + 1) * 3 + (uy - t1 + 1)) * 3 + (uz - t2 + 1) : 13;
                use[q] = inr && nitem[b] > 0;
                nb[q] = b;
                const int a = i - (TILE * ux - 1), bq = j - (TILE * uy - 1), cz = k - (TILE * uz - 1);
                off[q] = use[q] ? ((a * TJN + bq) * TJN + cz) * 3 : 0;
                if (use[q] && own < 0) own = (ux * tv.nt[1] + uy) * tv.nt[2] + uz;
            }
            if (own != t) continue;
            const long o = gidx(g, i, j, k);
            const double J0 = v.J[0][o], J1 = v.J[1][o], J2 = v.J[2][o];
            float f[8][3];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const float *blk = tj + (long)(use[q] ? nfirst[nb[q]] : 0) * TJB + off[q];
                f[q][0] = blk[0];
                f[q][1] = blk[1];
                f[q][2] = blk[2];
            }
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                if (!use[q]) continue;
                s0 += (double)f[q][0];
                s1 += (double)f[q][1];
                s2 += (double)f[q][2];
                const int n_it = nitem[nb[q]];
                for (int w = 1; w < n_it; ++w) {          /* tiles split into several items */
                    const float *blk = tj + (long)(nfirst[nb[q]] + w) * TJB + off[q];
                    s0 += (double)blk[0];
                    s1 += (double)blk[1];
                    s2 += (double)blk[2];
                }
            }
            v.J[0][o] = J0 + s0;
            v.J[1][o] = J1 + s1;
            v.J[2][o] = J2 + s2;
        }
    }
}

template <> struct KernelDef<JGatherArgs> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const JGatherArgs &a, VBlock b, VBlock gd)
    {
        jgather32(a.v, a.tv, a.tj, b, gd);
    }
};

int launch_push32(hipStream_t s, const Markers32 &m, const Markers32 &dst, const Fields32 &f,
                  const Grid &g, const PushConsts &pc, const float *qm, const TileView &tv,
                  float *tj, double *Wacc, int nblocks, int waves)
{
    F32View v;
    for (int a = 0; a < 3; ++a) {
        v.E[a] = f.E[a];
        v.H[a] = f.H[a];
        v.J[a] = f.J[a];
    }
    v.g = g;
    Push32Consts c;
    c.tau = (float)g.tau;
    for (int a = 0; a < 3; ++a) {
        c.dtn[a] = (float)(g.tau / g.h[a]);
        c.cj[a] = (float)(g.h[a] / g.tau);
    }
    c.vol = g.h[0] * g.h[1] * g.h[2];
    for (int q = 0; q < 4; ++q) c.qm[q] = qm[q];
    for (int b = 0; b < 6; ++b) c.bc[b] = pc.bc[b];
    const Push32Args a = {m, dst, v, c, tv, tj, Wacc};
    switch (waves) {
    case 1: return launch_def(s, Push32ArgsW<1>{a}, (unsigned)nblocks);
    case 2: return launch_def(s, Push32ArgsW<2>{a}, (unsigned)nblocks);
    case 4: return launch_def(s, Push32ArgsW<4>{a}, (unsigned)nblocks);
    default:
        mpic_set_error("launch_push32: %d wavefronts per tile (1, 2 or 4)", waves);
        return MPIC_ERR_ARG;
    }
}

int launch_jgather32(hipStream_t s, const Fields32 &f, const Grid &g, const TileView &tv,
                     const float *tj, int nblocks)
{
    F32View v;
    for (int a = 0; a < 3; ++a) {
        v.E[a] = f.E[a];
        v.H[a] = f.H[a];
        v.J[a] = f.J[a];
    }
    v.g = g;
    JGatherArgs a = {v, tv, tj};
    return launch_def(s, a, (unsigned)nblocks);
}

/* Keys and counts of the current buffer's first nmark[parity] markers
 * (the cell word holds the key; a deleted marker's is -1): the start of a
 * sort that moves nothing (loaded markers, or a re-sort). */
__global__ void k_keys32(Markers32 m, long n, TileView tv, int parity)
{
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long live = tv.nmark[parity];
    const int word = (p < n && p < live) ? ((const int *)m.pos)[4 * p] : -1;
    const int key = word < 0 ? -1 : (word & KEY_MASK);
    if (p < n) tv.key[p] = key;
    wave_run_count(tv.count, key);
}

int launch_keys32(hipStream_t s, const Markers32 &m, long n, const TileView &tv, int parity)
{
    if (n <= 0) return MPIC_OK;
    k_keys32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(m, n, tv, parity);
    MPIC_CHECK(hipGetLastError());
    return MPIC_OK;
}

/* CIC charge density of the split-precision markers of the order, FP64. */
__global__ void k_rho32(Markers32 m, long n, double *rho, Grid g, TileView tv, int parity)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n || (tv.perm && q >= tv.nmark[parity])) return;
    const long p = tv.perm ? tv.perm[q] : q;            /* no permutation: the loaded order */
    const float4 a = m.pos[p];
    const int word = __float_as_int(a.x);
    if (word < 0) return;
    int i, j, k;
    key_to_ijk(tv, word & KEY_MASK, i, j, k);
    const double sx = a.y, sy = a.z, sz = a.w, val = m.mom[p].w;
    unsafeAtomicAdd(rho + gidx(g, i, j, k), (1 - sx) * ((1 - sy) * ((1 - sz) * val)));
    unsafeAtomicAdd(rho + gidx(g, i + 1, j, k), sx * ((1 - sy) * ((1 - sz) * val)));
    unsafeAtomicAdd(rho + gidx(g, i, j + 1, k), (1 - sx) * (sy * ((1 - sz) * val)));
    unsafeAtomicAdd(rho + gidx(g, i + 1, j + 1, k), sx * (sy * ((1 - sz) * val)));
    unsafeAtomicAdd(rho + gidx(g, i, j, k + 1), (1 - sx) * ((1 - sy) * (sz * val)));
    unsafeAtomicAdd(rho + gidx(g, i + 1, j, k + 1), sx * ((1 - sy) * (sz * val)));
    unsafeAtomicAdd(rho + gidx(g, i, j + 1, k + 1), (1 - sx) * (sy * (sz * val)));
    unsafeAtomicAdd(rho + gidx(g, i + 1, j + 1, k + 1), sx * (sy * (sz * val)));
}

int launch_rho32(hipStream_t s, const Markers32 &m, long n, double *rho, const Grid &g,
                 const TileView &tv, int parity)
{
    if (n <= 0) return MPIC_OK;
    k_rho32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(m, n, rho, g, tv, parity);
    MPIC_CHECK(hipGetLastError());
    return MPIC_OK;
}

/* Natural cell index (i*n1 + j)*n2 + k of every marker of the order
 * (-1: deleted). */
__global__ void k_cells32(Markers32 m, long n, Grid g, TileView tv, int parity, int32_t *cell)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long p = tv.perm ? (q < tv.nmark[parity] ? tv.perm[q] : -1) : q;
    const int word = p >= 0 ? ((const int *)m.pos)[4 * p] : -1;
    if (word < 0) {
        cell[q] = -1;
        return;
    }
    int i, j, k;
    key_to_ijk(tv, word & KEY_MASK, i, j, k);
    cell[q] = (i * g.n[1] + j) * g.n[2] + k;
}

int launch_cells32(hipStream_t s, const Markers32 &m, long n, const Grid &g, const TileView &tv,
                   int parity, int32_t *cell)
{
    if (n <= 0) return MPIC_OK;
    k_cells32<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(m, n, g, tv, parity, cell);
    MPIC_CHECK(hipGetLastError());
    return MPIC_OK;
}
