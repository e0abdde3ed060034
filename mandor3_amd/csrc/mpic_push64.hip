/*
 * mpic_push64.hip -- FP64 plain-form marker push and current deposit.
 *
 * One thread per marker: trilinear gather of E and H on the staggered Yee
 * grid (source/misc_PIC.h:18-126), relativistic Boris rotation
 * (source/core/plasma.c:115-171), charge-conserving deposit of the straight
 * trajectory split at cell faces (source/core/plasma_currentKernel.c:20-279)
 * and the particle walls (source/core/plasma_walls.c:27-104).  Built with
 * -ffp-contract=off so each marker's arithmetic is the reference's, bit for
 * bit; J is accumulated with FP64 atomics, so only the summation order of J
 * differs from the CPU path.
 *
 * Roofline: HBM-bound.  Algorithmic bytes per marker-step: read 56 B +
 * write 52 B of SoA marker data (x y z ux uy uz, w qm, id read once), plus
 * per cell 48 B of E/H reads and 48 B of J read-modify-write divided by the
 * markers per cell.
 */
#include "mpic_internal.h"
#include "mpic_launch.h"
#include "mpic_tiles.h"

#define TPB 256

#define DBL_TO_INT(x) ((int)((x) + 128.0) - 128)

struct F64View {
    const double *E[3], *H[3];
    double *J[3];
    Grid g;
};

__device__ static inline double gather(const double *F, const Grid &g, int i, double sx, int j,
                                       double sy, int k, double sz)
{
    return F[gidx(g, i, j, k)] * (1 - sx) * ((1 - sy) * (1 - sz)) +
           F[gidx(g, i + 1, j, k)] * sx * ((1 - sy) * (1 - sz)) +
           F[gidx(g, i, j + 1, k)] * (1 - sx) * (sy * (1 - sz)) +
           F[gidx(g, i + 1, j + 1, k)] * sx * (sy * (1 - sz)) +
           F[gidx(g, i, j, k + 1)] * (1 - sx) * ((1 - sy) * sz) +
           F[gidx(g, i + 1, j, k + 1)] * sx * ((1 - sy) * sz) +
           F[gidx(g, i, j + 1, k + 1)] * (1 - sx) * (sy * sz) +
           F[gidx(g, i + 1, j + 1, k + 1)] * sx * (sy * sz);
}

/* Deposit targets.  GlobalSink adds straight to J in HBM with FP64
 * atomics; TileSink (tiled kernel) adds to a 7x7x7-node J block in LDS that
 * covers a 4x4x4-cell tile and its one-cell rim, and falls back to HBM for a
 * segment outside it. */
struct GlobalSink {
    double *J[3];
    Grid g;
    __device__ void seg(int i, int j, int k, const double vx[4], const double vy[4],
                        const double vz[4]) const
    {
        unsafeAtomicAdd(J[0] + gidx(g, i + 1, j, k), vx[0]);
        unsafeAtomicAdd(J[0] + gidx(g, i + 1, j + 1, k), vx[1]);
        unsafeAtomicAdd(J[0] + gidx(g, i + 1, j, k + 1), vx[2]);
        unsafeAtomicAdd(J[0] + gidx(g, i + 1, j + 1, k + 1), vx[3]);
        unsafeAtomicAdd(J[1] + gidx(g, i, j + 1, k), vy[0]);
        unsafeAtomicAdd(J[1] + gidx(g, i + 1, j + 1, k), vy[1]);
        unsafeAtomicAdd(J[1] + gidx(g, i, j + 1, k + 1), vy[2]);
        unsafeAtomicAdd(J[1] + gidx(g, i + 1, j + 1, k + 1), vy[3]);
        unsafeAtomicAdd(J[2] + gidx(g, i, j, k + 1), vz[0]);
        unsafeAtomicAdd(J[2] + gidx(g, i + 1, j, k + 1), vz[1]);
        unsafeAtomicAdd(J[2] + gidx(g, i, j + 1, k + 1), vz[2]);
        unsafeAtomicAdd(J[2] + gidx(g, i + 1, j + 1, k + 1), vz[3]);
    }
};

#define TILE 4
/* TJN (mpic_tiles.h): LDS J nodes per axis, the tile plus its rim */

struct TileSink {
    GlobalSink gs;
    double *lj;                    /* LDS: [TJN^3][3] */
    int o0, o1, o2;                /* node of lj[0] */
    __device__ void seg(int i, int j, int k, const double vx[4], const double vy[4],
                        const double vz[4]) const
    {
        const int a = i - o0, b = j - o1, c = k - o2;
        if (a < 0 || b < 0 || c < 0 || a > TJN - 2 || b > TJN - 2 || c > TJN - 2) {
            gs.seg(i, j, k, vx, vy, vz);
            return;
        }
        lds_add(a, b, c, vx, vy, vz);
    }
    /* the 12 values of a segment in the block's cell (a, b, c) */
    __device__ void lds_add(int a, int b, int c, const double vx[4], const double vy[4],
                            const double vz[4]) const
    {
#define L(aa, bb, cc, comp) (lj + 3 * (((aa) * TJN + (bb)) * TJN + (cc)) + (comp))
        atomicAdd(L(a + 1, b, c, 0), vx[0]);
        atomicAdd(L(a + 1, b + 1, c, 0), vx[1]);
        atomicAdd(L(a + 1, b, c + 1, 0), vx[2]);
        atomicAdd(L(a + 1, b + 1, c + 1, 0), vx[3]);
        atomicAdd(L(a, b + 1, c, 1), vy[0]);
        atomicAdd(L(a + 1, b + 1, c, 1), vy[1]);
        atomicAdd(L(a, b + 1, c + 1, 1), vy[2]);
        atomicAdd(L(a + 1, b + 1, c + 1, 1), vy[3]);
        atomicAdd(L(a, b, c + 1, 2), vz[0]);
        atomicAdd(L(a + 1, b, c + 1, 2), vz[1]);
        atomicAdd(L(a, b + 1, c + 1, 2), vz[2]);
        atomicAdd(L(a + 1, b + 1, c + 1, 2), vz[3]);
#undef L
    }
};

/* The 12 current values of one straight segment inside cell (i,j,k). */
__device__ static inline void seg_values(const Grid &g, int i, int j, int k, double x1, double x2,
                                         double y1, double y2, double z1, double z2, double rp,
                                         double vx[4], double vy[4], double vz[4])
{
    const double h1 = g.h[0], h2 = g.h[1], h3 = g.h[2], tau = g.tau;
    const double dX = x2 - x1, dY = y2 - y1, dZ = z2 - z1;
    const double tX = 0.5 * (x1 + x2) - i * h1;
    const double tY = 0.5 * (y1 + y2) - j * h2;
    const double tZ = 0.5 * (z1 + z2) - k * h3;
    const double yz = dY * dZ / 12, xz = dX * dZ / 12, xy = dX * dY / 12;
    const double ax = rp * dX / (tau * h2 * h3);
    const double ay = rp * dY / (tau * h1 * h3);
    const double az = rp * dZ / (tau * h1 * h2);
    vx[0] = ax * ((h2 - tY) * (h3 - tZ) + yz);
    vx[1] = ax * (tY * (h3 - tZ) - yz);
    vx[2] = ax * ((h2 - tY) * tZ - yz);
    vx[3] = ax * (tY * tZ + yz);
    vy[0] = ay * ((h1 - tX) * (h3 - tZ) + xz);
    vy[1] = ay * (tX * (h3 - tZ) - xz);
    vy[2] = ay * ((h1 - tX) * tZ - xz);
    vy[3] = ay * (tX * tZ + xz);
    vz[0] = az * ((h1 - tX) * (h2 - tY) + xy);
    vz[1] = az * (tX * (h2 - tY) - xy);
    vz[2] = az * ((h1 - tX) * tY - xy);
    vz[3] = az * (tX * tY + xy);
}

/* Deposit of one straight segment inside cell (i,j,k). */
template <class S>
__device__ static void deposit_seg(const S &v, const Grid &g, int i, int j, int k, double x1,
                                   double x2, double y1, double y2, double z1, double z2,
                                   double rp)
{
    double vx[4], vy[4], vz[4];
    seg_values(g, i, j, k, x1, x2, y1, y2, z1, z2, rp, vx, vy, vz);
    v.seg(i, j, k, vx, vy, vz);
}

/* The first straight segment of a marker's path, in its start cell: the
 * tiled push keeps it (kKeep) to sum it over the run of markers that share
 * the cell (the markers are sorted by cell) before it reaches LDS. */
struct FirstSeg {
    int    i, j, k, on;
    double x1, x2, y1, y2, z1, z2;
};

template <bool kKeep, class S>
__device__ static inline void first_seg(const S &v, const Grid &g, FirstSeg *fs, int i, int j,
                                        int k, double x1, double x2, double y1, double y2,
                                        double z1, double z2, double rp)
{
    if constexpr (kKeep) {
        fs->i = i;
        fs->j = j;
        fs->k = k;
        fs->on = 1;
        fs->x1 = x1;
        fs->x2 = x2;
        fs->y1 = y1;
        fs->y2 = y2;
        fs->z1 = z1;
        fs->z2 = z2;
    } else {
        deposit_seg(v, g, i, j, k, x1, x2, y1, y2, z1, z2, rp);
    }
}

/* Two or three face crossings, source/core/plasma_currentKernel.c:156-279. */
template <bool kKeep, class S>
__device__ static void deposit_multi(const S &v, const Grid &g, int i1, int j1, int k1, int i2,
                                     int j2, int k2, double x1, double y1, double z1, double x2,
                                     double y2, double z2, double rp, FirstSeg *fs)
{
    const double h1 = g.h[0], h2 = g.h[1], h3 = g.h[2];
    const int di = abs(i2 - i1), dj = abs(j2 - j1), dk = abs(k2 - k1);
    if (di + dj + dk == 2) {
        int i = i1, j = j1, k = k1;
        const double sgX = di * ((i1 + i2 + 1) * 0.5 * h1 - x1) / (x2 - x1 + ((1 - di) << 16));
        const double sgY = dj * ((j1 + j2 + 1) * 0.5 * h2 - y1) / (y2 - y1 + ((1 - dj) << 16));
        const double sgZ = dk * ((k1 + k2 + 1) * 0.5 * h3 - z1) / (z2 - z1 + ((1 - dk) << 16));
        double k1s = sgX + (1 - di) * sgY;
        double k2s = sgZ + (1 - dk) * sgY;
        /* the two crossings' cell steps, in the order they happen (selected
         * by value: an array picked by pointer would live in scratch) */
        const int dI0[3] = {di * (i2 - i1), (1 - di) * (j2 - j1), 0};
        const int dI1[3] = {0, (1 - dk) * (j2 - j1), dk * (k2 - k1)};
        int st = 0;
        if (k1s > k2s) {
            double t = k1s;
            k1s = k2s;
            k2s = t;
            st = 1;
        }
        const double xA = x1 + k1s * (x2 - x1), yA = y1 + k1s * (y2 - y1), zA = z1 + k1s * (z2 - z1);
        const double xB = x1 + k2s * (x2 - x1), yB = y1 + k2s * (y2 - y1), zB = z1 + k2s * (z2 - z1);
        first_seg<kKeep>(v, g, fs, i, j, k, x1, xA, y1, yA, z1, zA, rp);
        i += st ? dI1[0] : dI0[0];
        j += st ? dI1[1] : dI0[1];
        k += st ? dI1[2] : dI0[2];
        deposit_seg(v, g, i, j, k, xA, xB, yA, yB, zA, zB, rp);
        i += st ? dI0[0] : dI1[0];
        j += st ? dI0[1] : dI1[1];
        k += st ? dI0[2] : dI1[2];
        deposit_seg(v, g, i, j, k, xB, x2, yB, y2, zB, z2, rp);
        return;
    }
    const double xC = (i1 + i2 + 1) * 0.5 * h1, yC = (j1 + j2 + 1) * 0.5 * h2,
                 zC = (k1 + k2 + 1) * 0.5 * h3;
    double sg[3] = {(xC - x1) / (x2 - x1), (yC - y1) / (y2 - y1), (zC - z1) / (z2 - z1)};
    double rc[3][3];
    rc[0][0] = xC;
    rc[0][1] = y1 + sg[0] * (y2 - y1);
    rc[0][2] = z1 + sg[0] * (z2 - z1);
    rc[1][0] = x1 + sg[1] * (x2 - x1);
    rc[1][1] = yC;
    rc[1][2] = z1 + sg[1] * (z2 - z1);
    rc[2][0] = x1 + sg[2] * (x2 - x1);
    rc[2][1] = y1 + sg[2] * (y2 - y1);
    rc[2][2] = zC;
    const int dI[3] = {i2 - i1, j2 - j1, k2 - k1};
    int s1 = 0, s2 = 1, s3 = 2;
    double a1 = sg[0], a2 = sg[1], a3 = sg[2];
    if (a1 > a2) { int t = s1; s1 = s2; s2 = t; double d = a1; a1 = a2; a2 = d; }
    if (a1 > a3) { int t = s1; s1 = s3; s3 = t; double d = a1; a1 = a3; a3 = d; }
    if (a2 > a3) { int t = s2; s2 = s3; s3 = t; double d = a2; a2 = a3; a3 = d; }
    int ii[3] = {i1, j1, k1};
    first_seg<kKeep>(v, g, fs, ii[0], ii[1], ii[2], x1, rc[s1][0], y1, rc[s1][1], z1, rc[s1][2],
                     rp);
    ii[s1] += dI[s1];
    deposit_seg(v, g, ii[0], ii[1], ii[2], rc[s1][0], rc[s2][0], rc[s1][1], rc[s2][1], rc[s1][2],
                rc[s2][2], rp);
    ii[s2] += dI[s2];
    deposit_seg(v, g, ii[0], ii[1], ii[2], rc[s2][0], rc[s3][0], rc[s2][1], rc[s3][1], rc[s2][2],
                rc[s3][2], rp);
    deposit_seg(v, g, i2, j2, k2, rc[s3][0], x2, rc[s3][1], y2, rc[s3][2], z2, rp);
}

/* Particle walls in the reference order xMin yMin zMin xMax yMax zMax.
 * Returns 0 when an open wall deletes the marker. */
__device__ static inline int walls(const PushConsts &c, double *r, double *u)
{
    for (int b = 0; b < 6; ++b) {
        const int ax = b % 3, top = b / 3, bc = c.bc[b];
        const double bnd = top ? c.shell2[ax] : c.shell1[ax];
        if (bc == MPIC_BC_PERIODIC) {
            if (top ? (r[ax] >= bnd) : (r[ax] <= bnd)) r[ax] += top ? -c.L[ax] : c.L[ax];
        } else if (bc == MPIC_BC_MIRROR) {
            if (top ? (r[ax] >= bnd) : (r[ax] <= bnd)) {
                r[ax] = 2 * bnd - r[ax];
                u[ax] *= -1;
            }
        } else if (bc == MPIC_BC_OPEN) {
            const double lim = top ? bnd - c.h[ax] : bnd + c.h[ax];
            if (top ? (r[ax] >= lim) : (r[ax] <= lim)) return 0;
        }
    }
    return 1;
}

/* Gather, Boris and deposit of one live marker; returns 0 for a marker off
 * the grid (deleted and flagged), else applies the walls and returns
 * whether the marker survives.  r/u are in/out. */
template <bool kKeep = false, class S>
__device__ static int push_marker(const S &sink, const F64View &v, const PushConsts &c, double r[3],
                                  double u[3], float rhof, float qmf, double w3[3], unsigned *err,
                                  FirstSeg *fs = nullptr)
{
    const Grid &g = v.g;
    const double h1 = g.h[0], h2 = g.h[1], h3 = g.h[2], tau = g.tau;
    const double x1 = r[0], y1 = r[1], z1 = r[2];
    if (!(x1 / h1 > -1.0 && x1 / h1 < g.n[0] + 1.0 && y1 / h2 > -1.0 &&
          y1 / h2 < g.n[1] + 1.0 && z1 / h3 > -1.0 && z1 / h3 < g.n[2] + 1.0)) {
        atomicOr(err, 1u);
        return 0;
    }
    double sx, sy, sz, sx_, sy_, sz_;
    int i, j, k, i_, j_, k_;
    sx = x1 / h1; sx_ = sx + 0.5; i = DBL_TO_INT(sx); i_ = DBL_TO_INT(sx_); sx -= i; sx_ -= i_;
    sy = y1 / h2; sy_ = sy + 0.5; j = DBL_TO_INT(sy); j_ = DBL_TO_INT(sy_); sy -= j; sy_ -= j_;
    sz = z1 / h3; sz_ = sz + 0.5; k = DBL_TO_INT(sz); k_ = DBL_TO_INT(sz_); sz -= k; sz_ -= k_;
    const double Ex = gather(v.E[0], g, i_, sx_, j, sy, k, sz),
                 Ey = gather(v.E[1], g, i, sx, j_, sy_, k, sz),
                 Ez = gather(v.E[2], g, i, sx, j, sy, k_, sz_),
                 Hx = gather(v.H[0], g, i, sx, j_, sy_, k_, sz_),
                 Hy = gather(v.H[1], g, i_, sx_, j, sy, k_, sz_),
                 Hz = gather(v.H[2], g, i_, sx_, j_, sy_, k, sz);
    const double tauMass = 0.5 * tau * qmf, px = u[0] + tauMass * Ex, py = u[1] + tauMass * Ey,
                 pz = u[2] + tauMass * Ez, gamma = sqrt(1 + px * px + py * py + pz * pz);
    const double W = (gamma - 1.0) * rhof / (px * px + py * py + pz * pz + 1e-19) * h1 * h2 * h3 / qmf;
    w3[0] += W * px * px;
    w3[1] += W * py * py;
    w3[2] += W * pz * pz;
    const double tauGamma = tauMass / gamma;
    const double aHx = tauGamma * Hx, aHy = tauGamma * Hy, aHz = tauGamma * Hz;
    const double uu = px + py * aHz - pz * aHy, vv = py + pz * aHx - px * aHz,
                 ww = pz + px * aHy - py * aHx;
    const double bxx = aHx * aHx, byy = aHy * aHy, bzz = aHz * aHz, bxy = aHx * aHy,
                 byz = aHy * aHz, bxz = aHz * aHx;
    const double delta = 1.0 / (1 + bxx + byy + bzz);
    u[0] = tauMass * Ex + ((1 + bxx) * uu + (bxy + aHz) * vv + (bxz - aHy) * ww) * delta;
    u[1] = tauMass * Ey + ((bxy - aHz) * uu + (1 + byy) * vv + (byz + aHx) * ww) * delta;
    u[2] = tauMass * Ez + ((bxz + aHy) * uu + (byz - aHx) * vv + (1 + bzz) * ww) * delta;
    const double tg = tau / sqrt(1 + u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double x2 = x1 + u[0] * tg, y2 = y1 + u[1] * tg, z2 = z1 + u[2] * tg;
    int i2 = DBL_TO_INT(x2 / h1), j2 = DBL_TO_INT(y2 / h2), k2 = DBL_TO_INT(z2 / h3);
    int di = abs(i2 - i), dj = abs(j2 - j), dk = abs(k2 - k);
    if (fabs(x2 - x1) < 1e-10 * h1) { di = 0; i2 = i; }
    if (fabs(y2 - y1) < 1e-10 * h2) { dj = 0; j2 = j; }
    if (fabs(z2 - z1) < 1e-10 * h3) { dk = 0; k2 = k; }
    const double rp = rhof;
    if (di + dj + dk == 0) {
        first_seg<kKeep>(sink, g, fs, i, j, k, x1, x2, y1, y2, z1, z2, rp);
    } else if (di + dj + dk == 1) {
        const double alpha = dk * ((k + k2 + 1) * 0.5 * h3 - z1) / (z2 - z1 + ((1 - dk) << 16))
                           + dj * ((j + j2 + 1) * 0.5 * h2 - y1) / (y2 - y1 + ((1 - dj) << 16))
                           + di * ((i + i2 + 1) * 0.5 * h1 - x1) / (x2 - x1 + ((1 - di) << 16));
        const double xC = x1 + alpha * (x2 - x1), yC = y1 + alpha * (y2 - y1),
                     zC = z1 + alpha * (z2 - z1);
        first_seg<kKeep>(sink, g, fs, i, j, k, x1, xC, y1, yC, z1, zC, rp);
        deposit_seg(sink, g, i2, j2, k2, xC, x2, yC, y2, zC, z2, rp);
    } else {
        deposit_multi<kKeep>(sink, g, i, j, k, i2, j2, k2, x1, y1, z1, x2, y2, z2, rp, fs);
    }
    r[0] = x2;
    r[1] = y2;
    r[2] = z2;
    return walls(c, r, u);
}

__device__ static void block_sum3(double w3[3], double *Wacc)
{
    __shared__ double sw[3][TPB];
    sw[0][threadIdx.x] = w3[0];
    sw[1][threadIdx.x] = w3[1];
    sw[2][threadIdx.x] = w3[2];
    __syncthreads();
    for (int s = TPB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sw[0][threadIdx.x] += sw[0][threadIdx.x + s];
            sw[1][threadIdx.x] += sw[1][threadIdx.x + s];
            sw[2][threadIdx.x] += sw[2][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(Wacc + 0, sw[0][0]);
        atomicAdd(Wacc + 1, sw[1][0]);
        atomicAdd(Wacc + 2, sw[2][0]);
    }
}

/* Untiled push: one thread per marker, J straight to HBM.  mode 1 applies
 * the walls only (plasma_init).  A marker whose cell lies outside [-1, n]
 * (a state the reference never produces) is deleted and flagged in *err
 * instead of touching memory off the grid. */
struct Push64Args {
    Markers64  m;
    long       n;
    F64View    v;
    PushConsts c;
    double    *Wacc;
    int        mode;
    unsigned  *err;
};

__device__ __forceinline__ static void push64(const Markers64 &m, long n, const F64View &v,
                                     const PushConsts &c, double *Wacc, int mode, unsigned *err,
                                     long p)
{
    double w3[3] = {0, 0, 0};
    if (p < n && m.qm[p] != 0.0f) {
        double r[3] = {m.x[p], m.y[p], m.z[p]};
        double u[3] = {m.ux[p], m.uy[p], m.uz[p]};
        GlobalSink gs;
        for (int a = 0; a < 3; ++a) gs.J[a] = v.J[a];
        gs.g = v.g;
        int alive = (mode == 0) ? push_marker(gs, v, c, r, u, m.w[p], m.qm[p], w3, err)
                                : walls(c, r, u);
        if (!alive) {
            m.qm[p] = 0.0f;
            m.w[p] = 0.0f;
        } else {
            m.x[p] = r[0];
            m.y[p] = r[1];
            m.z[p] = r[2];
            m.ux[p] = u[0];
            m.uy[p] = u[1];
            m.uz[p] = u[2];
        }
    }
    if (mode != 0) return;
    block_sum3(w3, Wacc);
}

template <> struct KernelDef<Push64Args> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const Push64Args &a, VBlock b, VBlock)
    {
        push64(a.m, a.n, a.v, a.c, a.Wacc, a.mode, a.err, (long)b.x * TPB + threadIdx.x);
    }
};

int launch_push64(hipStream_t s, const Markers64 &m, long n, const Fields64 &f, const Grid &g,
                  const PushConsts &c, double *Wacc, int mode, unsigned *err)
{
    if (n <= 0) return MPIC_OK;
    F64View v;
    for (int a = 0; a < 3; ++a) {
        v.E[a] = f.E[a];
        v.H[a] = f.H[a];
        v.J[a] = f.J[a];
    }
    v.g = g;
    Push64Args a = {m, n, v, c, Wacc, mode, err};
    return launch_def(s, a, (unsigned)((n + TPB - 1) / TPB));
}

/* Row shift of a double by D lanes inside rows of 16 (DPP row_shr on
 * both halves; 0 from outside the row).  Every lane must be active. */
template <int D> __device__ static inline double row_shr(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x110 + D, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x110 + D, 0xf, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

/* The first segments of a batch (one marker per lane, lanes in cell
 * order): summed over each run of lanes in one cell within each row of 16
 * lanes, then added to LDS by the run's last lane in the row (at most 4
 * atomics per value and run instead of one per marker: LDS atomics to one
 * address serialise).  Called by every lane of the wavefront. */
__device__ static inline void first_run_add(const TileSink &ts, const Grid &g, const FirstSeg &fs,
                                            double rp)
{
    double vv[12];
    int cell = -1, a = 0, b = 0, c = 0;
    if (fs.on) {
        seg_values(g, fs.i, fs.j, fs.k, fs.x1, fs.x2, fs.y1, fs.y2, fs.z1, fs.z2, rp, vv, vv + 4,
                   vv + 8);
        a = fs.i - ts.o0;
        b = fs.j - ts.o1;
        c = fs.k - ts.o2;
        if (a < 0 || b < 0 || c < 0 || a > TJN - 2 || b > TJN - 2 || c > TJN - 2)
            ts.gs.seg(fs.i, fs.j, fs.k, vv, vv + 4, vv + 8);   /* outside the block: to HBM */
        else
            cell = (a * TJN + b) * TJN + c;
    }
    if (cell < 0)
        for (int q = 0; q < 12; ++q) vv[q] = 0.0;
    int start, len;
    bool head;
    wave_run(cell, start, len, head);
    const int lane = threadIdx.x & 63;
    /* fma(y, 1, x) is x + y rounded once, fma(y, 0, x) is x: one
     * instruction instead of a select of y and an add */
    const double m1 = lane - 1 >= start ? 1.0 : 0.0, m2 = lane - 2 >= start ? 1.0 : 0.0,
                 m4 = lane - 4 >= start ? 1.0 : 0.0, m8 = lane - 8 >= start ? 1.0 : 0.0;
#pragma unroll
    for (int q = 0; q < 12; ++q) {
        double x = vv[q];
        x = __builtin_fma(row_shr<1>(x), m1, x);
        x = __builtin_fma(row_shr<2>(x), m2, x);
        x = __builtin_fma(row_shr<4>(x), m4, x);
        x = __builtin_fma(row_shr<8>(x), m8, x);
        vv[q] = x;
    }
    const bool tail = lane == start + len - 1 || ((lane & 15) == 15 && lane < start + len - 1);
    if (tail && cell >= 0) ts.lds_add(a, b, c, vv, vv + 4, vv + 8);
}

/* Tiled push: persistent workgroups pull 4x4x4-cell tiles (the order's
 * markers by tile-major cell key, read through its permutation) from
 * per-XCD queues, accumulate the tile's J in LDS and flush it to HBM once;
 * every marker is written, with its id and new cell key, to the output
 * buffer at its position in the order, and the per-cell counts of the new
 * positions are kept for the next sort. */
struct Push64TArgs {
    Markers64  m, d;           /* current buffer (read through tv.perm), output buffer */
    F64View    v;
    PushConsts c;
    TileView   tv;
    double    *Wacc;
    unsigned  *err;
};

__device__ __forceinline__ static void push64_tiled(const Markers64 &m, const Markers64 &d,
                                           const F64View &v, const PushConsts &c,
                                           const TileView &tv, double *Wacc, unsigned *err,
                                           const VBlock &bk)
{
    __shared__ double lj[TJN * TJN * TJN * 3];
    __shared__ int s_tile;
    const Grid &g = v.g;
    const int xcd = bk.x % MPIC_XCDS;
    double w3[3] = {0, 0, 0};
    GlobalSink gs;
    for (int a = 0; a < 3; ++a) gs.J[a] = v.J[a];
    gs.g = g;
    for (;;) {
        if (threadIdx.x == 0) s_tile = tile_next(tv, xcd);
        __syncthreads();
        const int t = s_tile;
        if (t < 0) break;
        int t0, t1, t2;
        tile_coords(tv, t, t0, t1, t2);
        TileSink ts;
        ts.gs = gs;
        ts.lj = lj;
        ts.o0 = TILE * t0 - 1;
        ts.o1 = TILE * t1 - 1;
        ts.o2 = TILE * t2 - 1;
        for (int q = threadIdx.x; q < TJN * TJN * TJN * 3; q += TPB) lj[q] = 0.0;
        __syncthreads();
        const long beg = tv.toff[t], end = tv.toff[t + 1];
        for (long base = beg; base < end; base += TPB) {
            const long p = base + threadIdx.x;
            int key = -1;
            FirstSeg fs;
            fs.on = 0;
            float wf = 0.0f;
            if (p < end) {
                const long q = tv.perm[p];
                double r[3] = {m.x[q], m.y[q], m.z[q]};
                double u[3] = {m.ux[q], m.uy[q], m.uz[q]};
                const float qm = m.qm[q], w = m.w[q];
                const int32_t id = m.id[q];
                wf = w;
                int alive = (qm != 0.0f) && push_marker<true>(ts, v, c, r, u, w, qm, w3, err, &fs);
                if (!alive) {
                    d.qm[p] = 0.0f;
                    d.w[p] = 0.0f;
                } else {
                    d.x[p] = r[0];
                    d.y[p] = r[1];
                    d.z[p] = r[2];
                    d.ux[p] = u[0];
                    d.uy[p] = u[1];
                    d.uz[p] = u[2];
                    d.w[p] = w;
                    d.qm[p] = qm;
                    key = cell_key_of(tv, g, r);
                }
                d.id[p] = id;
                tv.key[p] = key;
            }
            first_run_add(ts, g, fs, wf);
            wave_run_count(tv.count, key);
        }
        __syncthreads();
        /* flush the tile's J block: one HBM atomic per non-zero node value */
        for (int q = threadIdx.x; q < TJN * TJN * TJN * 3; q += TPB) {
            const double val = lj[q];
            if (val == 0.0) continue;
            const int comp = q % 3, node = q / 3;
            const int cc = node % TJN, bb = (node / TJN) % TJN, aa = node / (TJN * TJN);
            const int i = ts.o0 + aa, j = ts.o1 + bb, k = ts.o2 + cc;
            if (i < -2 || j < -2 || k < -2 || i > g.n[0] + 2 || j > g.n[1] + 2 || k > g.n[2] + 2)
                continue;
            unsafeAtomicAdd(v.J[comp] + gidx(g, i, j, k), val);
        }
        __syncthreads();
    }
    block_sum3(w3, Wacc);
}

template <> struct KernelDef<Push64TArgs> {
    static constexpr int kBlock = TPB;
    static constexpr int kWaves = 3;   /* register budget for 3 wavefronts per SIMD */
    __device__ __forceinline__ static void body(const Push64TArgs &a, VBlock b, VBlock)
    {
        push64_tiled(a.m, a.d, a.v, a.c, a.tv, a.Wacc, a.err, b);
    }
};

/* Workgroups of the persistent push that fit on one CU at once. */
int push64_blocks_per_cu(void)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_one_w<Push64TArgs>, TPB, 0) !=
            hipSuccess ||
        nb < 1)
        nb = 1;
    return nb;
}

int launch_push64_tiled(hipStream_t s, const Markers64 &m, const Markers64 &dst,
                        const Fields64 &f, const Grid &g, const PushConsts &c,
                        const TileView &tv, double *Wacc, unsigned *err, int nblocks)
{
    F64View v;
    for (int a = 0; a < 3; ++a) {
        v.E[a] = f.E[a];
        v.H[a] = f.H[a];
        v.J[a] = f.J[a];
    }
    v.g = g;
    Push64TArgs a = {m, dst, v, c, tv, Wacc, err};
    return launch_def(s, a, (unsigned)nblocks);
}

/* CIC charge density of the live markers, source/core/plasma_rho.c:160-173:
 * the order's markers through its permutation (perm non-null: the first
 * nmark[parity] entries), else the first n records. */
__global__ void k_rho64(Markers64 m, long n, double *rho, Grid g, const int *perm,
                        const int *nmark, int parity)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n || (perm && q >= nmark[parity])) return;
    const long p = perm ? perm[q] : q;
    if (m.qm[p] == 0.0f) return;
    int i, j, k;
    double sx, sy, sz;
    sx = m.x[p] / g.h[0]; i = DBL_TO_INT(sx); sx -= i;
    sy = m.y[p] / g.h[1]; j = DBL_TO_INT(sy); sy -= j;
    sz = m.z[p] / g.h[2]; k = DBL_TO_INT(sz); sz -= k;
    const double val = m.w[p];
    unsafeAtomicAdd(rho + gidx(g, i, j, k), (1 - sx) * ((1 - sy) * ((1 - sz) * val)));
    unsafeAtomicAdd(rho + gidx(g, i + 1, j, k), sx * ((1 - sy) * ((1 - sz) * val)));
    unsafeAtomicAdd(rho + gidx(g, i, j + 1, k), (1 - sx) * (sy * ((1 - sz) * val)));
    unsafeAtomicAdd(rho + gidx(g, i + 1, j + 1, k), sx * (sy * ((1 - sz) * val)));
    unsafeAtomicAdd(rho + gidx(g, i, j, k + 1), (1 - sx) * ((1 - sy) * (sz * val)));
    unsafeAtomicAdd(rho + gidx(g, i + 1, j, k + 1), sx * ((1 - sy) * (sz * val)));
    unsafeAtomicAdd(rho + gidx(g, i, j + 1, k + 1), (1 - sx) * (sy * (sz * val)));
    unsafeAtomicAdd(rho + gidx(g, i + 1, j + 1, k + 1), sx * (sy * (sz * val)));
}

int launch_rho64(hipStream_t s, const Markers64 &m, long n, double *rho, const Grid &g,
                 const int *perm, const int *nmark, int parity)
{
    if (n <= 0) return MPIC_OK;
    k_rho64<<<(unsigned)((n + TPB - 1) / TPB), TPB, 0, s>>>(m, n, rho, g, perm, nmark, parity);
    MPIC_CHECK(hipGetLastError());
    return MPIC_OK;
}

/* Linear cell index of floor(x/h) for every marker of the order (through
 * its permutation when perm is non-null; -1 for deleted ones). */
__global__ void k_cells64(Markers64 m, long n, Grid g, const int *perm, const int *nmark,
                          int parity, int32_t *cell)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const long p = perm ? (q < nmark[parity] ? perm[q] : -1) : q;
    if (p < 0 || m.qm[p] == 0.0f) {
        cell[q] = -1;
        return;
    }
    int i = DBL_TO_INT(m.x[p] / g.h[0]), j = DBL_TO_INT(m.y[p] / g.h[1]),
        k = DBL_TO_INT(m.z[p] / g.h[2]);
    cell[q] = (i * g.n[1] + j) * g.n[2] + k;
}

int launch_cells64(hipStream_t s, const Markers64 &m, long n, const Grid &g, const int *perm,
                   const int *nmark, int parity, int32_t *cell)
{
    if (n <= 0) return MPIC_OK;
    k_cells64<<<(unsigned)((n + TPB - 1) / TPB), TPB, 0, s>>>(m, n, g, perm, nmark, parity, cell);
    MPIC_CHECK(hipGetLastError());
    return MPIC_OK;
}
