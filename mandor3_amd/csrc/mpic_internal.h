/*
 * mpic_internal.h -- device data layout and launch helpers of libmpic.
 *
 * HBM layout of one run (whole domain on one GPU):
 *   fields   structure-of-arrays, one array per component, on a padded grid
 *            of (n+5)^3 nodes covering node indices [-2, n+2] per axis
 *            (k fastest).  E/H of the reference carry one ghost layer and J
 *            two (source/core/main.c:109-115); one common padding keeps a
 *            single index formula for every kernel.
 *   caps     Hc/Ec: H^{n+1/2} and E^{n+1} of the boundary layer, computed
 *            ahead of the interior exactly as the reference's boundary caps
 *            (source/core/em_caps.c:405-725), stored on the same padded grid.
 *   markers  structure-of-arrays (FP64 plain form): x y z ux uy uz (double),
 *            w qm (float), id (int32).  A deleted marker has qm == 0.
 */
#ifndef MPIC_INTERNAL_H
#define MPIC_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpic.h"

#define MPIC_GHOST 2
#define MPIC_XCDS 8          /* MI355X: 8 XCDs, workgroups dispatched round-robin */

/* Workgroup L of 'total' as the hardware places it (XCD L % 8, its
 * (L / 8)-th there) mapped to a logical index such that every XCD takes one
 * contiguous range: neighbouring work then shares the XCD's L2. */
__device__ static inline unsigned xcd_block(unsigned L, unsigned total)
{
    const unsigned x = L % MPIC_XCDS, q = total / MPIC_XCDS, r = total % MPIC_XCDS;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + L / MPIC_XCDS;
}

struct Grid {
    int    n[3];
    int    GY, GZ;          /* padded extents along y and z */
    long   ncell;           /* padded node count */
    double h[3];
    double tau;
};

__host__ __device__ static inline long gidx(const Grid &g, int i, int j, int k)
{
    return ((long)(i + MPIC_GHOST) * g.GY + (j + MPIC_GHOST)) * g.GZ + (k + MPIC_GHOST);
}

struct Box {
    int lo[3], hi[3];
};

/* Up to 6 disjoint boxes enumerated by one launch. */
struct BoxSet {
    int  n;
    Box  b[6];
    long off[7];            /* prefix sums of the box volumes */
    int  shell;             /* >= 0: the boxes are exactly the nodes of [0, n+1]^3 outside
                             * [shell, n-2]^3 (a flush set; tested by ranges), else -1 */
};

/* One laser's soft-source plane (source/scpic/em_mirror.c:667-693): the
 * nodes (pos, j, k), j in [jmin, jmax], k in [kmin, kmax], of its Fy/Fz
 * pair get A * (map . {cos, sin}) added, A and the carrier's cos and sin of
 * the step from tab[step % cap]; map rows are numbered from (JMIN, KMIN)
 * over [KMIN, KMAX].  on = 0: no source. */
struct LaserPlane {
    const double *map = nullptr, *tab = nullptr;
    const int    *step = nullptr;
    int           on = 0, pos = 0, JMIN = 0, KMIN = 0, KMAX = 0, jmin = 0, jmax = -1, kmin = 0,
                  kmax = -1, cap = 1;
};

/* Field arrays of one precision; J is always accumulated in FP64. */
template <class T>
struct FieldsT {
    T      *E[3], *H[3];
    double *J[3];
    T      *Hc[3], *Ec[3];
};
typedef FieldsT<double> Fields64;
typedef FieldsT<float>  Fields32;

/* TFSF interface faces of one kind (E or H), source/core/em_TFSF.c:64-76:
 * the field component, the normal axis ea and tangential axes ep (outer) /
 * eq (inner) of the played frames, the node box, and the offset of the
 * face's frames in the device data (off < 0: face opened or absent). */
struct TfsfFace {
    int  comp, ea, ep, eq;
    int  lo[3], hi[3];
    int  dQ;
    long vol, off;
};
struct TfsfSet {
    TfsfFace f[12];
    int      lo_in;               /* sources act on nodes [lo_in, n-2] per axis */
};

struct JPtrs {
    double *J[3];
};

struct Markers64 {
    double  *x, *y, *z, *ux, *uy, *uz;
    float   *w, *qm;
    int32_t *id;
};

/* One wall of the boundary layer, applied to a 3-component array set. */
enum FaceKind {
    FACE_PERIODIC_H = 0,    /* X(0) = X(N), X(N+1) = X(1) */
    FACE_PERIODIC_E = 1,    /* X(N) = X(0), X(N+1) = X(1) */
    FACE_WALL_H     = 2,    /* mirror (zero normal) / Mur: tangential copied out */
    FACE_MIRROR_E   = 3,
    FACE_MUR_E      = 4,
    FACE_MUR_INIT   = 5
};

struct FaceOp {
    int     kind, ea, top, N, zeroN, fp32;
    int     plo, phi, qlo, qhi;  /* transverse extents along ep = (ea+1)%3, eq = (ea+2)%3 */
    double  alpha;
    void   *buf;                 /* Mur buffer, field precision */
    void   *X[3];                /* field arrays, field precision */
};

/* FP32 split-precision markers (see mpic_push32.hip). */
/* FP32 split markers: two 16-byte records per marker, each moved by one
 * 128-bit load or store, and the id:
 *   pos = (cell word as float bits, ox, oy, oz): cell word = (species << 29)
 *         | tile-major cell key, -1 = deleted; offsets in cell units
 *   mom = (ux, uy, uz, w) */
struct Markers32 {
    float4  *pos;
    float4  *mom;
    int32_t *id;
};

struct PushConsts {
    double h[3], tau;
    double L[3];
    double shell1[3], shell2[3];
    int    bc[6];
};

#define MPIC_CHECK(call)                                                    \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) {                                             \
            mpic_set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,      \
                           hipGetErrorString(e_));                          \
            return MPIC_ERR_DEVICE;                                         \
        }                                                                   \
    } while (0)

void mpic_set_error(const char *fmt, ...);

/* Launchers (mpic_fdtd.hip, mpic_push64.hip, mpic_push32.hip). */
template <class T> int launch_cap_H(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &bs);
template <class T> int launch_cap_E(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &bs);
int launch_face(hipStream_t s, const Grid &g, const FaceOp &op);
int launch_face_pair(hipStream_t s, const Grid &g, const FaceOp &a, const FaceOp &b);
int launch_face_planes(hipStream_t s, const Grid &g, const FaceOp *ops, int nop);
template <class T>
int launch_h_update_jclear(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &flush,
                           int full, double *part, double *J, double *acc, int *range,
                           const int *step, const int *active, const int *nactive, int tpp,
                           int jfull);
template <class T>
int launch_h_update_jfold(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &flush,
                          int full, const LaserPlane &lp, const JPtrs &jp, const int bc[6],
                          const int *jr);
template <class T>
int launch_h_update(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &flush,
                    int full, double *part = nullptr, const LaserPlane &lp = LaserPlane());
template <class T>
int launch_e_fused(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &flush,
                   const int *jr, const LaserPlane &lp = LaserPlane());
template <class T>
int launch_e_update(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &flush,
                    int interior, int do_flush);
template <class T>
int launch_e_tauJ(hipStream_t s, const FieldsT<T> &f, const Grid &g, int *tick, const int *jr);
template <class T>
int launch_laser(hipStream_t s, T *Fy, T *Fz, const Grid &g, const LaserPlane &lp);
int launch_tick(hipStream_t s, int *step);
template <class T>
int launch_tfsf(hipStream_t s, T *const F[3], const Grid &g, const TfsfSet &fs, long maxvol,
                const double *data, const int *tab, const int *step, int cap, int slot,
                double sign);
template <class T>
int launch_spot(hipStream_t s, T *Ey, T *Ez, const Grid &g, int i, const double *py,
                const double *pz, const double *tab, const int *step, int cap);
int launch_jfold(hipStream_t s, const JPtrs &f, const Grid &g, int ea, int bc0, int bc1,
                 const int *jr);
int launch_jfold_planes(hipStream_t s, const JPtrs &f, const Grid &g, const int bc[6],
                        const int *jr);
/* VSP (beta, alpha, beta) sweeps along z, y, x of ncomp arrays on the padded grid. */
int launch_vsp(hipStream_t s, double *const F[3], int ncomp, const Grid &g, double alpha);
template <class T>
int launch_node_fields(hipStream_t s, T *const F[3], const Grid &g, int kind, double *out);
template <class T>
int launch_tfsf_record(hipStream_t s, T *const E[3], T *const H[3], const Grid &g,
                       const TfsfSet &fe, const TfsfSet &fh, long maxvol, const double *factor,
                       double *out, const int *step, int step0, int cap);
template <class T>
int launch_probes(hipStream_t s, T *const E[3], T *const H[3], const Grid &g, const int *ijk,
                  int np, double *out);
template <class T> int launch_energy(hipStream_t s, const FieldsT<T> &f, const Grid &g, double *acc);
/* Loop-body diagnostics: the H update writes em_energy partial sums (two
 * doubles per workgroup, h_update_blocks(g, fp32) workgroups) when given
 * 'part'; the ring row {W_E sum, W_M sum, WTx, WTy, WTz} of the device step
 * counter's step sums them, and with 'tick' advances the counter. */
int h_update_blocks(const Grid &g, int fp32);
int launch_diag_ring(hipStream_t s, const double *part, int nparts, const double *wk,
                     double *ring, int *step, int cap, int tick);
int launch_zero(hipStream_t s, double *p, long n);
int launch_jclear(hipStream_t s, double *J, const Grid &g, double *acc, int *range,
                  const int *step, const int *active, const int *nactive, int tpp, int full);

int launch_push64(hipStream_t s, const Markers64 &m, long n, const Fields64 &f,
                  const Grid &g, const PushConsts &c, double *Wacc, int mode, unsigned *err);
int launch_rho64(hipStream_t s, const Markers64 &m, long n, double *rho,
                 const Grid &g, const int *perm, const int *nmark, int parity);
int launch_rho_fold(hipStream_t s, double *rho, const Grid &g, int b, int bc);
template <class T>
int launch_gauss(hipStream_t s, const FieldsT<T> &f, const double *rho1, const double *rho2,
                 const Grid &g, double *acc);
int launch_cells64(hipStream_t s, const Markers64 &m, long n, const Grid &g, const int *perm,
                   const int *nmark, int parity, int32_t *cell);

#endif
