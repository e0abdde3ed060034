/*
 * mpic_run.hip -- run state and the C ABI of include/mpic.h (FP64 plain form).
 *
 * Stage order and the per-step call sequence follow source/core/main.c:
 * 353-476; the emStage order check follows source/core/em.c:34,88-219.
 */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "mpic_internal.h"
#include "mpic_launch.h"
#include "mpic_tiles.h"

thread_local Recorder *g_recorder = nullptr;

static thread_local char g_err[512] = "";

void mpic_set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

extern "C" const char *mpic_last_error(void) { return g_err; }

extern "C" int mpic_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

#define TRY(x)                        \
    do {                              \
        int rc_ = (x);                \
        if (rc_ != MPIC_OK) return rc_; \
    } while (0)

static int scratch_setup(int dev);

/* Calls a field launcher with the run's field set (FP64 or FP32). */
#define FCALL(fn, ...)                                                   \
    do {                                                                 \
        if (r->fp32)                                                     \
            TRY(fn(r->stream, r->f32, __VA_ARGS__));                     \
        else                                                             \
            TRY(fn(r->stream, r->f, __VA_ARGS__));                       \
    } while (0)

/* Steps covered by one fill of the laser tables. */
#define LASER_CAP 1024
/* Steps the diagnostics ring holds. */
#define DIAG_RING 1024

struct Laser {
    double  A, T0, delay, dt, phi0;
    int     forward, plane;
    double *dmap;             /* device: E map then H map */
    long    nodes;
    double *dtab;             /* device: per step {A, cos, sin}, H table then E table */
    double *htab;             /* pinned staging of dtab */
};

/* Hard Gaussian spot source (source/core/em_gaussSpot.c:20-156). */
struct Spot {
    int     on, i;
    double  omega, front, duration, plateau, Ey, Ez;
    double *dprof;            /* device: profiles of Ey then Ez, (n1+3) x (n2+3) */
    double *dtab;             /* device: per step {Ey amplitude, Ez amplitude} */
    double *htab;             /* pinned staging of dtab */
};

/* TFSF interface playback: the whole recording lives in HBM; the frame of
 * each step comes from a host-filled table like the laser's. */
struct Tfsf {
    int     on, mode, last, frame0;
    long    step0;                 /* run step at which the frame counter was frame0 */
    TfsfSet fe, fh;                /* E faces (inner nodes from 2), H faces (from 3) */
    long    maxvol;
    double *ddata;
    int    *dtab;                  /* device: per step {H frame, E frame}, -1 off */
    int    *htab;                  /* pinned staging of dtab */
};

struct TfsfRec {
    int     on;
    long    step0;                 /* run step of frame 0 */
    int     cap;                   /* frames the device buffer holds */
    TfsfSet fe, fh;                /* faces; off = offset of the face's frames */
    long    maxvol, total;         /* nodes of the largest face; doubles per frame */
    double *dbuf;                  /* [slot][frame][node], E slots then H slots */
    double *dfac;                  /* factor per slot f: (1-2 top)(1-2 comp) tau/h[axis] */
};

struct CapBoxes {             /* one field's six caps */
    Box upd[6], fl[6];        /* original boxes */
    Box upd_s[6], fl_s[6];    /* after overlap removal */
};

struct mpic_run {
    mpic_config_t cfg;
    Grid          g;
    int           dev;
    hipStream_t   stream;
    int           fp32;       /* MPIC_FP32_SPLIT run */
    Fields64      f;          /* FP64 fields (plain form) */
    Fields32      f32;        /* FP32 fields (split form); J is FP64 in both */
    JPtrs         jp;
    void         *fblock;     /* all field arrays */
    Markers32     m32, m32buf[2];
    float         qm[8];      /* q/m per species of the split form */
    int           nspecies;
    double       *rho1, *rho2;
    Markers64     m;          /* current marker order */
    Markers64     mbuf[2];    /* double buffer of the per-step sort */
    int           cur;        /* m == mbuf[cur] */
    long          n, cap;     /* n: markers loaded (upper bound of the live count) */
    int           tiled;      /* 1: sorted tiled push (default), 0: untiled push */
    TileView      tv;
    unsigned long long *scanw;  /* the tile scan's published words and sequence number */
    int           parity;     /* tv.nmark[parity] = markers in the current order */
    int           push_blocks;  /* FP64: resident workgroups of the persistent push */
    int           tile_blocks;  /* FP32: workgroups of the tile push (work items taken in turn) */
    int           tile_waves;   /* FP32: wavefronts per workgroup of the tile push */
    int           gather_blocks;/* FP32: workgroups of the J gather */
    long          items_cap;    /* work items (and FP32 J blocks) the buffers hold */
    int           ncu;          /* compute units of the device */
    float        *tj;           /* FP32: per-tile J blocks, [ntiles][TJB] */
    CapBoxes      cH, cE;
    BoxSet        updH, updE, flH, flE;
    std::vector<FaceOp> bcH, bcE;
    std::vector<double *> murbuf;
    std::vector<Laser> lasers;
    PushConsts    pc;
    double       *dacc;       /* device scratch: [0..1] energy, [2..4] W, [5..6] gauss */
    double       *hacc;       /* pinned host mirror */
    unsigned     *derr;       /* device error flags (bit 0: marker off the grid, bit 1: a
                                 tile scan's wait for an earlier workgroup ran out) */
    int           stage;
    int           inited;
    int           timed;      /* >= 0: index of the timed step of mpic_advance */
    double        time;
    std::vector<hipEvent_t> evpool;  /* 5 events per timed step */
    int          *dstep;      /* device step counter (indexes the laser tables) */
    int          *djr;        /* device: x plane ranges of J written: by step parity [2][2],
                                 then the current step's [2] */
    Spot          spot;
    Tfsf          tfsf;
    TfsfRec       trec;
    long          step;       /* host copy of the step counter */
    long          tab_lo, tab_hi;  /* steps the laser tables hold */
    int           use_graphs; /* mpic_advance replays captured steps */
    int           graphs_ok;  /* exec[] match the current launch topology */
    hipGraphExec_t exec[2];   /* one step from each of the two marker-buffer parities */
    int           exec_parity;     /* parity at which exec[0] was captured */
    int           ev_steps;   /* timed steps of the last mpic_advance */
    double       *dnode;      /* node values of the Tecplot records, 3 (n+1)^3 doubles */
    int          *dprobe;     /* probe nodes, 3 ints each */
    double       *dsample;    /* probe samples, 6 doubles each */
    int           nprobes;
    /* diagnostics of the loop body (em_energy, plasma_temperature), one row
     * per step in a device ring indexed by the device step counter */
    double       *dpart;      /* energy partial sums, 2 per workgroup */
    int           nparts;
    double       *dring;      /* DIAG_RING rows {W_E sum, W_M sum, WTx, WTy, WTz} */
    std::vector<long>   ring_step;  /* step whose row a slot holds (-1: none) */
    std::vector<double> ring_time;  /* Time at the start of that step */
    int           capturing;  /* step_body is being recorded into graphs, not run */
    int           jclear_done;  /* the H half step cleared J and the temperature sums */
    double        vsp;        /* VSP convolution weight alpha (1: off) */
    long          topo;       /* bumped whenever the launch sequence of a step changes */
};

/* The launch sequence of the run's step changed: captured graphs and fused
 * ensembles holding it are stale. */
static void invalidate(mpic_run *r)
{
    r->graphs_ok = 0;
    ++r->topo;
}

/* ---------------- boundary-cap geometry (single rank) ---------------- */
/* Tables of source/core/em_caps.c:68-113 indexed [bc][top][lo/hi]. */
static const int kUpdE[3][2][2] = {{{0, 1}, {-1, -1}}, {{1, 1}, {-1, 0}}, {{1, 1}, {-1, 0}}};
static const int kFlE[3][2][2]  = {{{0, 1}, {-1, +1}}, {{0, 1}, {-1, +1}}, {{0, 1}, {-1, +1}}};
static const int kUpdH[3][2][2] = {{{0, 2}, {-1, 0}}, {{1, 2}, {-1, 0}}, {{0, 2}, {-1, 0}}};
static const int kFlH[3][2][2]  = {{{0, 2}, {-1, +1}}, {{0, 2}, {-1, +1}}, {{0, 2}, {-1, +1}}};

static void cap_geometry(const int n[3], const int bc[6], const int (*U)[2][2],
                         const int (*F)[2][2], CapBoxes &cb)
{
    const int side[2][3] = {{0, 0, 0}, {n[0], n[1], n[2]}};
    for (int b = 0; b < 6; ++b) {
        const int ea = b >> 1, top = b & 1;
        const int bn = bc[3 * top + ea];
        Box &u = cb.upd[b], &f = cb.fl[b];
        for (int a = 0; a < 3; ++a) {
            if (a == ea) {
                u.lo[a] = side[top][a] + U[bn][top][0];
                u.hi[a] = side[top][a] + U[bn][top][1];
                f.lo[a] = side[top][a] + F[bn][top][0];
                f.hi[a] = side[top][a] + F[bn][top][1];
            } else {
                u.lo[a] = side[0][a] + U[bc[a]][0][0];
                u.hi[a] = side[1][a] + U[bc[3 + a]][1][1];
                f.lo[a] = side[0][a] + F[bc[a]][0][0];
                f.hi[a] = side[1][a] + F[bc[3 + a]][1][1];
            }
        }
        cb.upd_s[b] = u;
        cb.fl_s[b] = f;
    }
    /* Overlap removal of cap_2_shrink (source/core/em_caps.c:205-259): a cap
     * gives up the rows already owned by an earlier cap of another axis. */
    for (int b = 2; b < 6; ++b) {
        const int ea = b >> 1, ax[2] = {(ea + 1) % 3, (ea + 2) % 3};
        for (int rd = 0; rd < b; ++rd) {
            if ((rd >> 1) == ea) continue;
            for (int pass = 0; pass < 2; ++pass) {
                Box &c = pass ? cb.fl_s[b] : cb.upd_s[b];
                const Box &k = pass ? cb.fl_s[rd] : cb.upd_s[rd];
                for (int t = 0; t < 2; ++t) {
                    const int a = ax[t];
                    if (c.hi[a] == k.hi[a] && c.lo[a] < k.lo[a]) c.hi[a] = k.lo[a] - 1;
                    if (c.lo[a] == k.lo[a] && c.hi[a] > k.hi[a]) c.lo[a] = k.hi[a] + 1;
                }
            }
        }
    }
}

static BoxSet make_set(const Box *bx)
{
    BoxSet s;
    s.n = 0;
    s.shell = -1;
    s.off[0] = 0;
    for (int b = 0; b < 6; ++b) {
        long v = 1;
        for (int a = 0; a < 3; ++a) {
            const int d = bx[b].hi[a] - bx[b].lo[a] + 1;
            v *= (d > 0) ? d : 0;
        }
        if (v <= 0) continue;
        s.b[s.n] = bx[b];
        s.off[s.n + 1] = s.off[s.n] + v;
        s.n++;
    }
    return s;
}

/* lo when the boxes of s are pairwise disjoint, lie in [0, n+1]^3, miss
 * [lo, n-2]^3 and together hold every other node of [0, n+1]^3 (so a node
 * of [0, n+1]^3 is in s exactly when it is outside [lo, n-2]^3), else -1. */
static int complement_shell(const BoxSet &s, const int n[3], int lo)
{
    long vol = 0;
    for (int b = 0; b < s.n; ++b) {
        const Box &B = s.b[b];
        bool hits_inner = true;
        for (int a = 0; a < 3; ++a) {
            if (B.lo[a] < 0 || B.hi[a] > n[a] + 1) return -1;
            if (B.hi[a] < lo || B.lo[a] > n[a] - 2) hits_inner = false;
        }
        if (hits_inner) return -1;
        for (int c = 0; c < b; ++c) {
            const Box &C = s.b[c];
            bool meet = true;
            for (int a = 0; a < 3; ++a)
                if (B.hi[a] < C.lo[a] || C.hi[a] < B.lo[a]) meet = false;
            if (meet) return -1;
        }
        vol += s.off[b + 1] - s.off[b];
    }
    long all = 1, inner = 1;
    for (int a = 0; a < 3; ++a) {
        all *= n[a] + 2;
        inner *= (n[a] - 2 - lo + 1 > 0) ? n[a] - 2 - lo + 1 : 0;
    }
    return vol == all - inner ? lo : -1;
}

template <class T>
static FaceOp face(int kind, int ea, int top, const Box &ext, T *const X[3], int N)
{
    FaceOp op;
    memset(&op, 0, sizeof(op));
    op.kind = kind;
    op.ea = ea;
    op.top = top;
    op.N = N;
    const int ep = (ea + 1) % 3, eq = (ea + 2) % 3;
    op.plo = ext.lo[ep];
    op.phi = ext.hi[ep];
    op.qlo = ext.lo[eq];
    op.qhi = ext.hi[eq];
    for (int c = 0; c < 3; ++c) op.X[c] = (void *)X[c];
    op.fp32 = (sizeof(T) == 4);
    return op;
}

/* Wall sequences of cap_cacheH / cap_cacheEH (source/core/em_caps.c:470-553,
 * 631-721): x, y, z; low side then high side. */
template <class T>
static int build_wall_ops_t(mpic_run *r, T *const Hc[3], T *const Ec[3])
{
    const int *bc = r->cfg.bc;
    r->bcH.clear();
    r->bcE.clear();
    for (int ea = 0; ea < 3; ++ea) {
        const int N = r->g.n[ea];
        for (int top = 0; top < 2; ++top) {
            const int b = bc[3 * top + ea];
            const Box &eH = r->cH.fl[(ea << 1) | top], &eE = r->cE.fl[(ea << 1) | top];
            if (b == MPIC_BC_PERIODIC) {
                if (top) continue;
                r->bcH.push_back(face(FACE_PERIODIC_H, ea, 0, eH, Hc, N));
                r->bcE.push_back(face(FACE_PERIODIC_E, ea, 0, eE, Ec, N));
            } else if (b == MPIC_BC_MIRROR) {
                FaceOp h = face(FACE_WALL_H, ea, top, eH, Hc, N);
                h.zeroN = 1;
                r->bcH.push_back(h);
                r->bcE.push_back(face(FACE_MIRROR_E, ea, top, eE, Ec, N));
            } else if (b == MPIC_BC_OPEN) {
                r->bcH.push_back(face(FACE_WALL_H, ea, top, eH, Hc, N));
                FaceOp e = face(FACE_MUR_E, ea, top, eE, Ec, N);
                const long nodes = (long)(e.phi - e.plo + 1) * (e.qhi - e.qlo + 1);
                double *buf = nullptr;
                MPIC_CHECK(hipMalloc(&buf, sizeof(T) * 4 * nodes));
                r->murbuf.push_back(buf);
                e.buf = buf;
                e.alpha = (r->g.tau - r->g.h[ea]) / (r->g.tau + r->g.h[ea]);
                r->bcE.push_back(e);
            } else {
                mpic_set_error("unknown boundary code %d", b);
                return MPIC_ERR_ARG;
            }
        }
    }
    return MPIC_OK;
}

static int build_wall_ops(mpic_run *r)
{
    return r->fp32 ? build_wall_ops_t<float>(r, r->f32.Hc, r->f32.Ec)
                   : build_wall_ops_t<double>(r, r->f.Hc, r->f.Ec);
}

/* ---------------- lifecycle ---------------- */
/* Row pitch along k: n + 5 nodes rounded up to a multiple of 4, so that
 * every row starts 16-byte aligned (the volume kernels move 16 bytes per
 * access). */
static int row_pitch(int n2) { return (n2 + 5 + 3) & ~3; }

static size_t field_bytes(const int n[3], int fp32)
{
    const size_t cells = (size_t)(n[0] + 5) * (n[1] + 5) * row_pitch(n[2]);
    /* E H Hc Ec in the field precision, J in FP64 */
    return (fp32 ? sizeof(float) : sizeof(double)) * 12 * cells + sizeof(double) * 3 * cells;
}

/* Markers per work item of the FP32 push (a tile's markers are split
 * evenly into items of at most this many; 0 = one item per tile, the
 * default: on the laser-foil workload splitting measured slower at every
 * item size from 512 to 1728).  MPIC_ITEM_MARKERS overrides. */
static int push32_item_markers(void)
{
    const char *e = getenv("MPIC_ITEM_MARKERS");
    const int v = e ? atoi(e) : 0;
    return v >= 64 ? v : 0;
}

/* Work items of a run: at most one per tile plus one per chm markers. */
static long items_capacity(int ntiles, long cap, int chm)
{
    return (long)ntiles + (chm > 0 ? cap / chm : 0) + 1;
}

extern "C" int mpic_footprint(const mpic_config_t *cfg, size_t *bytes)
{
    if (!cfg || !bytes) return MPIC_ERR_ARG;
    const size_t cells = (size_t)(cfg->n[0] + 5) * (cfg->n[1] + 5) * row_pitch(cfg->n[2]);
    const size_t tiles = (size_t)((cfg->n[0] + 3) / 4) * ((cfg->n[1] + 3) / 4) * ((cfg->n[2] + 3) / 4);
    /* per marker and buffer: records and id; plus the new key (and, FP32, the
     * order's index) once */
    const size_t mk = (cfg->precision == MPIC_FP32_SPLIT) ? (4 + 7 * 4 + 4 + 4) : (6 * 8 + 2 * 4 + 4 + 4);
    const int fp32 = cfg->precision == MPIC_FP32_SPLIT;
    const long cap = cfg->max_markers > 0 ? cfg->max_markers : 1;
    const size_t items = (size_t)items_capacity((int)tiles, cap, fp32 ? push32_item_markers() : 0);
    *bytes = field_bytes(cfg->n, fp32) + 2 * sizeof(double) * cells +
             2 * (size_t)cap * mk + tiles * 64 * 3 * sizeof(int) + tiles * 6 * sizeof(int) +
             items * sizeof(int) + (fp32 ? items * TJB * sizeof(float) : 0);
    return MPIC_OK;
}

extern "C" int mpic_create(const mpic_config_t *cfg, mpic_run_t **out)
{
    if (!cfg || !out) {
        mpic_set_error("mpic_create: null argument");
        return MPIC_ERR_ARG;
    }
    for (int a = 0; a < 3; ++a)
        if (cfg->n[a] < 7) {
            mpic_set_error("mpic_create: %d cells along axis %d; the reference needs >= 7 "
                           "(core region, source/core/plasma.c:404-407)", cfg->n[a], a);
            return MPIC_ERR_ARG;
        }
    if (!(1.0 / (cfg->tau * cfg->tau) > 1.0 / (cfg->h[0] * cfg->h[0]) +
                                            1.0 / (cfg->h[1] * cfg->h[1]) +
                                            1.0 / (cfg->h[2] * cfg->h[2]))) {
        mpic_set_error("mpic_create: Yee solver is unstable (Courant)");
        return MPIC_ERR_ARG;
    }
    if (cfg->precision != MPIC_FP64 && cfg->precision != MPIC_FP32_SPLIT) {
        mpic_set_error("mpic_create: unknown precision %d", cfg->precision);
        return MPIC_ERR_ARG;
    }
    if (cfg->precision == MPIC_FP32_SPLIT) {
        /* the split form's cell word keeps 29 bits for the cell */
        const long cells = (long)((cfg->n[0] + 3) / 4) * ((cfg->n[1] + 3) / 4) *
                           ((cfg->n[2] + 3) / 4) * 64;
        if (cells > (1L << 29)) {
            mpic_set_error("mpic_create: the FP32 split form holds at most 2^29 cells (in 4^3 "
                           "tiles)");
            return MPIC_ERR_ARG;
        }
    }
    /* marker offsets and counts are int32 */
    if (cfg->max_markers >= (1L << 31)) {
        mpic_set_error("mpic_create: capacity %ld exceeds 2^31 - 1 markers", cfg->max_markers);
        return MPIC_ERR_ARG;
    }
    MPIC_CHECK(hipSetDevice(cfg->device));
    TRY(scratch_setup(cfg->device));
    mpic_run *r = new mpic_run();
    r->cfg = *cfg;
    r->dev = cfg->device;
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
    Grid &g = r->g;
    for (int a = 0; a < 3; ++a) {
        g.n[a] = cfg->n[a];
        g.h[a] = cfg->h[a];
    }
    g.tau = cfg->tau;
    g.GY = cfg->n[1] + 5;
    g.GZ = row_pitch(cfg->n[2]);
    g.ncell = (long)(cfg->n[0] + 5) * g.GY * g.GZ;
    r->fp32 = (cfg->precision == MPIC_FP32_SPLIT);
    const size_t fb = field_bytes(cfg->n, r->fp32);
    MPIC_CHECK(hipMalloc(&r->fblock, fb));
    /* NaN fill of E/H and the caps, as the reference's NaN-filled meshes;
     * J starts at zero. */
    MPIC_CHECK(hipMemsetAsync(r->fblock, 0xff, fb, r->stream));
    double *jb = (double *)r->fblock;
    if (r->fp32) {
        float *p = (float *)r->fblock;
        for (int c = 0; c < 3; ++c) r->f32.E[c] = p + (0 + c) * g.ncell;
        for (int c = 0; c < 3; ++c) r->f32.H[c] = p + (3 + c) * g.ncell;
        for (int c = 0; c < 3; ++c) r->f32.Hc[c] = p + (6 + c) * g.ncell;
        for (int c = 0; c < 3; ++c) r->f32.Ec[c] = p + (9 + c) * g.ncell;
        jb = (double *)(p + 12 * g.ncell);
        for (int c = 0; c < 3; ++c) r->f32.J[c] = jb + c * g.ncell;
    } else {
        double *p = (double *)r->fblock;
        for (int c = 0; c < 3; ++c) r->f.E[c] = p + (0 + c) * g.ncell;
        for (int c = 0; c < 3; ++c) r->f.H[c] = p + (3 + c) * g.ncell;
        for (int c = 0; c < 3; ++c) r->f.Hc[c] = p + (6 + c) * g.ncell;
        for (int c = 0; c < 3; ++c) r->f.Ec[c] = p + (9 + c) * g.ncell;
        jb = p + 12 * g.ncell;
        for (int c = 0; c < 3; ++c) r->f.J[c] = jb + c * g.ncell;
    }
    for (int c = 0; c < 3; ++c) r->jp.J[c] = jb + c * g.ncell;
    MPIC_CHECK(hipMemsetAsync(jb, 0, sizeof(double) * 3 * g.ncell, r->stream));
    MPIC_CHECK(hipMalloc(&r->rho1, sizeof(double) * g.ncell));
    MPIC_CHECK(hipMalloc(&r->rho2, sizeof(double) * g.ncell));
    r->cap = cfg->max_markers > 0 ? cfg->max_markers : 1;
    for (int b = 0; b < 2 && r->fp32; ++b) {
        Markers32 &mb = r->m32buf[b];
        MPIC_CHECK(hipMalloc(&mb.pos, sizeof(float4) * r->cap));
        MPIC_CHECK(hipMalloc(&mb.mom, sizeof(float4) * r->cap));
        MPIC_CHECK(hipMalloc(&mb.id, sizeof(int32_t) * r->cap));
    }
    r->m32 = r->m32buf[0];
    for (int b = 0; b < 2 && !r->fp32; ++b) {
        Markers64 &mb = r->mbuf[b];
        MPIC_CHECK(hipMalloc(&mb.x, sizeof(double) * r->cap));
        MPIC_CHECK(hipMalloc(&mb.y, sizeof(double) * r->cap));
        MPIC_CHECK(hipMalloc(&mb.z, sizeof(double) * r->cap));
        MPIC_CHECK(hipMalloc(&mb.ux, sizeof(double) * r->cap));
        MPIC_CHECK(hipMalloc(&mb.uy, sizeof(double) * r->cap));
        MPIC_CHECK(hipMalloc(&mb.uz, sizeof(double) * r->cap));
        MPIC_CHECK(hipMalloc(&mb.w, sizeof(float) * r->cap));
        MPIC_CHECK(hipMalloc(&mb.qm, sizeof(float) * r->cap));
        MPIC_CHECK(hipMalloc(&mb.id, sizeof(int32_t) * r->cap));
    }
    r->cur = 0;
    r->m = r->mbuf[0];
    {
        const char *e = getenv("MPIC_UNTILED");
        r->tiled = r->fp32 || !(e && e[0] == '1');
    }
    TileView &tv = r->tv;
    tv.ntiles = 1;
    for (int a = 0; a < 3; ++a) {
        tv.nt[a] = (cfg->n[a] + 3) / 4;
        tv.ntiles *= tv.nt[a];
    }
    tv.ncellp = (long)tv.ntiles * TILE_CELLS;
    MPIC_CHECK(hipMalloc(&tv.off, sizeof(int) * (tv.ncellp + 1)));
    MPIC_CHECK(hipMalloc(&tv.cursor, sizeof(int) * tv.ncellp));
    MPIC_CHECK(hipMalloc(&tv.toff, sizeof(int) * ((size_t)tv.ntiles + 1)));
    MPIC_CHECK(hipMemsetAsync(tv.toff, 0, sizeof(int) * ((size_t)tv.ntiles + 1), r->stream));
    MPIC_CHECK(hipMalloc(&tv.count, sizeof(int) * tv.ncellp));
    MPIC_CHECK(hipMemsetAsync(tv.count, 0, sizeof(int) * tv.ncellp, r->stream));
    MPIC_CHECK(hipMemsetAsync(tv.off, 0, sizeof(int) * (tv.ncellp + 1), r->stream));
    MPIC_CHECK(hipMalloc(&tv.key, sizeof(int) * r->cap));
    tv.perm = nullptr;
    if (r->tiled) MPIC_CHECK(hipMalloc(&tv.perm, sizeof(int) * r->cap));
    MPIC_CHECK(hipMalloc(&tv.active, sizeof(int) * tv.ntiles));
    MPIC_CHECK(hipMalloc(&tv.slot, sizeof(int) * tv.ntiles));
    tv.chm = r->fp32 ? push32_item_markers() : 0;
    r->items_cap = items_capacity(tv.ntiles, r->cap, tv.chm);
    MPIC_CHECK(hipMalloc(&tv.item, sizeof(int) * r->items_cap));
    MPIC_CHECK(hipMalloc(&tv.istart, sizeof(int) * ((size_t)tv.ntiles + 1)));
    if (r->fp32) MPIC_CHECK(hipMalloc(&r->tj, sizeof(float) * TJB * (size_t)r->items_cap));
    MPIC_CHECK(hipMalloc(&tv.nactive, sizeof(int) * (4 + MPIC_XCDS)));
    tv.queue = tv.nactive + 1;
    tv.nmark = tv.nactive + 1 + MPIC_XCDS;
    tv.nitems = tv.nactive + 3 + MPIC_XCDS;
    MPIC_CHECK(hipMemsetAsync(tv.nactive, 0, sizeof(int) * (4 + MPIC_XCDS), r->stream));
    /* the tile scan's published words: zero never matches a scan's tag */
    MPIC_CHECK(hipMalloc(&r->scanw, sizeof(unsigned long long) * tile_scan_words(tv.ntiles)));
    MPIC_CHECK(hipMemsetAsync(r->scanw, 0, sizeof(unsigned long long) * tile_scan_words(tv.ntiles),
                              r->stream));
    r->parity = 0;
    {
        int ncu = 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, r->dev) == hipSuccess) ncu = prop.multiProcessorCount;
        /* exactly the resident workgroups of the persistent push (extra ones
         * only arrive once the tile queues are drained), and no more than
         * the tiles: a workgroup pushes whole tiles */
        r->push_blocks = ncu * push64_blocks_per_cu();
        if (r->push_blocks > tv.ntiles) r->push_blocks = tv.ntiles;
        /* the resident wavefronts of the tile push (16 per CU), one per
         * workgroup; a workgroup takes work items with stride = grid size */
        r->tile_blocks = 16 * ncu < r->items_cap ? 16 * ncu : (int)r->items_cap;
        r->tile_waves = 1;
        r->gather_blocks = 16 * ncu;
        r->ncu = ncu;
        const char *e = getenv("MPIC_PUSH_BLOCKS");
        if (e && atoi(e) > 0) r->push_blocks = r->tile_blocks = atoi(e);
        const char *w = getenv("MPIC_TILE_WAVES");
        if (w && (atoi(w) == 2 || atoi(w) == 4)) r->tile_waves = atoi(w);
    }
    r->n = 0;
    MPIC_CHECK(hipMalloc(&r->dacc, sizeof(double) * 8));
    MPIC_CHECK(hipHostMalloc(&r->hacc, sizeof(double) * 8, hipHostMallocDefault));
    MPIC_CHECK(hipMalloc(&r->derr, sizeof(unsigned)));
    MPIC_CHECK(hipMemsetAsync(r->derr, 0, sizeof(unsigned), r->stream));
    r->timed = -1;
    MPIC_CHECK(hipMalloc(&r->dstep, sizeof(int)));
    MPIC_CHECK(hipMemsetAsync(r->dstep, 0, sizeof(int), r->stream));
    {
        /* the first J clear clears the whole grid */
        const int full[6] = {-2, g.n[0] + 2, -2, g.n[0] + 2, -2, g.n[0] + 2};
        MPIC_CHECK(hipMalloc(&r->djr, sizeof(full)));
        MPIC_CHECK(hipMemcpy(r->djr, full, sizeof(full), hipMemcpyHostToDevice));
    }
    r->nparts = h_update_blocks(g, cfg->precision == MPIC_FP32_SPLIT);
    MPIC_CHECK(hipMalloc(&r->dpart, sizeof(double) * 2 * r->nparts));
    MPIC_CHECK(hipMalloc(&r->dring, sizeof(double) * 5 * DIAG_RING));
    r->ring_step.assign(DIAG_RING, -1);
    r->ring_time.assign(DIAG_RING, 0.0);
    cap_geometry(cfg->n, cfg->bc, kUpdH, kFlH, r->cH);
    cap_geometry(cfg->n, cfg->bc, kUpdE, kFlE, r->cE);
    r->updH = make_set(r->cH.upd_s);
    r->updE = make_set(r->cE.upd_s);
    r->flH = make_set(r->cH.fl_s);
    r->flE = make_set(r->cE.fl_s);
    /* the interiors of the H and E updates start at node 3 and 2 */
    r->flH.shell = complement_shell(r->flH, cfg->n, 3);
    r->flE.shell = complement_shell(r->flE, cfg->n, 2);
    int rc = build_wall_ops(r);
    if (rc != MPIC_OK) {
        mpic_destroy(r);
        return rc;
    }
    PushConsts &pc = r->pc;
    for (int a = 0; a < 3; ++a) {
        pc.h[a] = cfg->h[a];
        pc.L[a] = cfg->L[a];
        pc.shell1[a] = 0 * cfg->h[a] * 1;          /* source/core/plasma.c:379-384 */
        pc.shell2[a] = cfg->n[a] * cfg->h[a] * 1;
    }
    for (int b = 0; b < 6; ++b) pc.bc[b] = cfg->bc[b];
    pc.tau = cfg->tau;
    r->time = cfg->time;
    r->stage = 0;
    r->vsp = 1.0;
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    *out = r;
    return MPIC_OK;
}

static void ensemble_forget(mpic_run *r);

extern "C" int mpic_destroy(mpic_run_t *r)
{
    if (!r) return MPIC_OK;
    ensemble_forget(r);
    hipSetDevice(r->dev);
    if (r->stream) hipStreamSynchronize(r->stream);
    hipFree(r->fblock);
    hipFree(r->rho1);
    hipFree(r->rho2);
    for (int b = 0; b < 2; ++b) {
        Markers64 &mb = r->mbuf[b];
        hipFree(mb.x); hipFree(mb.y); hipFree(mb.z);
        hipFree(mb.ux); hipFree(mb.uy); hipFree(mb.uz);
        hipFree(mb.w); hipFree(mb.qm); hipFree(mb.id);
        Markers32 &m3 = r->m32buf[b];
        hipFree(m3.pos); hipFree(m3.mom); hipFree(m3.id);
    }
    hipFree(r->tv.off); hipFree(r->tv.toff); hipFree(r->tv.cursor); hipFree(r->tv.count); hipFree(r->tv.key);
    hipFree(r->tv.perm);
    hipFree(r->tv.active); hipFree(r->tv.slot); hipFree(r->tv.nactive); hipFree(r->scanw);
    hipFree(r->tv.item); hipFree(r->tv.istart);
    hipFree(r->tj);
    hipFree(r->dacc);
    hipFree(r->derr);
    if (r->hacc) hipHostFree(r->hacc);
    for (double *b : r->murbuf) hipFree(b);
    for (Laser &l : r->lasers) hipFree(l.dmap);
    for (hipEvent_t e : r->evpool) hipEventDestroy(e);
    for (Laser &l : r->lasers) {
        hipFree(l.dtab);
        hipHostFree(l.htab);
    }
    hipFree(r->dstep);
    hipFree(r->djr);
    hipFree(r->dpart);
    hipFree(r->dring);
    hipFree(r->dnode);
    hipFree(r->dprobe);
    hipFree(r->dsample);
    hipFree(r->spot.dprof);
    hipFree(r->tfsf.ddata);
    hipFree(r->trec.dbuf);
    hipFree(r->trec.dfac);
    hipFree(r->tfsf.dtab);
    if (r->tfsf.htab) hipHostFree(r->tfsf.htab);
    hipFree(r->spot.dtab);
    if (r->spot.htab) hipHostFree(r->spot.htab);
    for (int k = 0; k < 2; ++k)
        if (r->exec[k]) hipGraphExecDestroy(r->exec[k]);
    if (r->stream) hipStreamDestroy(r->stream);
    delete r;
    return MPIC_OK;
}

/* ---------------- host <-> device data ---------------- */
static int nodes_of(const Grid &g) { return (g.n[0] + 1) * (g.n[1] + 1) * (g.n[2] + 1); }

template <class T>
static int scatter_block(mpic_run *r, const double *src, T *const dst[3])
{
    /* node block [0,n]^3, interleaved xyz -> padded SoA (ghosts keep NaN) */
    const Grid &g = r->g;
    std::vector<T> tmp((size_t)g.ncell);
    for (int c = 0; c < 3; ++c) {
        MPIC_CHECK(hipMemcpy(tmp.data(), dst[c], sizeof(T) * g.ncell, hipMemcpyDeviceToHost));
        long q = 0;
        for (int i = 0; i <= g.n[0]; ++i)
            for (int j = 0; j <= g.n[1]; ++j)
                for (int k = 0; k <= g.n[2]; ++k, ++q) tmp[gidx(g, i, j, k)] = (T)src[3 * q + c];
        MPIC_CHECK(hipMemcpy(dst[c], tmp.data(), sizeof(T) * g.ncell, hipMemcpyHostToDevice));
    }
    return MPIC_OK;
}

template <class T>
static int gather_block(mpic_run *r, T *const src[3], double *out)
{
    const Grid &g = r->g;
    std::vector<T> tmp((size_t)g.ncell);
    for (int c = 0; c < 3; ++c) {
        MPIC_CHECK(hipMemcpy(tmp.data(), src[c], sizeof(T) * g.ncell, hipMemcpyDeviceToHost));
        long q = 0;
        for (int i = 0; i <= g.n[0]; ++i)
            for (int j = 0; j <= g.n[1]; ++j)
                for (int k = 0; k <= g.n[2]; ++k, ++q) out[3 * q + c] = (double)tmp[gidx(g, i, j, k)];
    }
    return MPIC_OK;
}

extern "C" int mpic_set_fields(mpic_run_t *r, const double *E, const double *H)
{
    if (!r || !E || !H) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    if (r->fp32) {
        TRY(scatter_block(r, E, r->f32.E));
        TRY(scatter_block(r, H, r->f32.H));
    } else {
        TRY(scatter_block(r, E, r->f.E));
        TRY(scatter_block(r, H, r->f.H));
    }
    return MPIC_OK;
}

extern "C" int mpic_get_fields(mpic_run_t *r, double *E, double *H)
{
    if (!r) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    if (r->fp32) {
        if (E) TRY(gather_block(r, r->f32.E, E));
        if (H) TRY(gather_block(r, r->f32.H, H));
    } else {
        if (E) TRY(gather_block(r, r->f.E, E));
        if (H) TRY(gather_block(r, r->f.H, H));
    }
    return MPIC_OK;
}

extern "C" int mpic_get_J(mpic_run_t *r, double *J)
{
    if (!r || !J) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    const Grid &g = r->g;
    std::vector<double> tmp((size_t)g.ncell);
    for (int c = 0; c < 3; ++c) {
        MPIC_CHECK(hipMemcpy(tmp.data(), r->jp.J[c], sizeof(double) * g.ncell, hipMemcpyDeviceToHost));
        long q = 0;
        for (int i = -2; i <= g.n[0] + 2; ++i)
            for (int j = -2; j <= g.n[1] + 2; ++j)
                for (int k = -2; k <= g.n[2] + 2; ++k, ++q) J[3 * q + c] = tmp[gidx(g, i, j, k)];
    }
    return MPIC_OK;
}

/* tecIO_saveFields / tecIO_saveCurrents payloads (source/IO_tecplot.c:
 * 131-266): the node values of E, H or J, computed on the device. */
extern "C" int mpic_node_fields(mpic_run_t *r, int what, double *out)
{
    if (!r || !out || what < 0 || what > 2) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    const Grid &g = r->g;
    const long nodes = (long)(g.n[0] + 1) * (g.n[1] + 1) * (g.n[2] + 1);
    if (!r->dnode) MPIC_CHECK(hipMalloc(&r->dnode, sizeof(double) * 3 * nodes));
    if (what == 2) {
        TRY(launch_node_fields<double>(r->stream, r->jp.J, g, 0, r->dnode));
    } else if (r->fp32) {
        TRY(launch_node_fields<float>(r->stream, what ? r->f32.H : r->f32.E, g, what, r->dnode));
    } else {
        TRY(launch_node_fields<double>(r->stream, what ? r->f.H : r->f.E, g, what, r->dnode));
    }
    MPIC_CHECK(hipMemcpyAsync(out, r->dnode, sizeof(double) * 3 * nodes, hipMemcpyDeviceToHost,
                              r->stream));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    return MPIC_OK;
}

extern "C" int mpic_set_probes(mpic_run_t *r, const int *ijk, int nprobes)
{
    if (!r || nprobes < 0 || (nprobes && !ijk)) return MPIC_ERR_ARG;
    for (int p = 0; p < nprobes; ++p)
        for (int a = 0; a < 3; ++a)
            if (ijk[3 * p + a] < 0 || ijk[3 * p + a] >= r->g.n[a]) {
                mpic_set_error("mpic_set_probes: probe %d node (%d, %d, %d) outside [0, n-1]", p,
                               ijk[3 * p], ijk[3 * p + 1], ijk[3 * p + 2]);
                return MPIC_ERR_ARG;
            }
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    hipFree(r->dprobe);
    hipFree(r->dsample);
    r->dprobe = nullptr;
    r->dsample = nullptr;
    r->nprobes = 0;
    if (!nprobes) return MPIC_OK;
    MPIC_CHECK(hipMalloc(&r->dprobe, sizeof(int) * 3 * nprobes));
    MPIC_CHECK(hipMalloc(&r->dsample, sizeof(double) * 6 * nprobes));
    MPIC_CHECK(hipMemcpy(r->dprobe, ijk, sizeof(int) * 3 * nprobes, hipMemcpyHostToDevice));
    r->nprobes = nprobes;
    return MPIC_OK;
}

/* probe_postData (source/core/diag_probe.c:361-383): the fields at the
 * probe nodes now (the reference samples right after em_HHalfStep). */
extern "C" int mpic_probe_sample(mpic_run_t *r, double *out)
{
    if (!r || (r->nprobes && !out)) return MPIC_ERR_ARG;
    if (!r->nprobes) return MPIC_OK;
    MPIC_CHECK(hipSetDevice(r->dev));
    if (r->fp32)
        TRY(launch_probes<float>(r->stream, r->f32.E, r->f32.H, r->g, r->dprobe, r->nprobes,
                                 r->dsample));
    else
        TRY(launch_probes<double>(r->stream, r->f.E, r->f.H, r->g, r->dprobe, r->nprobes,
                                  r->dsample));
    MPIC_CHECK(hipMemcpyAsync(out, r->dsample, sizeof(double) * 6 * r->nprobes,
                              hipMemcpyDeviceToHost, r->stream));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    return MPIC_OK;
}

/* ---------------- probe samples without allocation ---------------- */
/* Device scratch of mpic_probe_sample_at: the samples (6 doubles per probe)
 * then the probe nodes (3 ints per probe, 2 doubles reserved).  Every GPU
 * holds its own copy until the process ends; all calls on one GPU share it,
 * so calls that overlap on one GPU race. */
static __device__ __align__(256) double scratch[8 * MPIC_PROBE_AT_MAX];

#define MPIC_MAX_DEVICES 64
static void *g_scratch_ptr[MPIC_MAX_DEVICES];
static std::mutex g_scratch_mu;

/* The device address of this device's scratch, taken once per device with
 * that device current (mpic_create calls it for the run's device).  Host
 * code uses this address, never the array's name. */
static int scratch_setup(int dev)
{
    if (dev < 0 || dev >= MPIC_MAX_DEVICES) {
        mpic_set_error("device %d beyond the %d devices the library tracks", dev,
                       MPIC_MAX_DEVICES);
        return MPIC_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[dev]) return MPIC_OK;
    MPIC_CHECK(hipSetDevice(dev));
    void *p = nullptr;
    MPIC_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
        mpic_set_error("hipGetSymbolAddress returned no address for the probe scratch");
        return MPIC_ERR_DEVICE;
    }
    g_scratch_ptr[dev] = p;
    return MPIC_OK;
}

/* mpic_set_probes + mpic_probe_sample in one call that allocates nothing:
 * the nodes and the samples pass through the device's static scratch.  The
 * kernel writes every sample it hands back, so nothing is cleared. */
extern "C" int mpic_probe_sample_at(mpic_run_t *r, const int *ijk, int nprobes, double *out)
{
    if (!r || nprobes < 0 || (nprobes && (!ijk || !out))) return MPIC_ERR_ARG;
    if (nprobes > MPIC_PROBE_AT_MAX) {
        mpic_set_error("mpic_probe_sample_at: %d probes exceed the scratch limit of %d", nprobes,
                       MPIC_PROBE_AT_MAX);
        return MPIC_ERR_CAPACITY;
    }
    for (int p = 0; p < nprobes; ++p)
        for (int a = 0; a < 3; ++a)
            if (ijk[3 * p + a] < 0 || ijk[3 * p + a] >= r->g.n[a]) {
                mpic_set_error("mpic_probe_sample_at: probe %d node (%d, %d, %d) outside "
                               "[0, n-1]", p, ijk[3 * p], ijk[3 * p + 1], ijk[3 * p + 2]);
                return MPIC_ERR_ARG;
            }
    if (!nprobes) return MPIC_OK;
    MPIC_CHECK(hipSetDevice(r->dev));
    void *base = g_scratch_ptr[r->dev];
    if (!base) {
        mpic_set_error("mpic_probe_sample_at: scratch of device %d not set up", r->dev);
        return MPIC_ERR_STATE;
    }
    double *dout = (double *)base;
    int *dijk = (int *)(dout + 6 * MPIC_PROBE_AT_MAX);
    MPIC_CHECK(hipMemcpyAsync(dijk, ijk, sizeof(int) * 3 * nprobes, hipMemcpyHostToDevice,
                              r->stream));
    if (r->fp32)
        TRY(launch_probes<float>(r->stream, r->f32.E, r->f32.H, r->g, dijk, nprobes, dout));
    else
        TRY(launch_probes<double>(r->stream, r->f.E, r->f.H, r->g, dijk, nprobes, dout));
    MPIC_CHECK(hipMemcpyAsync(out, dout, sizeof(double) * 6 * nprobes, hipMemcpyDeviceToHost,
                              r->stream));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    return MPIC_OK;
}

/* Markers in the current device order (the sort drops deleted ones). */
static int live_n(mpic_run *r, long *n)
{
    if (!r->tiled || !r->inited) {
        *n = r->n;
        return MPIC_OK;
    }
    int v = 0;
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    MPIC_CHECK(hipMemcpy(&v, r->tv.nmark + r->parity, sizeof(int), hipMemcpyDeviceToHost));
    *n = v;
    return MPIC_OK;
}

struct RecView {              /* marker_t, source/type_marker.h:12-25 */
    double x, y, z, vx, vy, vz;
    float  rho, qDivM;
    int32_t id;
};

/* Host form of the walls for absolute FP64 positions (plasma_init applies
 * them to the loaded markers, source/core/plasma.c:409-414). */
static int host_walls(const mpic_run *r, double x[3], double u[3])
{
    for (int b = 0; b < 6; ++b) {
        const int ax = b % 3, top = b / 3, bc = r->pc.bc[b];
        const double bnd = top ? r->pc.shell2[ax] : r->pc.shell1[ax];
        if (bc == MPIC_BC_PERIODIC) {
            if (top ? (x[ax] >= bnd) : (x[ax] <= bnd)) x[ax] += top ? -r->pc.L[ax] : r->pc.L[ax];
        } else if (bc == MPIC_BC_MIRROR) {
            if (top ? (x[ax] >= bnd) : (x[ax] <= bnd)) {
                x[ax] = 2 * bnd - x[ax];
                u[ax] *= -1;
            }
        } else if (bc == MPIC_BC_OPEN) {
            const double lim = top ? bnd - r->pc.h[ax] : bnd + r->pc.h[ax];
            if (top ? (x[ax] >= lim) : (x[ax] <= lim)) return 0;
        }
    }
    return 1;
}

static int host_key(const mpic_run *r, int i, int j, int k)
{
    const int *n = r->g.n;
    i = i < 0 ? 0 : (i >= n[0] ? n[0] - 1 : i);
    j = j < 0 ? 0 : (j >= n[1] ? n[1] - 1 : j);
    k = k < 0 ? 0 : (k >= n[2] ? n[2] - 1 : k);
    const int t = ((i >> 2) * r->tv.nt[1] + (j >> 2)) * r->tv.nt[2] + (k >> 2);
    return t * 64 + (((i & 3) * 4 + (j & 3)) * 4 + (k & 3));
}

/* Split-precision load: walls in FP64, then cell index + FP32 offset. */
/* Converts n records to the split form and stores the live ones from
 * marker 'at' on. */
static int load_split(mpic_run *r, std::vector<double> *d, std::vector<float> &w,
                      std::vector<float> &qm, std::vector<int32_t> &id, long n, long at)
{
    std::vector<float> pos, mom;          /* 4 floats per marker each */
    pos.reserve(4 * n);
    mom.reserve(4 * n);
    std::vector<int32_t> ids;
    ids.reserve(n);
    for (long q = 0; q < n; ++q) {
        double x[3] = {d[0][q], d[1][q], d[2][q]}, u[3] = {d[3][q], d[4][q], d[5][q]};
        if (!host_walls(r, x, u)) continue;
        int sp = -1;
        for (int s2 = 0; s2 < r->nspecies; ++s2)
            if (r->qm[s2] == qm[q]) sp = s2;
        if (sp < 0) {
            if (r->nspecies >= 4) {
                mpic_set_error("split form holds at most 4 species (q/m values)");
                return MPIC_ERR_ARG;
            }
            sp = r->nspecies++;
            r->qm[sp] = qm[q];
        }
        int ic[3];
        float of[3];
        for (int a = 0; a < 3; ++a) {
            const double sgm = x[a] / r->g.h[a];
            ic[a] = (int)(sgm + 128.0) - 128;
            double o = sgm - ic[a];
            float fo = (float)o;
            if (fo >= 1.0f) fo = 0.99999994f;
            if (fo < 0.0f) fo = 0.0f;
            of[a] = fo;
            if (ic[a] < 0 || ic[a] >= r->g.n[a]) {
                mpic_set_error("marker %ld lies outside the box", q);
                return MPIC_ERR_ARG;
            }
        }
        const int32_t word = (sp << 29) | host_key(r, ic[0], ic[1], ic[2]);
        float wf;
        memcpy(&wf, &word, sizeof(wf));
        pos.push_back(wf); pos.push_back(of[0]); pos.push_back(of[1]); pos.push_back(of[2]);
        mom.push_back((float)u[0]); mom.push_back((float)u[1]); mom.push_back((float)u[2]);
        mom.push_back(w[q]);
        ids.push_back(id[q]);
    }
    const long m = (long)ids.size();
    Markers32 &M = r->m32;
    if (m > 0) {
        MPIC_CHECK(hipMemcpy(M.pos + at, pos.data(), sizeof(float4) * m, hipMemcpyHostToDevice));
        MPIC_CHECK(hipMemcpy(M.mom + at, mom.data(), sizeof(float4) * m, hipMemcpyHostToDevice));
        MPIC_CHECK(hipMemcpy(M.id + at, ids.data(), sizeof(int32_t) * m, hipMemcpyHostToDevice));
    }
    r->n = at + m;
    return MPIC_OK;
}

static int save_split(mpic_run *r, void *rec, long *n, int recsize, long N)
{
    Markers32 &M = r->m32;
    /* the N markers of the order: through the permutation once initialised
     * (the buffer's first E records are written), else the loaded records */
    long E = N;
    std::vector<int32_t> order;
    if (r->inited) {
        int ext = 0;
        MPIC_CHECK(hipMemcpy(&ext, r->tv.nmark + (r->parity ^ 1), sizeof(int), hipMemcpyDeviceToHost));
        E = ext;
        order.resize(N);
        if (N)
            MPIC_CHECK(hipMemcpy(order.data(), r->tv.perm, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
    }
    std::vector<int32_t> id(E);
    std::vector<float> pos(4 * E), mom(4 * E);
    if (E) {
        MPIC_CHECK(hipMemcpy(pos.data(), M.pos, sizeof(float4) * E, hipMemcpyDeviceToHost));
        MPIC_CHECK(hipMemcpy(mom.data(), M.mom, sizeof(float4) * E, hipMemcpyDeviceToHost));
        MPIC_CHECK(hipMemcpy(id.data(), M.id, sizeof(int32_t) * E, hipMemcpyDeviceToHost));
    }
    char *dst = (char *)rec;
    long o = 0;
    for (long u = 0; u < N; ++u) {
        const long q = r->inited ? order[u] : u;
        if (q < 0 || q >= E) {
            mpic_set_error("mpic_markers_save: order entry %ld names record %ld of %ld", u, q, E);
            return MPIC_ERR_STATE;
        }
        int32_t word;
        memcpy(&word, &pos[4 * q], sizeof(word));
        if (word < 0) continue;
        if (o >= *n) {
            mpic_set_error("mpic_markers_save: output holds %ld markers, more are live", *n);
            return MPIC_ERR_CAPACITY;
        }
        const int key = word & ((1 << 29) - 1), sp = (int)((unsigned)word >> 29);
        const int t = key >> 6, loc = key & 63;
        const int t2 = t % r->tv.nt[2], t1 = (t / r->tv.nt[2]) % r->tv.nt[1],
                  t0 = t / (r->tv.nt[2] * r->tv.nt[1]);
        const int ic[3] = {4 * t0 + (loc >> 4), 4 * t1 + ((loc >> 2) & 3), 4 * t2 + (loc & 3)};
        RecView v;
        memset(&v, 0, sizeof(v));
        v.x = (ic[0] + (double)pos[4 * q + 1]) * r->g.h[0];
        v.y = (ic[1] + (double)pos[4 * q + 2]) * r->g.h[1];
        v.z = (ic[2] + (double)pos[4 * q + 3]) * r->g.h[2];
        v.vx = mom[4 * q]; v.vy = mom[4 * q + 1]; v.vz = mom[4 * q + 2];
        v.rho = mom[4 * q + 3];
        v.qDivM = r->qm[sp];
        v.id = id[q];
        memset(dst + o * recsize, 0, recsize);
        memcpy(dst + o * recsize, &v, recsize == 64 ? 60 : 56);
        ++o;
    }
    *n = o;
    return MPIC_OK;
}

/* Stores n records from marker 'at' on (at = 0: a fresh load). */
static int load_records(mpic_run *r, const void *rec, long n, int recsize, long at,
                        const char *who)
{
    if (!r || (!rec && n) || n < 0 || (recsize != 56 && recsize != 64)) {
        mpic_set_error("%s: bad arguments", who);
        return MPIC_ERR_ARG;
    }
    invalidate(r);
    if (n > r->cap - at) {
        mpic_set_error("%s: %ld markers exceed the capacity %ld", who, at + n, r->cap);
        return MPIC_ERR_CAPACITY;
    }
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    std::vector<double> d[6];
    std::vector<float> w(n), qm(n);
    std::vector<int32_t> id(n);
    for (int c = 0; c < 6; ++c) d[c].resize(n);
    const char *src = (const char *)rec;
    for (long q = 0; q < n; ++q) {
        RecView v;
        memset(&v, 0, sizeof(v));
        memcpy(&v, src + q * recsize, recsize == 64 ? 60 : 56);
        d[0][q] = v.x; d[1][q] = v.y; d[2][q] = v.z;
        d[3][q] = v.vx; d[4][q] = v.vy; d[5][q] = v.vz;
        w[q] = v.rho;
        qm[q] = v.qDivM;
        id[q] = (recsize == 64) ? v.id : (int32_t)(at + q + 1);
    }
    if (r->fp32) return load_split(r, d, w, qm, id, n, at);
    r->n = at + n;
    if (n == 0) return MPIC_OK;
    double *dst[6] = {r->m.x, r->m.y, r->m.z, r->m.ux, r->m.uy, r->m.uz};
    for (int c = 0; c < 6; ++c)
        MPIC_CHECK(hipMemcpy(dst[c] + at, d[c].data(), sizeof(double) * n, hipMemcpyHostToDevice));
    MPIC_CHECK(hipMemcpy(r->m.w + at, w.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    MPIC_CHECK(hipMemcpy(r->m.qm + at, qm.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    MPIC_CHECK(hipMemcpy(r->m.id + at, id.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    return MPIC_OK;
}

extern "C" int mpic_markers_load(mpic_run_t *r, const void *rec, long n, int recsize)
{
    /* after mpic_init the tile ranges, live counts and work items describe
     * the loaded order: a new record would be pushed through stale ones */
    if (r && r->inited) {
        mpic_set_error("mpic_markers_load: markers are loaded before mpic_init");
        return MPIC_ERR_STATE;
    }
    return load_records(r, rec, n, recsize, 0, "mpic_markers_load");
}

extern "C" int mpic_markers_append(mpic_run_t *r, const void *rec, long n, int recsize)
{
    if (r && r->inited) {
        mpic_set_error("mpic_markers_append: markers are appended before mpic_init");
        return MPIC_ERR_STATE;
    }
    return load_records(r, rec, n, recsize, r ? r->n : 0, "mpic_markers_append");
}

extern "C" int mpic_markers_save(mpic_run_t *r, void *rec, long *n, int recsize)
{
    if (!r || !n || (recsize != 56 && recsize != 64)) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    long N = 0;
    TRY(live_n(r, &N));
    if (r->fp32) return save_split(r, rec, n, recsize, N);
    /* the N markers of the order: through the permutation once a tiled run
     * is initialised (the buffer's first E records are written), else the
     * first N records */
    const bool through = r->tiled && r->inited;
    long E = N;
    std::vector<int32_t> order;
    if (through) {
        int ext = 0;
        MPIC_CHECK(hipMemcpy(&ext, r->tv.nmark + (r->parity ^ 1), sizeof(int), hipMemcpyDeviceToHost));
        E = ext;
        order.resize(N);
        if (N)
            MPIC_CHECK(hipMemcpy(order.data(), r->tv.perm, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
    }
    std::vector<double> d[6];
    std::vector<float> w(E), qm(E);
    std::vector<int32_t> id(E);
    double *src[6] = {r->m.x, r->m.y, r->m.z, r->m.ux, r->m.uy, r->m.uz};
    for (int c = 0; c < 6; ++c) {
        d[c].resize(E);
        if (E) MPIC_CHECK(hipMemcpy(d[c].data(), src[c], sizeof(double) * E, hipMemcpyDeviceToHost));
    }
    if (E) {
        MPIC_CHECK(hipMemcpy(w.data(), r->m.w, sizeof(float) * E, hipMemcpyDeviceToHost));
        MPIC_CHECK(hipMemcpy(qm.data(), r->m.qm, sizeof(float) * E, hipMemcpyDeviceToHost));
        MPIC_CHECK(hipMemcpy(id.data(), r->m.id, sizeof(int32_t) * E, hipMemcpyDeviceToHost));
    }
    long o = 0;
    char *dst = (char *)rec;
    for (long u = 0; u < N; ++u) {
        const long q = through ? order[u] : u;
        if (q < 0 || q >= E) {
            mpic_set_error("mpic_markers_save: order entry %ld names record %ld of %ld", u, q, E);
            return MPIC_ERR_STATE;
        }
        if (qm[q] == 0.0f) continue;
        if (o >= *n) {
            mpic_set_error("mpic_markers_save: output holds %ld markers, more are live", *n);
            return MPIC_ERR_CAPACITY;
        }
        RecView v;
        memset(&v, 0, sizeof(v));
        v.x = d[0][q]; v.y = d[1][q]; v.z = d[2][q];
        v.vx = d[3][q]; v.vy = d[4][q]; v.vz = d[5][q];
        v.rho = w[q];
        v.qDivM = qm[q];
        v.id = id[q];
        memset(dst + o * recsize, 0, recsize);
        memcpy(dst + o * recsize, &v, recsize == 64 ? 60 : 56);
        ++o;
    }
    *n = o;
    return MPIC_OK;
}

extern "C" long mpic_marker_count(mpic_run_t *r)
{
    if (!r) return -1;
    if (hipSetDevice(r->dev) != hipSuccess) return -1;
    long n = 0;
    if (live_n(r, &n) != MPIC_OK) return -1;
    if (r->tiled && r->inited) return n;          /* the sort keeps live markers only */
    if (r->fp32) return n;
    std::vector<float> qm(n);
    if (hipStreamSynchronize(r->stream) != hipSuccess) return -1;
    if (n && hipMemcpy(qm.data(), r->m.qm, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess)
        return -1;
    long c = 0;
    for (long q = 0; q < n; ++q) c += (qm[q] != 0.0f);
    return c;
}

extern "C" int mpic_add_laser(mpic_run_t *r, double A, double T0, double delay, double dt,
                              double phi0, int forward, int plane, const double *map)
{
    if (!r || !map) return MPIC_ERR_ARG;
    if (r->inited) {
        mpic_set_error("mpic_add_laser: add sources before mpic_init");
        return MPIC_ERR_STATE;
    }
    const Grid &g = r->g;
    /* The source plane and its transverse extent [5, n-5] must lie inside
     * the interior the caps do not cover (the reference keeps sources 5
     * cells off the walls for that reason, source/scpic/em_mirror.c:27-35). */
    const int I = forward ? plane : g.n[0] - plane + 1;
    if (I - 1 < 3 || I > g.n[0] - 2 || g.n[1] < 11 || g.n[2] < 11) {
        mpic_set_error("mpic_add_laser: source plane %d outside the interior", plane);
        return MPIC_ERR_ARG;
    }
    Laser l;
    l.A = A; l.T0 = T0; l.delay = delay; l.dt = dt; l.phi0 = phi0;
    l.forward = forward; l.plane = plane;
    l.nodes = (long)(g.n[1] + 1 - 10) * (g.n[2] + 1 - 10);
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipMalloc(&l.dmap, sizeof(double) * 8 * l.nodes));
    MPIC_CHECK(hipMemcpy(l.dmap, map, sizeof(double) * 8 * l.nodes, hipMemcpyHostToDevice));
    MPIC_CHECK(hipMalloc(&l.dtab, sizeof(double) * 6 * LASER_CAP));
    MPIC_CHECK(hipHostMalloc(&l.htab, sizeof(double) * 6 * LASER_CAP, hipHostMallocDefault));
    r->lasers.push_back(l);
    r->tab_hi = r->tab_lo = 0;      /* refill before the next step */
    invalidate(r);
    return MPIC_OK;
}

/* gaussSpot_init (source/core/em_gaussSpot.c:42-120): p = the 11 values of
 * .gaussSpot.cfg in file order.  The spatial profiles are evaluated here,
 * on the host, as gaussSpot_forceE does per step. */
extern "C" int mpic_add_gauss_spot(mpic_run_t *r, const double *p)
{
    if (!r || !p) return MPIC_ERR_ARG;
    if (r->spot.on) {
        mpic_set_error("mpic_add_gauss_spot: one spot per run");
        return MPIC_ERR_STATE;
    }
    const Grid &g = r->g;
    Spot &sp = r->spot;
    const double ampl = p[0], x = p[2], y = p[3], z = p[4], width = p[5];
    sp.omega = 2 * M_PI * p[1];
    sp.duration = p[6];
    sp.front = p[7];
    sp.plateau = p[8];
    const int ey = (int)p[9], ez = (int)p[10];
    sp.Ey = ((ey > 0) - (ey < 0)) * ampl;
    sp.Ez = ((ez > 0) - (ez < 0)) * ampl;
    sp.i = (int)(x / g.h[0] + 0.5);
    if (!(width > 0) || !(sp.duration > 0)) {
        mpic_set_error("mpic_add_gauss_spot: width and duration must be positive");
        return MPIC_ERR_ARG;
    }
    const int nj = g.n[1] + 3, nk = g.n[2] + 3;
    std::vector<double> prof(2 * (size_t)nj * nk);
    for (int j = -1; j <= g.n[1] + 1; ++j) {
        const double dy_ = ((j - 0.5) * g.h[1] - y) / width, dy = (j * g.h[1] - y) / width;
        for (int k = -1; k <= g.n[2] + 1; ++k) {
            const double dz = (k * g.h[2] - z) / width, dz_ = ((k - 0.5) * g.h[2] - z) / width;
            const size_t o = (size_t)(j + 1) * nk + (k + 1);
            prof[o] = exp(-dy_ * dy_ - dz * dz);
            prof[(size_t)nj * nk + o] = exp(-dy * dy - dz_ * dz_);
        }
    }
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipMalloc(&sp.dprof, sizeof(double) * prof.size()));
    MPIC_CHECK(hipMemcpy(sp.dprof, prof.data(), sizeof(double) * prof.size(), hipMemcpyHostToDevice));
    MPIC_CHECK(hipMalloc(&sp.dtab, sizeof(double) * 2 * LASER_CAP));
    MPIC_CHECK(hipHostMalloc(&sp.htab, sizeof(double) * 2 * LASER_CAP, hipHostMallocDefault));
    /* outside the E mesh [-1, n+1] the reference skips the spot */
    sp.on = sp.i >= -1 && sp.i <= g.n[0] + 1;
    r->tab_hi = r->tab_lo = 0;
    invalidate(r);
    return MPIC_OK;
}

/* Face geometry of the interface box reg = {imin, imax, jmin, jmax, kmin,
 * kmax}, TFSF_init (source/core/em_TFSF.c:500-537): E face of slot
 * f = axis*4 + top*2 + comp on the box side 'top' along 'axis', component
 * (axis + 1 + comp) % 3 with its lower corner moved up by one node along the
 * component; the H face of the slot is the E face moved by 'top' along the
 * axis, component 3 - cE - axis.  Every face keeps 5 cells from the domain
 * edges (TFSF_turnPlay / tag_TFSF.c:106-117). */
static int tfsf_faces(const Grid &g, const int reg[6], TfsfSet &E, TfsfSet &H, long &maxvol)
{
    const int side[2][3] = {{reg[0], reg[2], reg[4]}, {reg[1], reg[3], reg[5]}};
    maxvol = 0;
    for (int ea = 0; ea < 3; ++ea) {
        const int ep = (ea + 1) % 3, eq = (ea + 2) % 3;
        for (int top = 0; top < 2; ++top)
            for (int comp = 0; comp < 2; ++comp) {
                const int f = ea * 4 + top * 2 + comp;
                TfsfFace &fe = E.f[f], &fh = H.f[f];
                for (int a = 0; a < 3; ++a) {
                    fe.lo[a] = side[0][a];
                    fe.hi[a] = side[1][a];
                }
                fe.lo[ea] = fe.hi[ea] = side[top][ea];
                fe.comp = (ea + 1 + comp) % 3;
                fe.lo[fe.comp] += 1;
                fe.ea = ea;
                fe.ep = ep;
                fe.eq = eq;
                fh = fe;
                fh.lo[ea] += top;
                fh.hi[ea] += top;
                fh.comp = 3 - fe.comp - ea;
                fe.dQ = fh.dQ = fe.hi[eq] - fe.lo[eq] + 1;
                fe.vol = fh.vol = (long)(fe.hi[ep] - fe.lo[ep] + 1) * fe.dQ;
                for (int a = 0; a < 3; ++a)
                    if (fe.lo[a] < 5 || fe.hi[a] > g.n[a] - 5 || fh.lo[a] < 5 || fh.hi[a] > g.n[a] - 5) {
                        mpic_set_error("TFSF interface is too close to a boundary (less than 5 cell)");
                        return MPIC_ERR_ARG;
                    }
                if (fe.vol <= 0) {
                    mpic_set_error("TFSF: empty interface box");
                    return MPIC_ERR_ARG;
                }
                if (fe.vol > maxvol) maxvol = fe.vol;
            }
    }
    return MPIC_OK;
}

extern "C" int mpic_add_tfsf(mpic_run_t *r, const int reg[6], int mode, int nframes,
                             const double *const dataE[12], const double *const dataH[12])
{
    if (!r || !reg || !dataE || !dataH) return MPIC_ERR_ARG;
    if (r->inited) {
        mpic_set_error("mpic_add_tfsf: add sources before mpic_init");
        return MPIC_ERR_STATE;
    }
    if (r->tfsf.on) {
        mpic_set_error("mpic_add_tfsf: one interface per run");
        return MPIC_ERR_STATE;
    }
    if ((mode != 0 && mode != 2) || nframes < 1) {
        mpic_set_error("mpic_add_tfsf: mode must be 0 (playForward) or 2 (playBackward), "
                       "frames >= 1");
        return MPIC_ERR_ARG;
    }
    const Grid &g = r->g;
    Tfsf &tf = r->tfsf;
    TRY(tfsf_faces(g, reg, tf.fe, tf.fh, tf.maxvol));
    long total = 0;
    for (int f = 0; f < 12; ++f) {
        TfsfFace &fe = tf.fe.f[f], &fh = tf.fh.f[f];
        fe.off = dataE[f] ? total : -1;
        total += dataE[f] ? fe.vol * nframes : 0;
        fh.off = dataH[f] ? total : -1;
        total += dataH[f] ? fh.vol * nframes : 0;
    }
    tf.fe.lo_in = 2;                     /* innerRegE, source/core/em.c:197-198 */
    tf.fh.lo_in = 3;                     /* innerRegH, source/core/em.c:122-123 */
    MPIC_CHECK(hipSetDevice(r->dev));
    if (total > 0) {
        MPIC_CHECK(hipMalloc(&tf.ddata, sizeof(double) * total));
        for (int f = 0; f < 12; ++f) {
            if (dataE[f])
                MPIC_CHECK(hipMemcpy(tf.ddata + tf.fe.f[f].off, dataE[f],
                                     sizeof(double) * tf.fe.f[f].vol * nframes, hipMemcpyHostToDevice));
            if (dataH[f])
                MPIC_CHECK(hipMemcpy(tf.ddata + tf.fh.f[f].off, dataH[f],
                                     sizeof(double) * tf.fh.f[f].vol * nframes, hipMemcpyHostToDevice));
        }
    }
    MPIC_CHECK(hipMalloc(&tf.dtab, sizeof(int) * 2 * LASER_CAP));
    MPIC_CHECK(hipHostMalloc(&tf.htab, sizeof(int) * 2 * LASER_CAP, hipHostMallocDefault));
    tf.mode = mode;
    tf.last = nframes;
    tf.frame0 = (int)(r->time / g.tau + 0.1);   /* TFSF_turnPlay, :466 */
    tf.step0 = r->step;
    tf.on = 1;
    r->tab_hi = r->tab_lo = 0;
    invalidate(r);
    return MPIC_OK;
}

/* TFSF record mode (TFSF_turnRec, source/core/em_TFSF.c:273-330). */
extern "C" int mpic_tfsf_record(mpic_run_t *r, const int reg[6], int max_frames)
{
    if (!r || !reg || max_frames < 1) return MPIC_ERR_ARG;
    if (r->inited) {
        mpic_set_error("mpic_tfsf_record: start the recording before mpic_init");
        return MPIC_ERR_STATE;
    }
    if (r->tfsf.on || r->trec.on) {
        mpic_set_error("mpic_tfsf_record: one interface per run");
        return MPIC_ERR_STATE;
    }
    const Grid &g = r->g;
    TfsfRec &tr = r->trec;
    TRY(tfsf_faces(g, reg, tr.fe, tr.fh, tr.maxvol));
    long off = 0;
    for (int f = 0; f < 12; ++f) {
        tr.fe.f[f].off = off;
        off += tr.fe.f[f].vol * max_frames;
    }
    for (int f = 0; f < 12; ++f) {
        tr.fh.f[f].off = off;
        off += tr.fh.f[f].vol * max_frames;
    }
    double fac[12];
    for (int f = 0; f < 12; ++f) {
        const int ea = f >> 2, top = (f >> 1) & 1, comp = f & 1;
        fac[f] = (1 - 2 * top) * (1 - 2 * comp) * (g.tau / g.h[ea]);
    }
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipMalloc(&tr.dbuf, sizeof(double) * off));
    MPIC_CHECK(hipMalloc(&tr.dfac, sizeof(double) * 12));
    MPIC_CHECK(hipMemcpy(tr.dfac, fac, sizeof(fac), hipMemcpyHostToDevice));
    tr.cap = max_frames;
    tr.step0 = r->step;
    tr.on = 1;
    invalidate(r);
    return MPIC_OK;
}

extern "C" int mpic_tfsf_recorded(mpic_run_t *r, int *nframes, long vol[24], double *out)
{
    if (!r || !nframes) return MPIC_ERR_ARG;
    const TfsfRec &tr = r->trec;
    if (!tr.on) {
        mpic_set_error("mpic_tfsf_recorded: no recording");
        return MPIC_ERR_STATE;
    }
    long done = r->step - tr.step0 + (r->stage != 0);
    if (done > tr.cap) done = tr.cap;
    *nframes = (int)done;
    if (vol)
        for (int f = 0; f < 24; ++f) vol[f] = f < 12 ? tr.fe.f[f].vol : tr.fh.f[f - 12].vol;
    if (!out) return MPIC_OK;
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    long at = 0;
    for (int q = 0; q < 24; ++q) {
        const TfsfFace &fc = q < 12 ? tr.fe.f[q] : tr.fh.f[q - 12];
        MPIC_CHECK(hipMemcpy(out + at, tr.dbuf + fc.off, sizeof(double) * fc.vol * done,
                             hipMemcpyDeviceToHost));
        at += fc.vol * done;
    }
    return MPIC_OK;
}

/* ---------------- stages ---------------- */
/* The walls in list order (x, then y, then z): the x walls with the min and
 * max wall of an axis with at least 8 cells (adjacent in the list, disjoint
 * node layers) sharing a launch, then the y and z walls in one launch of a
 * workgroup per x plane (their lines keep i fixed). */
static int walls_on(mpic_run *r, std::vector<FaceOp> &ops)
{
    size_t nx = 0;
    while (nx < ops.size() && ops[nx].ea == 0) ++nx;
    if (ops.size() - nx <= 4) {
        for (size_t q = nx; q < ops.size(); ++q)
            if (ops[q].ea == 0) goto generic;       /* not in x, y, z order */
        for (size_t q = 0; q < nx; ++q) {
            const FaceOp &a = ops[q];
            if (q + 1 < nx) {
                const FaceOp &b = ops[q + 1];
                if (a.top == 0 && b.top == 1 && a.kind == b.kind && a.kind != FACE_PERIODIC_H &&
                    a.kind != FACE_PERIODIC_E && a.N >= 8) {
                    TRY(launch_face_pair(r->stream, r->g, a, b));
                    ++q;
                    continue;
                }
            }
            TRY(launch_face(r->stream, r->g, a));
        }
        return launch_face_planes(r->stream, r->g, ops.data() + nx, (int)(ops.size() - nx));
    }
generic:
    for (size_t q = 0; q < ops.size(); ++q) {
        const FaceOp &a = ops[q];
        if (q + 1 < ops.size()) {
            const FaceOp &b = ops[q + 1];
            if (a.ea == b.ea && a.top == 0 && b.top == 1 && a.kind == b.kind &&
                a.kind != FACE_PERIODIC_H && a.kind != FACE_PERIODIC_E && a.N >= 8) {
                TRY(launch_face_pair(r->stream, r->g, a, b));
                ++q;
                continue;
            }
        }
        TRY(launch_face(r->stream, r->g, a));
    }
    return MPIC_OK;
}

template <class T>
static int init_ghosts(mpic_run *r, T *const E[3], T *const H[3]);

/* The sorted array becomes the current one. */
static void swap_markers(mpic_run *r)
{
    const int other = r->cur ^ 1;
    if (r->fp32)
        r->m32 = r->m32buf[other];
    else
        r->m = r->mbuf[other];
    r->cur = other;
    r->parity ^= 1;
}

/* Counting sort of the current markers by cell key: with keys == 1 the
 * keys are computed here (loaded or untouched markers), else they come
 * from the push that just ran. */
static int sort_pass(mpic_run *r, int keys)
{
    const int other = r->cur ^ 1;
    if (r->tiled) {
        /* The order is a permutation of the current buffer: tv.nmark[parity]
         * markers, the buffer's first tv.nmark[parity ^ 1] records written.
         * After a push (keys == 0) the push's output buffer and keys are
         * ranked and become current; keys == 1 ranks the current buffer in
         * place (loaded markers: nmark[parity] set by the caller, or a
         * re-sort: the buffer's extent), moving nothing. */
        if (keys) {
            if (r->inited)
                MPIC_CHECK(hipMemcpyAsync(r->tv.nmark + r->parity, r->tv.nmark + (r->parity ^ 1),
                                          sizeof(int), hipMemcpyDeviceToDevice, r->stream));
            if (r->fp32)
                TRY(launch_keys32(r->stream, r->m32, r->n, r->tv, r->parity));
            else
                TRY(launch_keys64(r->stream, r->m, r->n, r->g, r->tv, r->parity));
        }
        TRY(launch_tile_scan(r->stream, r->tv, r->scanw, r->derr, r->parity ^ 1));
        TRY(launch_rank(r->stream, r->n, r->tv, r->parity));
        if (keys)
            r->parity ^= 1;
        else
            swap_markers(r);
        return MPIC_OK;
    }
    /* the untiled FP64 push keeps no order: a sort compacts its array */
    if (keys) TRY(launch_keys64(r->stream, r->m, r->n, r->g, r->tv, r->parity));
    TRY(launch_tile_scan(r->stream, r->tv, r->scanw, r->derr, r->parity ^ 1));
    TRY(launch_scatter64(r->stream, r->m, r->mbuf[other], r->n, r->tv, r->parity));
    swap_markers(r);
    return MPIC_OK;
}

extern "C" int mpic_init(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    if (r->inited) {
        mpic_set_error("mpic_init: double initialisation");
        return MPIC_ERR_STATE;
    }
    MPIC_CHECK(hipSetDevice(r->dev));
    const Grid &g = r->g;
    const int *bc = r->cfg.bc;
    if (r->fp32)
        TRY(init_ghosts(r, r->f32.E, r->f32.H));
    else
        TRY(init_ghosts(r, r->f.E, r->f.H));
    /* plasma_init: walls applied once to every loaded marker
     * (source/core/plasma.c:409-414). */
    if (!r->fp32) TRY(launch_push64(r->stream, r->m, r->n, r->f, g, r->pc, r->dacc, 1, r->derr));
    if (r->tiled) {
        /* first sort of the loaded markers by tile-major cell key */
        const int nn = (int)r->n;
        MPIC_CHECK(hipMemcpyAsync(r->tv.nmark + r->parity, &nn, sizeof(int), hipMemcpyHostToDevice,
                                  r->stream));
        TRY(sort_pass(r, 1));
    }
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    r->inited = 1;
    r->stage = 0;
    return MPIC_OK;
}

/* cap_resetGhosts on the main meshes (source/core/em_caps.c:266-308; the
 * main meshes carry one ghost layer: transverse extent [-1, n+1]) and the
 * Mur buffers from the main E (capBC_Mur_init, source/core/emCap_Mur.c:
 * 92-132). */
template <class T>
static int init_ghosts(mpic_run *r, T *const E[3], T *const H[3])
{
    const Grid &g = r->g;
    const int *bc = r->cfg.bc;
    for (int ax = 0; ax < 3; ++ax) {
        Box full;
        for (int a = 0; a < 3; ++a) {
            full.lo[a] = -1;
            full.hi[a] = g.n[a] + 1;
        }
        if (bc[ax] == MPIC_BC_PERIODIC) {
            TRY(launch_face(r->stream, g, face(FACE_PERIODIC_H, ax, 0, full, H, g.n[ax])));
            TRY(launch_face(r->stream, g, face(FACE_PERIODIC_E, ax, 0, full, E, g.n[ax])));
        } else if (bc[ax] == MPIC_BC_MIRROR) {
            FaceOp h = face(FACE_WALL_H, ax, 0, full, H, g.n[ax]);
            h.zeroN = 1;
            TRY(launch_face(r->stream, g, h));
            TRY(launch_face(r->stream, g, face(FACE_MIRROR_E, ax, 0, full, E, g.n[ax])));
        }
        if (bc[3 + ax] == MPIC_BC_MIRROR) {
            FaceOp h = face(FACE_WALL_H, ax, 1, full, H, g.n[ax]);
            h.zeroN = 1;
            TRY(launch_face(r->stream, g, h));
            TRY(launch_face(r->stream, g, face(FACE_MIRROR_E, ax, 1, full, E, g.n[ax])));
        }
    }
    for (FaceOp &op : r->bcE) {
        if (op.kind != FACE_MUR_E) continue;
        FaceOp ini = op;
        ini.kind = FACE_MUR_INIT;
        for (int c = 0; c < 3; ++c) ini.X[c] = (void *)E[c];
        TRY(launch_face(r->stream, g, ini));
    }
    return MPIC_OK;
}

static int check_stage(mpic_run *r, int want)
{
    if (!r->inited) {
        mpic_set_error("run is not initialised (mpic_init)");
        return MPIC_ERR_STATE;
    }
    if (r->stage != want) {
        mpic_set_error("stage order violated: at stage %d, call needs stage %d", r->stage, want);
        return MPIC_ERR_ORDER;
    }
    return MPIC_OK;
}

/* The laser tables for steps [r->step, r->step + n): amplitude x tau/h,
 * cos and sin of the phase of each step, for the H source (time t) and the
 * E source (t + tau/2), with t advanced exactly as the run's clock
 * (source/scpic/em_mirror.c:667-717).  Refilled LASER_CAP steps at a time
 * on the run's stream. */
static int laser_tables(mpic_run *r, long n)
{
    if ((r->lasers.empty() && !r->spot.on && !r->tfsf.on) ||
        (r->step >= r->tab_lo && r->step + n <= r->tab_hi))
        return MPIC_OK;
    if (n > LASER_CAP) {
        mpic_set_error("laser tables hold %d steps", LASER_CAP);
        return MPIC_ERR_ARG;
    }
    const Grid &g = r->g;
    MPIC_CHECK(hipStreamSynchronize(r->stream));     /* the staging buffers are reused */
    for (Laser &l : r->lasers) {
        double time = r->time;
        for (long q = 0; q < LASER_CAP; ++q) {
            const long idx = (r->step + q) % LASER_CAP;
            for (int half = 0; half < 2; ++half) {
                const double t = (half ? time + 0.5 * g.tau : time) - l.T0 + l.delay;
                const double ampl = exp(-pow(2 * t / l.dt, 2));
                const double phase = 2 * M_PI * t + l.phi0;
                double *e = l.htab + 3 * (half * LASER_CAP + idx);
                e[0] = ampl * g.tau / g.h[0];
                e[1] = cos(phase);
                e[2] = sin(phase);
            }
            time = time + g.tau;
        }
        MPIC_CHECK(hipMemcpyAsync(l.dtab, l.htab, sizeof(double) * 6 * LASER_CAP,
                                  hipMemcpyHostToDevice, r->stream));
    }
    if (r->spot.on) {
        /* gaussSpot_forceE at Time + tau: the time envelope and oscillation
         * (source/core/em_gaussSpot.c:133-147) */
        const Spot &sp = r->spot;
        double time = r->time;
        for (long q = 0; q < LASER_CAP; ++q) {
            const double t = time + g.tau;
            double tf = 1;
            if (t < sp.front) {
                const double l = (t - sp.front) / sp.duration;
                tf = exp(-l * l);
            }
            if (t > sp.front + sp.plateau) {
                const double l = (t - sp.front - sp.plateau) / sp.duration;
                tf = exp(-l * l);
            }
            double *e = sp.htab + 2 * ((r->step + q) % LASER_CAP);
            e[0] = sp.Ey * sin(sp.omega * t) * tf;
            e[1] = sp.Ez * cos(sp.omega * t) * tf;
            time = time + g.tau;
        }
        MPIC_CHECK(hipMemcpyAsync(sp.dtab, sp.htab, sizeof(double) * 2 * LASER_CAP,
                                  hipMemcpyHostToDevice, r->stream));
    }
    if (r->tfsf.on) {
        /* TFSF_cacheFrame / post-steps (source/core/em_TFSF.c:780-850): at
         * frame counter F the interface is live while F + 1 < last - 1 (the
         * caps' TFSF_postEStep at the start of a step switches it off,
         * source/core/em_caps.c:607); H faces play F (forward) or
         * last - 1 - F, E faces F + 1 or last - 1 - F. */
        const Tfsf &tf = r->tfsf;
        for (long q = 0; q < LASER_CAP; ++q) {
            const long F = tf.frame0 + (r->step + q - tf.step0);
            const bool live = F >= 0 && F + 1 < tf.last - 1;   /* F < 0: before the recording */
            int *e = tf.htab + 2 * ((r->step + q) % LASER_CAP);
            e[0] = live ? (int)(tf.mode == 0 ? F : tf.last - 1 - F) : -1;
            e[1] = live ? (int)(tf.mode == 0 ? F + 1 : tf.last - 1 - F) : -1;
        }
        MPIC_CHECK(hipMemcpyAsync(tf.dtab, tf.htab, sizeof(int) * 2 * LASER_CAP,
                                  hipMemcpyHostToDevice, r->stream));
    }
    r->tab_lo = r->step;
    r->tab_hi = r->step + LASER_CAP;
    return MPIC_OK;
}

/* Plane and node ranges of laser l's H (h = 1) or E source within box tu
 * (source/scpic/em_mirror.c:667-693: rows [5, n-5] of the map's
 * [-1, n+1] x [-1, n+1] frame, H on the source plane, E one plane before it
 * for a forward beam); on = 0 when the plane lies outside tu. */
static LaserPlane laser_plane(const mpic_run *r, const Laser &l, int h, const Box &tu)
{
    const Grid &g = r->g;
    LaserPlane lp;
    const int I = l.forward ? (h ? l.plane : l.plane - 1) : g.n[0] - l.plane + 1;
    if (I < tu.lo[0] || I > tu.hi[0]) return lp;
    lp.JMIN = (int)(fmax(-1, 5) + 0.1);
    const int JMAX = (int)(fmin(g.n[1] + 1, g.n[1] - 5) + 0.1);
    lp.KMIN = (int)(fmax(-1, 5) + 0.1);
    lp.KMAX = (int)(fmin(g.n[2] + 1, g.n[2] - 5) + 0.1);
    lp.jmin = (int)(fmax(tu.lo[1], lp.JMIN) + 0.1);
    lp.jmax = (int)(fmin(tu.hi[1], JMAX) + 0.1);
    lp.kmin = (int)(fmax(tu.lo[2], lp.KMIN) + 0.1);
    lp.kmax = (int)(fmin(tu.hi[2], lp.KMAX) + 0.1);
    lp.pos = I;
    lp.map = h ? l.dmap : l.dmap + 4 * l.nodes;
    lp.tab = h ? l.dtab : l.dtab + 3 * LASER_CAP;
    lp.step = r->dstep;
    lp.cap = LASER_CAP;
    lp.on = lp.jmax >= lp.jmin && lp.kmax >= lp.kmin;
    return lp;
}

/* The box over which the H source applies (em_HStep's update box) and the
 * one of the E source. */
static Box laser_box_H(const mpic_run *r)
{
    return Box{{3, 3, 3}, {r->g.n[0] - 2, r->g.n[1] - 2, r->g.n[2] - 2}};
}
static Box laser_box_E(const mpic_run *r)
{
    return Box{{-1, -1, -1}, {r->g.n[0] + 1, r->g.n[1] + 1, r->g.n[2] + 1}};
}

/* The run's laser folded into the H step (h = 1) or the fused E update:
 * only with exactly one laser and no TFSF interface, whose source would
 * come between the update and the laser's. */
static LaserPlane laser_fold(const mpic_run *r, int h)
{
    if (r->lasers.size() != 1 || r->tfsf.on) return LaserPlane();
    return laser_plane(r, r->lasers[0], h, h ? laser_box_H(r) : laser_box_E(r));
}

template <class T>
static int laser_H(mpic_run *r, T *Fy, T *Fz, const Box &tu)
{
    for (const Laser &l : r->lasers)
        TRY(launch_laser(r->stream, Fy, Fz, r->g, laser_plane(r, l, 1, tu)));
    return MPIC_OK;
}

template <class T>
static int laser_E(mpic_run *r, T *Ey, T *Ez, const Box &tu)
{
    for (const Laser &l : r->lasers)
        TRY(launch_laser(r->stream, Ey, Ez, r->g, laser_plane(r, l, 0, tu)));
    return MPIC_OK;
}

/* em_HHalfStep; with 'part' the H update also writes em_energy's partial
 * sums of E and of the H it leaves (H^n). */
static int h_half(mpic_run *r, double *part, bool with_clear = false);

extern "C" int mpic_em_h_half(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    return h_half(r, nullptr);
}

/* The J clear of the push that follows the H half step; with_clear runs it
 * in the H update's launch (nothing between them reads or writes J). */
static int h_half(mpic_run *r, double *part, bool with_clear)
{
    TRY(check_stage(r, 0));
    MPIC_CHECK(hipSetDevice(r->dev));
    TRY(laser_tables(r, 1));
    if (r->trec.on) {                    /* TFSF_preHStep in record mode, source/core/em.c:92 */
        const TfsfRec &tr = r->trec;
        if (r->fp32)
            TRY(launch_tfsf_record<float>(r->stream, r->f32.E, r->f32.H, r->g, tr.fe, tr.fh, tr.maxvol,
                                          tr.dfac, tr.dbuf, r->dstep, (int)tr.step0, tr.cap));
        else
            TRY(launch_tfsf_record<double>(r->stream, r->f.E, r->f.H, r->g, tr.fe, tr.fh, tr.maxvol,
                                           tr.dfac, tr.dbuf, r->dstep, (int)tr.step0, tr.cap));
    }
    /* cap_cacheEH: cap H, walls on cap H, cap E, walls on cap E.  Sources
     * never reach the cap layer (checked in mpic_add_laser). */
    FCALL(launch_cap_H, r->g, r->updH);
    TRY(walls_on(r, r->bcH));
    FCALL(launch_cap_E, r->g, r->updE);
    TRY(walls_on(r, r->bcE));
    /* em_HSmallStep + cap_flushHalfH */
    if (with_clear) {
        FCALL(launch_h_update_jclear, r->g, r->flH, 0, part, r->jp.J[0], r->dacc + 2, r->djr,
              r->dstep, r->tiled ? r->tv.active : nullptr, r->tiled ? r->tv.nactive : nullptr,
              r->tiled ? r->tv.nt[1] * r->tv.nt[2] : 1, r->tiled ? 0 : 1);
        r->jclear_done = 1;
    } else {
        FCALL(launch_h_update, r->g, r->flH, 0, part);
    }
    r->stage = 1;
    return MPIC_OK;
}

extern "C" int mpic_em_energy(mpic_run_t *r, double *W_E, double *W_M)
{
    if (!r) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipMemsetAsync(r->dacc, 0, 2 * sizeof(double), r->stream));
    FCALL(launch_energy, r->g, r->dacc);
    MPIC_CHECK(hipMemcpyAsync(r->hacc, r->dacc, 2 * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    const double *h = r->g.h;
    if (W_E) *W_E = 0.5 * r->hacc[0] * h[0] * h[1] * h[2];
    if (W_M) *W_M = 0.5 * r->hacc[1] * h[0] * h[1] * h[2];
    return MPIC_OK;
}

/* Event k of the timed step r->timed. */
static int record(mpic_run *r, int k)
{
    MPIC_CHECK(hipEventRecord(r->evpool[5 * r->timed + k], r->stream));
    return MPIC_OK;
}

extern "C" int mpic_push_deposit(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    TRY(check_stage(r, 1));
    MPIC_CHECK(hipSetDevice(r->dev));
    /* J (the planes the last deposit wrote) and the temperature sums: a
     * kernel, not a memset (the step may be replayed from a graph), unless
     * the H half step's launch cleared them */
    if (!r->jclear_done)
        TRY(launch_jclear(r->stream, r->jp.J[0], r->g, r->dacc + 2, r->djr, r->dstep,
                          r->tiled ? r->tv.active : nullptr, r->tiled ? r->tv.nactive : nullptr,
                          r->tiled ? r->tv.nt[1] * r->tv.nt[2] : 1, r->tiled ? 0 : 1));
    r->jclear_done = 0;
    if (r->timed >= 0) TRY(record(r, 1));
    if (r->fp32) {
        TRY(launch_push32(r->stream, r->m32, r->m32buf[r->cur ^ 1], r->f32, r->g, r->pc, r->qm,
                          r->tv, r->tj, r->dacc + 2,
                          r->tile_blocks, r->tile_waves));
        if (r->timed >= 0) TRY(record(r, 2));
        TRY(launch_jgather32(r->stream, r->f32, r->g, r->tv, r->tj, r->gather_blocks));
    } else if (r->tiled) {
        TRY(launch_push64_tiled(r->stream, r->m, r->mbuf[r->cur ^ 1], r->f, r->g, r->pc, r->tv,
                                r->dacc + 2, r->derr,
                                r->push_blocks));
    } else {
        TRY(launch_push64(r->stream, r->m, r->n, r->f, r->g, r->pc, r->dacc + 2, 0, r->derr));
    }
    if (r->timed >= 0) {
        if (!r->fp32) TRY(record(r, 2));
        TRY(record(r, 3));
    }
    /* per-step marker sort by cell (keys written by the push) */
    if (r->tiled) TRY(sort_pass(r, 0));
    return MPIC_OK;
}

extern "C" int mpic_temperature(mpic_run_t *r, double *WTx, double *WTy, double *WTz)
{
    if (!r) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipMemcpyAsync(r->hacc + 2, r->dacc + 2, 3 * sizeof(double), hipMemcpyDeviceToHost,
                              r->stream));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    if (WTx) *WTx = r->hacc[2];
    if (WTy) *WTy = r->hacc[3];
    if (WTz) *WTz = r->hacc[4];
    return MPIC_OK;
}

/* The H step; with_fold also runs the y and z J folds in the H update's
 * launch (the H step neither reads nor writes J) and then the VSP sweep,
 * the rest of mpic_j_finish after its x folds. */
static int h_step(mpic_run *r, bool with_fold)
{
    TRY(check_stage(r, 1));
    MPIC_CHECK(hipSetDevice(r->dev));
    /* the laser's H source in the same pass where nothing comes between */
    const LaserPlane lf = laser_fold(r, 1);
    if (with_fold) {
        FCALL(launch_h_update_jfold, r->g, r->flH, 1, lf, r->jp, r->cfg.bc, r->djr + 4);
        if (fabs(r->vsp - 1.0) >= 1e-10) TRY(launch_vsp(r->stream, r->jp.J, 3, r->g, r->vsp));
    } else {
        FCALL(launch_h_update, r->g, r->flH, 1, nullptr, lf);
    }
    if (r->tfsf.on) {                    /* TFSF_postHStep, source/core/em.c:125 */
        const Tfsf &tf = r->tfsf;
        if (r->fp32)
            TRY(launch_tfsf(r->stream, r->f32.H, r->g, tf.fh, tf.maxvol, tf.ddata, tf.dtab,
                            r->dstep, LASER_CAP, 0, 1.0));
        else
            TRY(launch_tfsf(r->stream, r->f.H, r->g, tf.fh, tf.maxvol, tf.ddata, tf.dtab,
                            r->dstep, LASER_CAP, 0, 1.0));
    }
    if (!lf.on) {
        const Box inner = laser_box_H(r);
        if (r->fp32)
            TRY(laser_H(r, r->f32.H[1], r->f32.H[2], inner));
        else
            TRY(laser_H(r, r->f.H[1], r->f.H[2], inner));
    }
    r->stage = 2;
    return MPIC_OK;
}

extern "C" int mpic_em_h_step(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    return h_step(r, false);
}

extern "C" int mpic_em_e_start(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    TRY(check_stage(r, 2));
    MPIC_CHECK(hipSetDevice(r->dev));
    /* interior E step; the cap flush (which em_EStep_finish does) touches
     * disjoint nodes and reads only Ec, so it is done in the same pass */
    FCALL(launch_e_update, r->g, r->flE, 1, 1);
    r->stage = 3;
    return MPIC_OK;
}

/* jbc_finish's x walls: the first part of mpic_j_finish. */
static int j_fold_x(mpic_run *r)
{
    const int *bc = r->cfg.bc;
    return launch_jfold(r->stream, r->jp, r->g, 0, bc[0], bc[3], r->djr + 4);
}

extern "C" int mpic_j_finish(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    const int *bc = r->cfg.bc;
    /* x walls, then the y and z walls plane by plane */
    TRY(j_fold_x(r));
    TRY(launch_jfold_planes(r->stream, r->jp, r->g, bc, r->djr + 4));
    /* VSP_vectorSweep after the boundary fold (source/core/plasma_currentWalls.c:204) */
    if (fabs(r->vsp - 1.0) >= 1e-10) TRY(launch_vsp(r->stream, r->jp.J, 3, r->g, r->vsp));
    return MPIC_OK;
}

/* VSP_configure (source/core/plasma_VSP.c:18-27). */
extern "C" int mpic_set_vsp(mpic_run_t *r, double alpha)
{
    if (!r) return MPIC_ERR_ARG;
    if (!(alpha > -0.1 && alpha <= 1.1)) {
        mpic_set_error("mpic_set_vsp: weight %g outside (-0.1, 1.1]", alpha);
        return MPIC_ERR_ARG;
    }
    if (r->stage != 0) {
        mpic_set_error("mpic_set_vsp: set the weight between steps");
        return MPIC_ERR_ORDER;
    }
    r->vsp = alpha;
    invalidate(r);
    return MPIC_OK;
}

/* The sources em_EStep_finish adds to E: the TFSF interface
 * (TFSF_postEStep, source/core/em.c:200) and the lasers. */
static int e_sources(mpic_run *r, bool laser_done = false)
{
    if (r->tfsf.on) {                    /* TFSF_postEStep, source/core/em.c:200 */
        const Tfsf &tf = r->tfsf;
        const double sign = tf.mode == 0 ? 1.0 : -1.0;
        if (r->fp32)
            TRY(launch_tfsf(r->stream, r->f32.E, r->g, tf.fe, tf.maxvol, tf.ddata, tf.dtab,
                            r->dstep, LASER_CAP, 1, sign));
        else
            TRY(launch_tfsf(r->stream, r->f.E, r->g, tf.fe, tf.maxvol, tf.ddata, tf.dtab,
                            r->dstep, LASER_CAP, 1, sign));
    }
    if (laser_done) return MPIC_OK;     /* added by the fused E update */
    const Box dom = laser_box_E(r);
    if (r->fp32)
        TRY(laser_E(r, r->f32.E[1], r->f32.E[2], dom));
    else
        TRY(laser_E(r, r->f.E[1], r->f.E[2], dom));
    return MPIC_OK;
}

/* The hard spot after E -= tau J (gaussSpot_forceE). */
static int e_spot(mpic_run *r)
{
    if (r->spot.on) {
        const Spot &sp = r->spot;
        const long np = (long)(r->g.n[1] + 3) * (r->g.n[2] + 3);
        if (r->fp32)
            TRY(launch_spot(r->stream, r->f32.E[1], r->f32.E[2], r->g, sp.i, sp.dprof, sp.dprof + np,
                            sp.dtab, r->dstep, LASER_CAP));
        else
            TRY(launch_spot(r->stream, r->f.E[1], r->f.E[2], r->g, sp.i, sp.dprof, sp.dprof + np,
                            sp.dtab, r->dstep, LASER_CAP));
    }
    return MPIC_OK;
}

/* The host clock of a finished step. */
static void step_done(mpic_run *r)
{
    r->time = r->time + r->g.tau;          /* parameter_setTime (Time + tau) */
    ++r->step;
    r->stage = 0;
}

extern "C" int mpic_em_e_finish(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    TRY(check_stage(r, 3));
    MPIC_CHECK(hipSetDevice(r->dev));
    TRY(e_sources(r));
    /* E -= tau J; the step counter advances in the same launch unless the
     * spot source still has to read it */
    FCALL(launch_e_tauJ, r->g, r->spot.on ? nullptr : r->dstep, r->djr + 4);
    TRY(e_spot(r));
    if (r->spot.on) TRY(launch_tick(r->stream, r->dstep));
    step_done(r);
    return MPIC_OK;
}

/* The ring slot of the step about to run holds that step's row. */
static void mark_ring(mpic_run *r)
{
    const long slot = r->step % DIAG_RING;
    r->ring_step[slot] = r->step;
    r->ring_time[slot] = r->time;
}

/* One step; with r->timed >= 0 its events (before the push, after it,
 * after the deposit: slots 1, 2, 3 of the step's five) are recorded. */
static int step_body(mpic_run *r)
{
    if (!r->capturing) mark_ring(r);
    /* em_energy after the H half step (its sums come out of the H update),
     * plasma_temperature after the push (source/core/main.c:371,383): one
     * ring row per step, nothing copied back, so the step stays
     * asynchronous */
    /* the J clear joins the H half step's launch, the y and z J folds the
     * H step's */
    TRY(h_half(r, r->dpart, true));
    TRY(mpic_push_deposit(r));
    /* jbc_finish reads and writes J only and the H step does not read J:
     * the fold goes first, so that E -= tau J can join the E update */
    TRY(j_fold_x(r));
    TRY(h_step(r, true));
    /* one pass for em_EStep_start, the laser's E plane and E -= tau J where
     * no other source adds to E between them (FP64 form: no TFSF interface,
     * at most one laser, which the pass adds in the reference's order),
     * always in the FP32 split form (its TFSF and further laser sources
     * then add after E -= tau J) */
    const bool fused = r->fp32 || (!r->tfsf.on && r->lasers.size() <= 1);
    if (fused) {
        const LaserPlane lf = laser_fold(r, 0);
        FCALL(launch_e_fused, r->g, r->flE, r->djr + 4, lf);
        TRY(e_sources(r, lf.on != 0));
    } else {
        TRY(mpic_em_e_start(r));
        TRY(e_sources(r));
        FCALL(launch_e_tauJ, r->g, nullptr, r->djr + 4);
    }
    TRY(e_spot(r));
    /* the ring row of this step, then the device step counter advances */
    TRY(launch_diag_ring(r->stream, r->dpart, r->nparts, r->dacc + 2, r->dring, r->dstep,
                         DIAG_RING, 1));
    step_done(r);
    return MPIC_OK;
}

static int step_async(mpic_run *r, int idx)
{
    r->timed = idx;
    const int rc = step_body(r);
    r->timed = -1;
    return rc;
}

extern "C" int mpic_step(mpic_run_t *r, double *diag)
{
    if (!r) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    if (!diag) return step_async(r, -1);
    const double t0 = r->time;
    double WE, WM;
    TRY(mpic_em_h_half(r));
    TRY(mpic_em_energy(r, &WE, &WM));
    TRY(mpic_push_deposit(r));
    TRY(mpic_temperature(r, diag + 3, diag + 4, diag + 5));
    TRY(mpic_em_h_step(r));
    TRY(mpic_em_e_start(r));
    TRY(mpic_j_finish(r));
    TRY(mpic_em_e_finish(r));
    diag[0] = t0;
    diag[1] = WE;
    diag[2] = WM;
    return MPIC_OK;
}

/* n steps with a timed event record per step (kernel timing). */
static int advance_timed(mpic_run *r, int nsteps)
{
    while ((int)r->evpool.size() < 5 * nsteps) {
        hipEvent_t e;
        MPIC_CHECK(hipEventCreate(&e));
        r->evpool.push_back(e);
    }
    for (int s = 0; s < nsteps; ++s) TRY(step_async(r, s));
    r->ev_steps = nsteps;
    return MPIC_OK;
}

/* Host bookkeeping of a step that a replayed graph performs on the device. */
static void host_step(mpic_run *r)
{
    mark_ring(r);
    if (r->tiled) swap_markers(r);
    r->time = r->time + r->g.tau;
    ++r->step;
    r->stage = 0;
}

static int record_steps(mpic_run *r, Recorder rec[2]);

/* The recorded launches as a chain of kernel nodes, instantiated. */
static int graph_of(const Recorder &rec, hipGraphExec_t *exec)
{
    hipGraph_t gr = nullptr;
    MPIC_CHECK(hipGraphCreate(&gr, 0));
    hipGraphNode_t prev = nullptr;
    int rc = MPIC_OK;
    for (const LaunchRec &x : rec.recs) {
        hipGraphNode_t node = nullptr;
        rc = x.one_node(gr, prev, &node, x.args.data(), x.grid);
        if (rc != MPIC_OK) break;
        prev = node;
    }
    if (rc == MPIC_OK && hipGraphInstantiate(exec, gr, nullptr, nullptr, 0) != hipSuccess) {
        mpic_set_error("hipGraphInstantiate failed");
        rc = MPIC_ERR_DEVICE;
    }
    hipGraphDestroy(gr);
    return rc;
}

/* Graphs of one step from each marker-buffer parity in exec[0..1]: the
 * kernel arguments of a step differ only by which buffer holds the sorted
 * markers, so two graphs replay every step.  The step is recorded, not run,
 * and the host state is restored afterwards. */
static int capture(mpic_run *r)
{
    if (!r->tiled || !r->inited || r->stage != 0) {
        mpic_set_error("graph capture needs an initialised, tiled run between steps");
        return MPIC_ERR_STATE;
    }
    TRY(laser_tables(r, 2));
    for (int k = 0; k < 2; ++k)
        if (r->exec[k]) {
            hipGraphExecDestroy(r->exec[k]);
            r->exec[k] = nullptr;
        }
    r->exec_parity = r->parity;
    Recorder two[2];
    int rc = record_steps(r, two);
    for (int k = 0; k < 2 && rc == MPIC_OK; ++k) rc = graph_of(two[k], &r->exec[k]);
    r->graphs_ok = rc == MPIC_OK;
    return rc;
}

static int replay(mpic_run *r)
{
    const int k = (r->parity == r->exec_parity) ? 0 : 1;
    MPIC_CHECK(hipGraphLaunch(r->exec[k], r->stream));
    host_step(r);
    return MPIC_OK;
}

extern "C" int mpic_set_graphs(mpic_run_t *r, int on)
{
    if (!r) return MPIC_ERR_ARG;
    if (on && !r->tiled) {
        mpic_set_error("graph replay needs the sorted (tiled) push");
        return MPIC_ERR_ARG;
    }
    r->use_graphs = on != 0;
    return MPIC_OK;
}

static int batch_streams(mpic_run *const *runs, int nruns, long chunk);

extern "C" int mpic_advance(mpic_run_t *r, int nsteps)
{
    if (!r || nsteps < 0) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    if (!r->use_graphs) return advance_timed(r, nsteps);
    if (!r->inited || r->stage != 0) {
        mpic_set_error("stage order violated: run is at stage %d, a replayed step starts at "
                       "stage 0", r->stage);
        return r->inited ? MPIC_ERR_ORDER : MPIC_ERR_STATE;
    }
    /* the run's own captured step, replayed on its stream */
    r->ev_steps = 0;
    mpic_run *one[1] = {r};
    for (long left = nsteps; left > 0;) {
        const long chunk = left < LASER_CAP - 2 ? left : LASER_CAP - 2;
        TRY(laser_tables(r, chunk + 2));
        TRY(batch_streams(one, 1, chunk));
        left -= chunk;
    }
    return MPIC_OK;
}

/* ---------------- fused ensembles ---------------- */
/* One launch of the fused step: the runs' launches of one kind as one
 * k_multi launch (or the run's own launch when only one run has it). */
struct FusedLaunch {
    const LaunchRec *rec;          /* the first run's record (kind, launchers) */
    int              n;            /* runs in the launch */
    unsigned         blocks;       /* total blocks */
    size_t           args_off, tab_off;   /* in the device block */
};

/* The fused step of a set of runs: per marker-buffer phase (the runs
 * double-buffer their markers, so the step's arguments alternate) the
 * launches in order and a captured graph of them. */
struct Ensemble {
    std::vector<mpic_run *> runs;
    std::vector<long>       topo;      /* runs' topo counters when built */
    std::vector<int>        rel;       /* parity of run q relative to run 0 */
    int                     dev;
    int                     parity0;   /* run 0's parity at phase 0 */
    hipStream_t             stream;
    void                   *dmem;
    hipGraphExec_t          exec[2];
    int                     nlaunch[2];
    std::vector<hipEvent_t> ev;        /* one per run, plus the ensemble's */
};

static Ensemble *g_ens = nullptr;        /* the last fused ensemble */
static std::mutex g_ens_mu;

static void ensemble_free(Ensemble *e)
{
    if (!e) return;
    hipSetDevice(e->dev);
    if (e->stream) hipStreamSynchronize(e->stream);
    for (int k = 0; k < 2; ++k)
        if (e->exec[k]) hipGraphExecDestroy(e->exec[k]);
    hipFree(e->dmem);
    for (hipEvent_t v : e->ev) hipEventDestroy(v);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
}

/* A run is destroyed: a fused ensemble holding it goes with it. */
static void ensemble_forget(mpic_run *r)
{
    std::lock_guard<std::mutex> lk(g_ens_mu);
    if (!g_ens) return;
    for (mpic_run *q : g_ens->runs)
        if (q == r) {
            ensemble_free(g_ens);
            g_ens = nullptr;
            return;
        }
}

static bool ensemble_matches(const Ensemble *e, mpic_run *const *runs, int nruns)
{
    if (!e || (int)e->runs.size() != nruns) return false;
    for (int q = 0; q < nruns; ++q)
        if (e->runs[q] != runs[q] || e->topo[q] != runs[q]->topo ||
            e->rel[q] != (runs[q]->parity ^ runs[0]->parity))
            return false;
    return true;
}

/* Records one step of run r at each marker-buffer phase (k = 0: its current
 * parity) as launch data; the host state is restored afterwards. */
static int record_steps(mpic_run *r, Recorder rec[2])
{
    const double time = r->time;
    const long step = r->step;
    const int cur = r->cur, parity = r->parity;
    const Markers32 m32 = r->m32;
    const Markers64 m = r->m;
    int rc = MPIC_OK;
    for (int k = 0; k < 2 && rc == MPIC_OK; ++k) {
        g_recorder = &rec[k];
        r->capturing = 1;
        rc = step_async(r, -1);
        r->capturing = 0;
        g_recorder = nullptr;
    }
    r->time = time;
    r->step = step;
    r->cur = cur;
    r->parity = parity;
    r->m32 = m32;
    r->m = m;
    r->stage = 0;
    return rc;
}

/* Orders the runs' launches into fused launches: repeatedly the kind that
 * is next for the most runs is launched for all of them, so every run's
 * launches keep their order (runs are independent of each other). */
static void fuse(const std::vector<Recorder> &recs, std::vector<std::vector<const LaunchRec *>> &out)
{
    const int n = (int)recs.size();
    std::vector<size_t> at(n, 0);
    for (;;) {
        const void *best = nullptr;
        int bestc = 0;
        for (int q = 0; q < n; ++q) {
            if (at[q] >= recs[q].recs.size()) continue;
            const void *k = recs[q].recs[at[q]].kind;
            int c = 0;
            for (int u = 0; u < n; ++u)
                if (at[u] < recs[u].recs.size() && recs[u].recs[at[u]].kind == k) ++c;
            if (c > bestc) {
                bestc = c;
                best = k;
            }
        }
        if (!best) return;
        std::vector<const LaunchRec *> grp;
        for (int q = 0; q < n; ++q)
            if (at[q] < recs[q].recs.size() && recs[q].recs[at[q]].kind == best)
                grp.push_back(&recs[q].recs[at[q]++]);
        out.push_back(grp);
    }
}

static int ensemble_build(mpic_run *const *runs, int nruns, Ensemble **out)
{
    Ensemble *e = new Ensemble();
    e->dev = runs[0]->dev;
    e->parity0 = runs[0]->parity;
    for (int q = 0; q < nruns; ++q) {
        e->runs.push_back(runs[q]);
        e->topo.push_back(runs[q]->topo);
        e->rel.push_back(runs[q]->parity ^ runs[0]->parity);
    }
    int rc = MPIC_OK;
    std::vector<Recorder> recs[2];
    recs[0].resize(nruns);
    recs[1].resize(nruns);
    /* The FP64 tiled push of a run in the fused launch takes 2 workgroups
     * per CU (at most its tiles), not the 3 that fit: the runs' workgroups
     * then overlap more in the one launch (64-run scan ensemble: 7.30 ms per
     * FP64 step against 7.78 ms; a share of the resident workgroups in
     * proportion to the run's markers measured 8.07 ms). */
    std::vector<int> own_blocks(nruns);
    for (int q = 0; q < nruns; ++q) {
        mpic_run *r = runs[q];
        own_blocks[q] = r->push_blocks;
        if (!r->fp32 && r->tiled && r->push_blocks > 2 * r->ncu) r->push_blocks = 2 * r->ncu;
    }
    for (int q = 0; q < nruns && rc == MPIC_OK; ++q) {
        Recorder two[2];
        rc = record_steps(runs[q], two);
        recs[0][q] = std::move(two[0]);
        recs[1][q] = std::move(two[1]);
    }
    for (int q = 0; q < nruns; ++q) runs[q]->push_blocks = own_blocks[q];
    std::vector<std::vector<const LaunchRec *>> seq[2];
    std::vector<FusedLaunch> fl[2];
    size_t bytes = 0;
    for (int k = 0; k < 2 && rc == MPIC_OK; ++k) {
        fuse(recs[k], seq[k]);
        for (auto &grp : seq[k]) {
            FusedLaunch f;
            f.rec = grp[0];
            f.n = (int)grp.size();
            f.blocks = 0;
            for (const LaunchRec *x : grp) f.blocks += x->grid[0] * x->grid[1] * x->grid[2];
            bytes = (bytes + 255) & ~(size_t)255;
            f.args_off = bytes;
            bytes += grp[0]->args.size() * grp.size();
            bytes = (bytes + 255) & ~(size_t)255;
            f.tab_off = bytes;
            bytes += sizeof(unsigned) * (4 * grp.size() + 1);
            fl[k].push_back(f);
        }
        e->nlaunch[k] = (int)fl[k].size();
    }
    std::vector<unsigned char> host(bytes ? bytes : 1);
    for (int k = 0; k < 2 && rc == MPIC_OK; ++k)
        for (size_t u = 0; u < fl[k].size(); ++u) {
            const auto &grp = seq[k][u];
            const FusedLaunch &f = fl[k][u];
            const size_t as = grp[0]->args.size();
            unsigned *tab = (unsigned *)(host.data() + f.tab_off);
            unsigned acc = 0;
            for (size_t q = 0; q < grp.size(); ++q) {
                memcpy(host.data() + f.args_off + q * as, grp[q]->args.data(), as);
                tab[q] = acc;
                for (int d = 0; d < 3; ++d) tab[grp.size() + 1 + 3 * q + d] = grp[q]->grid[d];
                acc += grp[q]->grid[0] * grp[q]->grid[1] * grp[q]->grid[2];
            }
            tab[grp.size()] = acc;
        }
    if (rc == MPIC_OK && hipSetDevice(e->dev) != hipSuccess) rc = MPIC_ERR_DEVICE;
    if (rc == MPIC_OK && (hipMalloc(&e->dmem, host.size()) != hipSuccess ||
                          hipMemcpy(e->dmem, host.data(), host.size(), hipMemcpyHostToDevice) !=
                              hipSuccess ||
                          hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)) {
        mpic_set_error("fused ensemble: device set-up failed");
        rc = MPIC_ERR_DEVICE;
    }
    for (int q = 0; q <= nruns && rc == MPIC_OK; ++q) {
        hipEvent_t v;
        if (hipEventCreateWithFlags(&v, hipEventDisableTiming) != hipSuccess) {
            mpic_set_error("fused ensemble: event creation failed");
            rc = MPIC_ERR_DEVICE;
            break;
        }
        e->ev.push_back(v);
    }
    /* one graph per phase: the fused launches in order, as kernel nodes */
    for (int k = 0; k < 2 && rc == MPIC_OK; ++k) {
        hipGraph_t gr = nullptr;
        if (hipGraphCreate(&gr, 0) != hipSuccess) {
            mpic_set_error("fused ensemble: hipGraphCreate failed");
            rc = MPIC_ERR_DEVICE;
            break;
        }
        hipGraphNode_t prev = nullptr;
        for (size_t u = 0; u < fl[k].size() && rc == MPIC_OK; ++u) {
            const FusedLaunch &f = fl[k][u];
            const unsigned char *base = (const unsigned char *)e->dmem;
            hipGraphNode_t node = nullptr;
            if (f.n == 1)
                rc = f.rec->one_node(gr, prev, &node, f.rec->args.data(), f.rec->grid);
            else
                rc = f.rec->multi_node(gr, prev, &node, base + f.args_off,
                                       (const unsigned *)(base + f.tab_off), f.n, f.blocks);
            prev = node;
        }
        if (rc == MPIC_OK && hipGraphInstantiate(&e->exec[k], gr, nullptr, nullptr, 0) != hipSuccess) {
            mpic_set_error("fused ensemble: hipGraphInstantiate failed");
            rc = MPIC_ERR_DEVICE;
        }
        hipGraphDestroy(gr);
    }
    if (rc != MPIC_OK) {
        ensemble_free(e);
        return rc;
    }
    *out = e;
    return MPIC_OK;
}

/* The runs' steps as replays of per-run graphs, each run on its own stream
 * (MPIC_BATCH_STREAMS=1, and runs on several devices). */
static int batch_streams(mpic_run *const *runs, int nruns, long chunk)
{
    for (int q = 0; q < nruns; ++q)
        if (!runs[q]->graphs_ok) TRY(capture(runs[q]));
    int dev = -1;
    for (long s = 0; s < chunk; ++s)
        for (int q = 0; q < nruns; ++q) {
            mpic_run *r = runs[q];
            if (r->dev != dev) {
                MPIC_CHECK(hipSetDevice(r->dev));
                dev = r->dev;
            }
            TRY(replay(r));
        }
    return MPIC_OK;
}

/* The runs' steps as replays of the fused ensemble graph on the ensemble's
 * stream, ordered after the work already on the runs' streams and before
 * whatever they get next. */
static int batch_fused(mpic_run *const *runs, int nruns, long chunk)
{
    std::lock_guard<std::mutex> lk(g_ens_mu);
    if (!ensemble_matches(g_ens, runs, nruns)) {
        ensemble_free(g_ens);
        g_ens = nullptr;
        TRY(ensemble_build(runs, nruns, &g_ens));
    }
    Ensemble *e = g_ens;
    MPIC_CHECK(hipSetDevice(e->dev));
    for (int q = 0; q < nruns; ++q) {
        MPIC_CHECK(hipEventRecord(e->ev[q], runs[q]->stream));
        MPIC_CHECK(hipStreamWaitEvent(e->stream, e->ev[q], 0));
    }
    for (long s = 0; s < chunk; ++s) {
        const int k = runs[0]->parity == e->parity0 ? 0 : 1;
        MPIC_CHECK(hipGraphLaunch(e->exec[k], e->stream));
        for (int q = 0; q < nruns; ++q) host_step(runs[q]);
    }
    MPIC_CHECK(hipEventRecord(e->ev[nruns], e->stream));
    for (int q = 0; q < nruns; ++q) MPIC_CHECK(hipStreamWaitEvent(runs[q]->stream, e->ev[nruns], 0));
    return MPIC_OK;
}

extern "C" int mpic_batch_advance(mpic_run_t *const *runs, int nruns, int nsteps)
{
    if (!runs || nruns < 0 || nsteps < 0) return MPIC_ERR_ARG;
    bool one_dev = true;
    for (int q = 0; q < nruns; ++q) {
        mpic_run *r = runs[q];
        if (!r || !r->inited || !r->tiled) {
            mpic_set_error("mpic_batch_advance: run %d is not an initialised tiled run", q);
            return MPIC_ERR_ARG;
        }
        /* a replay runs whole steps: a step begun through the stage calls
         * has to be finished through them first */
        if (r->stage != 0) {
            mpic_set_error("stage order violated: run %d is at stage %d, a replayed step "
                           "starts at stage 0", q, r->stage);
            return MPIC_ERR_ORDER;
        }
        for (int u = 0; u < q; ++u)
            if (runs[u] == r) {
                mpic_set_error("mpic_batch_advance: run %d is listed twice", q);
                return MPIC_ERR_ARG;
            }
        one_dev = one_dev && r->dev == runs[0]->dev;
    }
    if (nruns == 0 || nsteps == 0) return MPIC_OK;
    const char *env = getenv("MPIC_BATCH_STREAMS");
    const bool streams = !one_dev || (env && env[0] == '1');
    for (int q = 0; q < nruns; ++q) runs[q]->ev_steps = 0;
    for (long left = nsteps; left > 0;) {
        const long chunk = left < LASER_CAP - 2 ? left : LASER_CAP - 2;
        for (int q = 0; q < nruns; ++q) {
            MPIC_CHECK(hipSetDevice(runs[q]->dev));
            TRY(laser_tables(runs[q], chunk + 2));
        }
        TRY(streams ? batch_streams(runs, nruns, chunk) : batch_fused(runs, nruns, chunk));
        left -= chunk;
    }
    return MPIC_OK;
}

/* Work-item list and FP32 J blocks for items of at most chm markers (the
 * buffers only grow; graphs and fused ensembles holding the old ones are
 * dropped). */
static int ensure_items(mpic_run *r, int chm)
{
    const long need = items_capacity(r->tv.ntiles, r->cap, chm);
    if (need <= r->items_cap) return MPIC_OK;
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    int *item = nullptr;
    float *tj = nullptr;
    if (hipMalloc(&item, sizeof(int) * need) != hipSuccess ||
        hipMalloc(&tj, sizeof(float) * TJB * (size_t)need) != hipSuccess) {
        hipFree(item);
        mpic_set_error("work items of %d markers: %ld items do not fit on the device", chm, need);
        return MPIC_ERR_CAPACITY;
    }
    hipFree(r->tv.item);
    hipFree(r->tj);
    r->tv.item = item;
    r->tj = tj;
    r->items_cap = need;
    invalidate(r);
    return MPIC_OK;
}

/* The FP32 push's launch shape: markers per work item (a tile's markers
 * split into even items of at most that many, or one item per tile),
 * wavefronts per workgroup (one workgroup per item at a time) and workgroup
 * count.  Each candidate runs 'steps' real steps, timed with HIP events
 * (push + J gather); item sizes whose J blocks would take more than an
 * eighth of the marker buffers' memory are not tried. */
extern "C" int mpic_autotune(mpic_run_t *r, int steps, int *chosen)
{
    if (!r || steps < 1) return MPIC_ERR_ARG;
    if (!r->inited || r->stage != 0) {
        mpic_set_error("mpic_autotune: the run must be initialised and between steps");
        return r->inited ? MPIC_ERR_ORDER : MPIC_ERR_STATE;
    }
    MPIC_CHECK(hipSetDevice(r->dev));
    if (!r->fp32 || !r->tiled) {
        if (chosen) {
            chosen[0] = r->push_blocks;
            chosen[1] = 1;
            chosen[2] = 0;
        }
        return MPIC_OK;
    }
    const int chms[3] = {0, 2048, 1024};
    const int wavess[3] = {1, 2, 4};
    const char *verbose = getenv("MPIC_AUTOTUNE_VERBOSE");
    int best_nb = r->tile_blocks, best_w = r->tile_waves, best_chm = r->tv.chm;
    double best_ms = 1e300;
    for (int c = 0; c < 3; ++c) {
        const int chm = chms[c];
        if (chm > 0 && (double)TJB * sizeof(float) / chm > 0.125 * 2 * 36.0) continue;
        if (ensure_items(r, chm) != MPIC_OK) continue;
        /* the item list of the current order follows the item size */
        r->tv.chm = chm;
        TRY(sort_pass(r, 1));
        TRY(advance_timed(r, 1));          /* untimed: first launches of the new lists */
        for (int wi = 0; wi < 3; ++wi) {
            const int w = wavess[wi];
            for (int per_cu = 4; per_cu <= 64; per_cu *= 2) {
                int nb = per_cu * r->ncu / w;
                if (nb > r->items_cap) nb = (int)r->items_cap;
                r->tile_blocks = nb;
                r->tile_waves = w;
                TRY(advance_timed(r, steps));
                double p, d, rest;
                TRY(mpic_last_timing(r, &p, &d, &rest));
                if (verbose && verbose[0] == '1')
                    fprintf(stderr, "mpic_autotune: %d markers per item, %d x %d wavefronts: push "
                            "%.1f us, J gather %.1f us per step\n", chm, nb, w, 1e3 * p / steps,
                            1e3 * d / steps);
                if (p + d < best_ms) {
                    best_ms = p + d;
                    best_nb = nb;
                    best_w = w;
                    best_chm = chm;
                }
                if (nb == r->items_cap) break;
            }
        }
    }
    r->tv.chm = best_chm;
    r->tile_blocks = best_nb;
    r->tile_waves = best_w;
    TRY(sort_pass(r, 1));
    invalidate(r);
    r->ev_steps = 0;
    if (chosen) {
        chosen[0] = best_nb;
        chosen[1] = best_w;
        chosen[2] = best_chm;
    }
    return MPIC_OK;
}

static int check_flags(mpic_run *r)
{
    unsigned f = 0;
    MPIC_CHECK(hipMemcpy(&f, r->derr, sizeof(unsigned), hipMemcpyDeviceToHost));
    if (f & 2u) {
        mpic_set_error("the tile scan's wait for an earlier workgroup ran out: the marker order "
                       "is invalid (flags %#x)", f);
        return MPIC_ERR_DEVICE;
    }
    if (f) {
        mpic_set_error("markers outside the grid were deleted (flags %#x)", f);
        return MPIC_ERR_STATE;
    }
    return MPIC_OK;
}

extern "C" int mpic_sync(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    return check_flags(r);
}

extern "C" double mpic_time(mpic_run_t *r) { return r ? r->time : NAN; }

extern "C" int mpic_last_timing(mpic_run_t *r, double *push_ms, double *deposit_ms,
                                double *rest_ms)
{
    if (!r) return MPIC_ERR_ARG;
    double t[3] = {0, 0, 0};
    const int n = r->ev_steps;
    if (n > 0) {
        MPIC_CHECK(hipSetDevice(r->dev));
        MPIC_CHECK(hipEventSynchronize(r->evpool[5 * (n - 1) + 3]));
        for (int s = 0; s < n; ++s) {
            const hipEvent_t *e = &r->evpool[5 * s];
            float b, c;
            MPIC_CHECK(hipEventElapsedTime(&b, e[1], e[2]));
            MPIC_CHECK(hipEventElapsedTime(&c, e[2], e[3]));
            t[0] += b;
            t[1] += c;
            if (s + 1 < n) {              /* from this step's deposit end to the next push */
                float a;
                MPIC_CHECK(hipEventElapsedTime(&a, e[3], r->evpool[5 * (s + 1) + 1]));
                t[2] += a;
            }
        }
        if (n > 1) t[2] *= (double)n / (n - 1);
    }
    if (push_ms) *push_ms = t[0];
    if (deposit_ms) *deposit_ms = t[1];
    if (rest_ms) *rest_ms = t[2];
    return MPIC_OK;
}

/* wDensity_addPoint's rows (source/core/diag_WDensity.c:136) of the steps
 * the loop body ran (mpic_step without diag, mpic_advance, replays): the
 * last min(max_rows, DIAG_RING) such steps, oldest first, each {Time, W_E,
 * W_M, WTx, WTy, WTz}. */
extern "C" int mpic_diag_history(mpic_run_t *r, double *rows, int max_rows, int *nrows)
{
    if (!r || !nrows || max_rows < 0 || (max_rows && !rows)) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    std::vector<double> ring(5 * DIAG_RING);
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    MPIC_CHECK(hipMemcpy(ring.data(), r->dring, sizeof(double) * ring.size(),
                         hipMemcpyDeviceToHost));
    const double vol = 0.5 * r->g.h[0] * r->g.h[1] * r->g.h[2];
    std::vector<long> steps;
    for (long s = r->step - 1; s >= 0 && s >= r->step - DIAG_RING && (int)steps.size() < max_rows;
         --s)
        if (r->ring_step[s % DIAG_RING] == s) steps.push_back(s);
    const int m = (int)steps.size();
    for (int q = 0; q < m; ++q) {
        const long s = steps[m - 1 - q], slot = s % DIAG_RING;
        double *o = rows + 6 * q;
        const double *v = ring.data() + 5 * slot;
        o[0] = r->ring_time[slot];
        o[1] = vol * v[0];
        o[2] = vol * v[1];
        o[3] = v[2];
        o[4] = v[3];
        o[5] = v[4];
    }
    *nrows = m;
    return MPIC_OK;
}

/* ---------------- rho, Gauss residuals, cells ---------------- */
/* The tile view for kernels that visit the markers of the order: before
 * mpic_init the loaded records are the order (no permutation yet). */
static TileView order_view(const mpic_run *r)
{
    TileView tv = r->tv;
    if (!r->inited || !r->tiled) tv.perm = nullptr;
    return tv;
}

static int rho_into(mpic_run *r, double *rho)
{
    long n = 0;
    TRY(live_n(r, &n));
    TRY(launch_zero(r->stream, rho, r->g.ncell));
    if (r->fp32)
        TRY(launch_rho32(r->stream, r->m32, n, rho, r->g, order_view(r), r->parity));
    else
        TRY(launch_rho64(r->stream, r->m, n, rho, r->g, order_view(r).perm, r->tv.nmark, r->parity));
    const int *bc = r->cfg.bc;
    const int BCv[6] = {bc[0], bc[3], bc[1], bc[4], bc[2], bc[5]};
    for (int b = 0; b < 6; ++b) TRY(launch_rho_fold(r->stream, rho, r->g, b, BCv[b]));
    /* VSP_scalarSweep after the boundary fold (source/core/plasma_rho.c:153) */
    if (fabs(r->vsp - 1.0) >= 1e-10) {
        double *F[3] = {rho, nullptr, nullptr};
        TRY(launch_vsp(r->stream, F, 1, r->g, r->vsp));
    }
    return MPIC_OK;
}

extern "C" int mpic_rho(mpic_run_t *r, double *rho)
{
    if (!r || !rho) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    TRY(rho_into(r, r->rho2));
    const Grid &g = r->g;
    std::vector<double> tmp((size_t)g.ncell);
    MPIC_CHECK(hipMemcpyAsync(tmp.data(), r->rho2, sizeof(double) * g.ncell, hipMemcpyDeviceToHost,
                              r->stream));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    long q = 0;
    for (int i = -2; i <= g.n[0] + 2; ++i)
        for (int j = -2; j <= g.n[1] + 2; ++j)
            for (int k = -2; k <= g.n[2] + 2; ++k, ++q) rho[q] = tmp[gidx(g, i, j, k)];
    return MPIC_OK;
}

extern "C" int mpic_gauss_before(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    return rho_into(r, r->rho1);
}

extern "C" int mpic_gauss_after(mpic_run_t *r, double *res)
{
    if (!r || !res) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    TRY(rho_into(r, r->rho2));
    MPIC_CHECK(hipMemsetAsync(r->dacc + 5, 0, 2 * sizeof(double), r->stream));
    FCALL(launch_gauss, r->rho1, r->rho2, r->g, r->dacc + 5);
    MPIC_CHECK(hipMemcpyAsync(r->hacc + 5, r->dacc + 5, 2 * sizeof(double), hipMemcpyDeviceToHost,
                              r->stream));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    res[0] = r->hacc[5];
    res[1] = r->hacc[6];
    return MPIC_OK;
}

extern "C" int mpic_marker_cells(mpic_run_t *r, int32_t *cells, long *n)
{
    if (!r || !cells || !n) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    long N = 0;
    TRY(live_n(r, &N));
    int32_t *d = nullptr;
    MPIC_CHECK(hipMallocAsync((void **)&d, sizeof(int32_t) * (N ? N : 1), r->stream));
    if (r->fp32)
        TRY(launch_cells32(r->stream, r->m32, N, r->g, order_view(r), r->parity, d));
    else
        TRY(launch_cells64(r->stream, r->m, N, r->g, order_view(r).perm, r->tv.nmark, r->parity, d));
    std::vector<int32_t> all(N);
    MPIC_CHECK(hipMemcpyAsync(all.data(), d, sizeof(int32_t) * N, hipMemcpyDeviceToHost,
                              r->stream));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    MPIC_CHECK(hipFreeAsync(d, r->stream));
    long o = 0;
    for (long q = 0; q < N; ++q) {
        if (all[q] < 0) continue;
        if (o >= *n) return MPIC_ERR_CAPACITY;
        cells[o++] = all[q];
    }
    *n = o;
    return MPIC_OK;
}

/* mpic_sort with the per-cell counts staged in 'stage' (device memory of at
 * least n0*n1*n2 int32, or nullptr: one is allocated for the call). */
static int sort_counts(mpic_run *r, int32_t *counts, int32_t *stage)
{
    if (!r->inited) {
        mpic_set_error("mpic_sort: run is not initialised");
        return MPIC_ERR_STATE;
    }
    MPIC_CHECK(hipSetDevice(r->dev));
    if (!r->tiled) {
        long n = r->n;
        const int nn = (int)n;
        MPIC_CHECK(hipMemcpyAsync(r->tv.nmark + r->parity, &nn, sizeof(int), hipMemcpyHostToDevice,
                                  r->stream));
    }
    TRY(sort_pass(r, 1));
    if (!r->tiled) {
        /* the untiled push keeps working on the compacted array */
        int v = 0;
        MPIC_CHECK(hipStreamSynchronize(r->stream));
        MPIC_CHECK(hipMemcpy(&v, r->tv.nmark + r->parity, sizeof(int), hipMemcpyDeviceToHost));
        r->n = v;
    }
    if (counts) {
        const long nc = (long)r->g.n[0] * r->g.n[1] * r->g.n[2];
        int32_t *d = stage;
        if (!d) MPIC_CHECK(hipMallocAsync((void **)&d, sizeof(int32_t) * nc, r->stream));
        TRY(launch_counts_natural(r->stream, r->tv, r->g, d));
        MPIC_CHECK(hipMemcpyAsync(counts, d, sizeof(int32_t) * nc, hipMemcpyDeviceToHost,
                                  r->stream));
        MPIC_CHECK(hipStreamSynchronize(r->stream));
        if (!stage) MPIC_CHECK(hipFreeAsync(d, r->stream));
    }
    return MPIC_OK;
}

extern "C" int mpic_sort(mpic_run_t *r, int32_t *counts)
{
    if (!r) return MPIC_ERR_ARG;
    return sort_counts(r, counts, nullptr);
}

