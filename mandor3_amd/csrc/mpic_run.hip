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

#include <vector>

#include "mpic_internal.h"

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

struct Laser {
    double  A, T0, delay, dt, phi0;
    int     forward, plane;
    double *dmap;             /* device: E map then H map */
    long    nodes;
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
    Fields64      f;
    double       *fblock;     /* all field arrays */
    double       *rho1, *rho2;
    Markers64     m;
    long          n, cap;
    CapBoxes      cH, cE;
    BoxSet        updH, updE, flH, flE;
    std::vector<FaceOp> bcH, bcE;
    std::vector<double *> murbuf;
    std::vector<Laser> lasers;
    PushConsts    pc;
    double       *dacc;       /* device scratch: [0..1] energy, [2..4] W, [5..6] gauss */
    double       *hacc;       /* pinned host mirror */
    unsigned     *derr;       /* device error flags (bit 0: marker off the grid) */
    int           stage;
    int           inited;
    int           timed;      /* record push-kernel events (mpic_advance) */
    double        time;
    hipEvent_t    ev[4];
    double        push_ms, field_ms;
};

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

static FaceOp face(int kind, int ea, int top, const Box &ext, double *const X[3], int N)
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
    for (int c = 0; c < 3; ++c) op.X[c] = X[c];
    return op;
}

/* Wall sequences of cap_cacheH / cap_cacheEH (source/core/em_caps.c:470-553,
 * 631-721): x, y, z; low side then high side. */
static int build_wall_ops(mpic_run *r)
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
                r->bcH.push_back(face(FACE_PERIODIC_H, ea, 0, eH, r->f.Hc, N));
                r->bcE.push_back(face(FACE_PERIODIC_E, ea, 0, eE, r->f.Ec, N));
            } else if (b == MPIC_BC_MIRROR) {
                FaceOp h = face(FACE_WALL_H, ea, top, eH, r->f.Hc, N);
                h.zeroN = 1;
                r->bcH.push_back(h);
                r->bcE.push_back(face(FACE_MIRROR_E, ea, top, eE, r->f.Ec, N));
            } else if (b == MPIC_BC_OPEN) {
                r->bcH.push_back(face(FACE_WALL_H, ea, top, eH, r->f.Hc, N));
                FaceOp e = face(FACE_MUR_E, ea, top, eE, r->f.Ec, N);
                const long nodes = (long)(e.phi - e.plo + 1) * (e.qhi - e.qlo + 1);
                double *buf = nullptr;
                MPIC_CHECK(hipMalloc(&buf, sizeof(double) * 4 * nodes));
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

/* ---------------- lifecycle ---------------- */
static size_t field_bytes(const int n[3])
{
    return sizeof(double) * 15 * (size_t)(n[0] + 5) * (n[1] + 5) * (n[2] + 5);
}

extern "C" int mpic_footprint(const mpic_config_t *cfg, size_t *bytes)
{
    if (!cfg || !bytes) return MPIC_ERR_ARG;
    const size_t cells = (size_t)(cfg->n[0] + 5) * (cfg->n[1] + 5) * (cfg->n[2] + 5);
    *bytes = field_bytes(cfg->n) + 2 * sizeof(double) * cells +
             (size_t)cfg->max_markers * (6 * sizeof(double) + 2 * sizeof(float) + 4);
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
    if (cfg->precision != MPIC_FP64) {
        mpic_set_error("mpic_create: this library build supports MPIC_FP64 runs only");
        return MPIC_ERR_ARG;
    }
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
    g.GZ = cfg->n[2] + 5;
    g.ncell = (long)(cfg->n[0] + 5) * g.GY * g.GZ;
    MPIC_CHECK(hipMalloc(&r->fblock, field_bytes(cfg->n)));
    /* NaN fill of E/H and the caps, as the reference's NaN-filled meshes;
     * J starts at zero. */
    MPIC_CHECK(hipMemsetAsync(r->fblock, 0xff, field_bytes(cfg->n), r->stream));
    double *p = r->fblock;
    for (int c = 0; c < 3; ++c) r->f.E[c] = p + (0 + c) * g.ncell;
    for (int c = 0; c < 3; ++c) r->f.H[c] = p + (3 + c) * g.ncell;
    for (int c = 0; c < 3; ++c) r->f.J[c] = p + (6 + c) * g.ncell;
    for (int c = 0; c < 3; ++c) r->f.Hc[c] = p + (9 + c) * g.ncell;
    for (int c = 0; c < 3; ++c) r->f.Ec[c] = p + (12 + c) * g.ncell;
    MPIC_CHECK(hipMemsetAsync(r->f.J[0], 0, sizeof(double) * 3 * g.ncell, r->stream));
    MPIC_CHECK(hipMalloc(&r->rho1, sizeof(double) * g.ncell));
    MPIC_CHECK(hipMalloc(&r->rho2, sizeof(double) * g.ncell));
    r->cap = cfg->max_markers > 0 ? cfg->max_markers : 1;
    MPIC_CHECK(hipMalloc(&r->m.x, sizeof(double) * r->cap));
    MPIC_CHECK(hipMalloc(&r->m.y, sizeof(double) * r->cap));
    MPIC_CHECK(hipMalloc(&r->m.z, sizeof(double) * r->cap));
    MPIC_CHECK(hipMalloc(&r->m.ux, sizeof(double) * r->cap));
    MPIC_CHECK(hipMalloc(&r->m.uy, sizeof(double) * r->cap));
    MPIC_CHECK(hipMalloc(&r->m.uz, sizeof(double) * r->cap));
    MPIC_CHECK(hipMalloc(&r->m.w, sizeof(float) * r->cap));
    MPIC_CHECK(hipMalloc(&r->m.qm, sizeof(float) * r->cap));
    MPIC_CHECK(hipMalloc(&r->m.id, sizeof(int32_t) * r->cap));
    r->n = 0;
    MPIC_CHECK(hipMalloc(&r->dacc, sizeof(double) * 8));
    MPIC_CHECK(hipHostMalloc(&r->hacc, sizeof(double) * 8, hipHostMallocDefault));
    MPIC_CHECK(hipMalloc(&r->derr, sizeof(unsigned)));
    MPIC_CHECK(hipMemsetAsync(r->derr, 0, sizeof(unsigned), r->stream));
    for (int e = 0; e < 4; ++e) MPIC_CHECK(hipEventCreate(&r->ev[e]));
    cap_geometry(cfg->n, cfg->bc, kUpdH, kFlH, r->cH);
    cap_geometry(cfg->n, cfg->bc, kUpdE, kFlE, r->cE);
    r->updH = make_set(r->cH.upd_s);
    r->updE = make_set(r->cE.upd_s);
    r->flH = make_set(r->cH.fl_s);
    r->flE = make_set(r->cE.fl_s);
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
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    *out = r;
    return MPIC_OK;
}

extern "C" int mpic_destroy(mpic_run_t *r)
{
    if (!r) return MPIC_OK;
    hipSetDevice(r->dev);
    if (r->stream) hipStreamSynchronize(r->stream);
    hipFree(r->fblock);
    hipFree(r->rho1);
    hipFree(r->rho2);
    hipFree(r->m.x); hipFree(r->m.y); hipFree(r->m.z);
    hipFree(r->m.ux); hipFree(r->m.uy); hipFree(r->m.uz);
    hipFree(r->m.w); hipFree(r->m.qm); hipFree(r->m.id);
    hipFree(r->dacc);
    hipFree(r->derr);
    if (r->hacc) hipHostFree(r->hacc);
    for (double *b : r->murbuf) hipFree(b);
    for (Laser &l : r->lasers) hipFree(l.dmap);
    for (int e = 0; e < 4; ++e)
        if (r->ev[e]) hipEventDestroy(r->ev[e]);
    if (r->stream) hipStreamDestroy(r->stream);
    delete r;
    return MPIC_OK;
}

/* ---------------- host <-> device data ---------------- */
static int nodes_of(const Grid &g) { return (g.n[0] + 1) * (g.n[1] + 1) * (g.n[2] + 1); }

static int scatter_block(mpic_run *r, const double *src, double *const dst[3])
{
    /* node block [0,n]^3, interleaved xyz -> padded SoA */
    const Grid &g = r->g;
    std::vector<double> tmp((size_t)g.ncell);
    for (int c = 0; c < 3; ++c) {
        MPIC_CHECK(hipMemcpy(tmp.data(), dst[c], sizeof(double) * g.ncell, hipMemcpyDeviceToHost));
        long q = 0;
        for (int i = 0; i <= g.n[0]; ++i)
            for (int j = 0; j <= g.n[1]; ++j)
                for (int k = 0; k <= g.n[2]; ++k, ++q) tmp[gidx(g, i, j, k)] = src[3 * q + c];
        MPIC_CHECK(hipMemcpy(dst[c], tmp.data(), sizeof(double) * g.ncell, hipMemcpyHostToDevice));
    }
    return MPIC_OK;
}

static int gather_block(mpic_run *r, double *const src[3], double *out)
{
    const Grid &g = r->g;
    std::vector<double> tmp((size_t)g.ncell);
    for (int c = 0; c < 3; ++c) {
        MPIC_CHECK(hipMemcpy(tmp.data(), src[c], sizeof(double) * g.ncell, hipMemcpyDeviceToHost));
        long q = 0;
        for (int i = 0; i <= g.n[0]; ++i)
            for (int j = 0; j <= g.n[1]; ++j)
                for (int k = 0; k <= g.n[2]; ++k, ++q) out[3 * q + c] = tmp[gidx(g, i, j, k)];
    }
    return MPIC_OK;
}

extern "C" int mpic_set_fields(mpic_run_t *r, const double *E, const double *H)
{
    if (!r || !E || !H) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    TRY(scatter_block(r, E, r->f.E));
    TRY(scatter_block(r, H, r->f.H));
    return MPIC_OK;
}

extern "C" int mpic_get_fields(mpic_run_t *r, double *E, double *H)
{
    if (!r) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    if (E) TRY(gather_block(r, r->f.E, E));
    if (H) TRY(gather_block(r, r->f.H, H));
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
        MPIC_CHECK(hipMemcpy(tmp.data(), r->f.J[c], sizeof(double) * g.ncell, hipMemcpyDeviceToHost));
        for (long q = 0; q < g.ncell; ++q) J[3 * q + c] = tmp[q];
    }
    return MPIC_OK;
}

struct RecView {              /* marker_t, source/type_marker.h:12-25 */
    double x, y, z, vx, vy, vz;
    float  rho, qDivM;
    int32_t id;
};

extern "C" int mpic_markers_load(mpic_run_t *r, const void *rec, long n, int recsize)
{
    if (!r || (!rec && n) || n < 0 || (recsize != 56 && recsize != 64)) {
        mpic_set_error("mpic_markers_load: bad arguments");
        return MPIC_ERR_ARG;
    }
    if (n > r->cap) {
        mpic_set_error("mpic_markers_load: %ld markers exceed the capacity %ld", n, r->cap);
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
        id[q] = (recsize == 64) ? v.id : (int32_t)(q + 1);
    }
    double *dst[6] = {r->m.x, r->m.y, r->m.z, r->m.ux, r->m.uy, r->m.uz};
    for (int c = 0; c < 6; ++c)
        MPIC_CHECK(hipMemcpy(dst[c], d[c].data(), sizeof(double) * n, hipMemcpyHostToDevice));
    MPIC_CHECK(hipMemcpy(r->m.w, w.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    MPIC_CHECK(hipMemcpy(r->m.qm, qm.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    MPIC_CHECK(hipMemcpy(r->m.id, id.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
    r->n = n;
    return MPIC_OK;
}

extern "C" int mpic_markers_save(mpic_run_t *r, void *rec, long *n, int recsize)
{
    if (!r || !n || (recsize != 56 && recsize != 64)) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    const long N = r->n;
    std::vector<double> d[6];
    std::vector<float> w(N), qm(N);
    std::vector<int32_t> id(N);
    double *src[6] = {r->m.x, r->m.y, r->m.z, r->m.ux, r->m.uy, r->m.uz};
    for (int c = 0; c < 6; ++c) {
        d[c].resize(N);
        MPIC_CHECK(hipMemcpy(d[c].data(), src[c], sizeof(double) * N, hipMemcpyDeviceToHost));
    }
    MPIC_CHECK(hipMemcpy(w.data(), r->m.w, sizeof(float) * N, hipMemcpyDeviceToHost));
    MPIC_CHECK(hipMemcpy(qm.data(), r->m.qm, sizeof(float) * N, hipMemcpyDeviceToHost));
    MPIC_CHECK(hipMemcpy(id.data(), r->m.id, sizeof(int32_t) * N, hipMemcpyDeviceToHost));
    long o = 0;
    char *dst = (char *)rec;
    for (long q = 0; q < N; ++q) {
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
    std::vector<float> qm(r->n);
    if (hipStreamSynchronize(r->stream) != hipSuccess) return -1;
    if (hipMemcpy(qm.data(), r->m.qm, sizeof(float) * r->n, hipMemcpyDeviceToHost) != hipSuccess)
        return -1;
    long c = 0;
    for (long q = 0; q < r->n; ++q) c += (qm[q] != 0.0f);
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
    r->lasers.push_back(l);
    return MPIC_OK;
}

/* ---------------- stages ---------------- */
static int walls_on(mpic_run *r, std::vector<FaceOp> &ops)
{
    for (FaceOp &op : ops) TRY(launch_face(r->stream, r->g, op));
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
    /* cap_resetGhosts on the main meshes (source/core/em_caps.c:266-308);
     * the main meshes carry one ghost layer: transverse extent [-1, n+1]. */
    for (int ax = 0; ax < 3; ++ax) {
        Box full;
        for (int a = 0; a < 3; ++a) {
            full.lo[a] = -1;
            full.hi[a] = g.n[a] + 1;
        }
        if (bc[ax] == MPIC_BC_PERIODIC) {
            TRY(launch_face(r->stream, g, face(FACE_PERIODIC_H, ax, 0, full, r->f.H, g.n[ax])));
            TRY(launch_face(r->stream, g, face(FACE_PERIODIC_E, ax, 0, full, r->f.E, g.n[ax])));
        } else if (bc[ax] == MPIC_BC_MIRROR) {
            FaceOp h = face(FACE_WALL_H, ax, 0, full, r->f.H, g.n[ax]);
            h.zeroN = 1;
            TRY(launch_face(r->stream, g, h));
            TRY(launch_face(r->stream, g, face(FACE_MIRROR_E, ax, 0, full, r->f.E, g.n[ax])));
        }
        if (bc[3 + ax] == MPIC_BC_MIRROR) {
            FaceOp h = face(FACE_WALL_H, ax, 1, full, r->f.H, g.n[ax]);
            h.zeroN = 1;
            TRY(launch_face(r->stream, g, h));
            TRY(launch_face(r->stream, g, face(FACE_MIRROR_E, ax, 1, full, r->f.E, g.n[ax])));
        }
    }
    /* Mur buffers from the main E (capBC_Mur_init, source/core/emCap_Mur.c:92-132). */
    for (FaceOp &op : r->bcE) {
        if (op.kind != FACE_MUR_E) continue;
        FaceOp ini = op;
        ini.kind = FACE_MUR_INIT;
        for (int c = 0; c < 3; ++c) ini.X[c] = r->f.E[c];
        TRY(launch_face(r->stream, g, ini));
    }
    /* plasma_init: walls applied once to every loaded marker
     * (source/core/plasma.c:409-414). */
    TRY(launch_push64(r->stream, r->m, r->n, r->f, g, r->pc, r->dacc, 1, r->derr));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    r->inited = 1;
    r->stage = 0;
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

static int laser_H(mpic_run *r, double *Fy, double *Fz, const Box &tu)
{
    const Grid &g = r->g;
    for (const Laser &l : r->lasers) {
        const double t = r->time - l.T0 + l.delay;
        const double ampl = exp(-pow(2 * t / l.dt, 2));
        const int I = l.forward ? l.plane : g.n[0] - l.plane + 1;
        if (I < tu.lo[0] || I > tu.hi[0]) continue;
        const int JMIN = (int)(fmax(-1, 5) + 0.1), JMAX = (int)(fmin(g.n[1] + 1, g.n[1] - 5) + 0.1);
        const int KMIN = (int)(fmax(-1, 5) + 0.1), KMAX = (int)(fmin(g.n[2] + 1, g.n[2] - 5) + 0.1);
        const int jmin = (int)(fmax(tu.lo[1], JMIN) + 0.1), jmax = (int)(fmin(tu.hi[1], JMAX) + 0.1);
        const int kmin = (int)(fmax(tu.lo[2], KMIN) + 0.1), kmax = (int)(fmin(tu.hi[2], KMAX) + 0.1);
        const double phase = 2 * M_PI * t + l.phi0;
        TRY(launch_laser(r->stream, Fy, Fz, g, l.dmap, I, JMIN, KMIN, KMAX, jmin, jmax, kmin, kmax,
                         ampl * g.tau / g.h[0], cos(phase), sin(phase)));
    }
    return MPIC_OK;
}

static int laser_E(mpic_run *r, const Box &tu)
{
    const Grid &g = r->g;
    for (const Laser &l : r->lasers) {
        const double t = r->time + 0.5 * g.tau - l.T0 + l.delay;
        const double ampl = exp(-pow(2 * t / l.dt, 2));
        const int I = l.forward ? l.plane - 1 : g.n[0] - l.plane + 1;
        if (I < tu.lo[0] || I > tu.hi[0]) continue;
        const int JMIN = (int)(fmax(-1, 5) + 0.1), JMAX = (int)(fmin(g.n[1] + 1, g.n[1] - 5) + 0.1);
        const int KMIN = (int)(fmax(-1, 5) + 0.1), KMAX = (int)(fmin(g.n[2] + 1, g.n[2] - 5) + 0.1);
        const int jmin = (int)(fmax(tu.lo[1], JMIN) + 0.1), jmax = (int)(fmin(tu.hi[1], JMAX) + 0.1);
        const int kmin = (int)(fmax(tu.lo[2], KMIN) + 0.1), kmax = (int)(fmin(tu.hi[2], KMAX) + 0.1);
        const double phase = 2 * M_PI * t + l.phi0;
        TRY(launch_laser(r->stream, r->f.E[1], r->f.E[2], g, l.dmap + 4 * l.nodes, I, JMIN, KMIN,
                         KMAX, jmin, jmax, kmin, kmax, ampl * g.tau / g.h[0], cos(phase),
                         sin(phase)));
    }
    return MPIC_OK;
}

extern "C" int mpic_em_h_half(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    TRY(check_stage(r, 0));
    MPIC_CHECK(hipSetDevice(r->dev));
    /* cap_cacheEH: cap H, walls on cap H, cap E, walls on cap E.  Sources
     * never reach the cap layer (checked in mpic_add_laser). */
    TRY(launch_cap_H(r->stream, r->f, r->g, r->updH));
    TRY(walls_on(r, r->bcH));
    TRY(launch_cap_E(r->stream, r->f, r->g, r->updE));
    TRY(walls_on(r, r->bcE));
    /* em_HSmallStep + cap_flushHalfH */
    TRY(launch_h_update(r->stream, r->f, r->g, r->flH, 0));
    r->stage = 1;
    return MPIC_OK;
}

extern "C" int mpic_em_energy(mpic_run_t *r, double *W_E, double *W_M)
{
    if (!r) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    MPIC_CHECK(hipMemsetAsync(r->dacc, 0, 2 * sizeof(double), r->stream));
    TRY(launch_energy(r->stream, r->f, r->g, r->dacc));
    MPIC_CHECK(hipMemcpyAsync(r->hacc, r->dacc, 2 * sizeof(double), hipMemcpyDeviceToHost, r->stream));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    const double *h = r->g.h;
    if (W_E) *W_E = 0.5 * r->hacc[0] * h[0] * h[1] * h[2];
    if (W_M) *W_M = 0.5 * r->hacc[1] * h[0] * h[1] * h[2];
    return MPIC_OK;
}

extern "C" int mpic_push_deposit(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    TRY(check_stage(r, 1));
    MPIC_CHECK(hipSetDevice(r->dev));
    TRY(launch_zero(r->stream, r->f.J[0], 3 * r->g.ncell));
    MPIC_CHECK(hipMemsetAsync(r->dacc + 2, 0, 3 * sizeof(double), r->stream));
    if (r->timed) MPIC_CHECK(hipEventRecord(r->ev[1], r->stream));
    TRY(launch_push64(r->stream, r->m, r->n, r->f, r->g, r->pc, r->dacc + 2, 0, r->derr));
    if (r->timed) MPIC_CHECK(hipEventRecord(r->ev[2], r->stream));
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

extern "C" int mpic_em_h_step(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    TRY(check_stage(r, 1));
    MPIC_CHECK(hipSetDevice(r->dev));
    TRY(launch_h_update(r->stream, r->f, r->g, r->flH, 1));
    Box inner = {{3, 3, 3}, {r->g.n[0] - 2, r->g.n[1] - 2, r->g.n[2] - 2}};
    TRY(laser_H(r, r->f.H[1], r->f.H[2], inner));
    r->stage = 2;
    return MPIC_OK;
}

extern "C" int mpic_em_e_start(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    TRY(check_stage(r, 2));
    MPIC_CHECK(hipSetDevice(r->dev));
    /* interior E step; the cap flush (which em_EStep_finish does) touches
     * disjoint nodes and reads only Ec, so it is done in the same pass */
    TRY(launch_e_update(r->stream, r->f, r->g, r->flE, 1, 1));
    r->stage = 3;
    return MPIC_OK;
}

extern "C" int mpic_j_finish(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    const int *bc = r->cfg.bc;
    const int BCv[6] = {bc[0], bc[3], bc[1], bc[4], bc[2], bc[5]};
    for (int b = 0; b < 6; ++b) TRY(launch_jfold(r->stream, r->f, r->g, b, BCv[b]));
    return MPIC_OK;
}

extern "C" int mpic_em_e_finish(mpic_run_t *r)
{
    if (!r) return MPIC_ERR_ARG;
    TRY(check_stage(r, 3));
    MPIC_CHECK(hipSetDevice(r->dev));
    Box dom = {{-1, -1, -1}, {r->g.n[0] + 1, r->g.n[1] + 1, r->g.n[2] + 1}};
    TRY(laser_E(r, dom));
    TRY(launch_e_tauJ(r->stream, r->f, r->g));
    r->time = r->time + r->g.tau;          /* parameter_setTime (Time + tau) */
    r->stage = 0;
    return MPIC_OK;
}

static int step_async(mpic_run *r, int timed)
{
    r->timed = timed;
    if (timed) MPIC_CHECK(hipEventRecord(r->ev[0], r->stream));
    TRY(mpic_em_h_half(r));
    TRY(mpic_push_deposit(r));
    r->timed = 0;
    TRY(mpic_em_h_step(r));
    TRY(mpic_em_e_start(r));
    TRY(mpic_j_finish(r));
    TRY(mpic_em_e_finish(r));
    if (timed) MPIC_CHECK(hipEventRecord(r->ev[3], r->stream));
    return MPIC_OK;
}

extern "C" int mpic_step(mpic_run_t *r, double *diag)
{
    if (!r) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    if (!diag) return step_async(r, 0);
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

extern "C" int mpic_advance(mpic_run_t *r, int nsteps)
{
    if (!r || nsteps < 0) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    double push = 0, field = 0;
    for (int s = 0; s < nsteps; ++s) {
        TRY(step_async(r, 1));
        MPIC_CHECK(hipEventSynchronize(r->ev[3]));
        float a = 0, b = 0, c = 0;
        MPIC_CHECK(hipEventElapsedTime(&a, r->ev[0], r->ev[1]));
        MPIC_CHECK(hipEventElapsedTime(&b, r->ev[1], r->ev[2]));
        MPIC_CHECK(hipEventElapsedTime(&c, r->ev[2], r->ev[3]));
        push += b;          /* the push+deposit kernel alone */
        field += a + c;     /* everything else of the step */
    }
    r->push_ms = push;
    r->field_ms = field;
    return MPIC_OK;
}

static int check_flags(mpic_run *r)
{
    unsigned f = 0;
    MPIC_CHECK(hipMemcpy(&f, r->derr, sizeof(unsigned), hipMemcpyDeviceToHost));
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

extern "C" int mpic_last_timing(mpic_run_t *r, double *push_ms, double *field_ms)
{
    if (!r) return MPIC_ERR_ARG;
    if (push_ms) *push_ms = r->push_ms;
    if (field_ms) *field_ms = r->field_ms;
    return MPIC_OK;
}

/* ---------------- rho, Gauss residuals, cells ---------------- */
static int rho_into(mpic_run *r, double *rho)
{
    TRY(launch_zero(r->stream, rho, r->g.ncell));
    TRY(launch_rho64(r->stream, r->m, r->n, rho, r->g));
    const int *bc = r->cfg.bc;
    const int BCv[6] = {bc[0], bc[3], bc[1], bc[4], bc[2], bc[5]};
    for (int b = 0; b < 6; ++b) TRY(launch_rho_fold(r->stream, rho, r->g, b, BCv[b]));
    return MPIC_OK;
}

extern "C" int mpic_rho(mpic_run_t *r, double *rho)
{
    if (!r || !rho) return MPIC_ERR_ARG;
    MPIC_CHECK(hipSetDevice(r->dev));
    TRY(rho_into(r, r->rho2));
    MPIC_CHECK(hipMemcpyAsync(rho, r->rho2, sizeof(double) * r->g.ncell, hipMemcpyDeviceToHost,
                              r->stream));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
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
    TRY(launch_gauss(r->stream, r->f, r->rho1, r->rho2, r->g, r->dacc + 5));
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
    int32_t *d = nullptr;
    MPIC_CHECK(hipMallocAsync((void **)&d, sizeof(int32_t) * (r->n ? r->n : 1), r->stream));
    TRY(launch_cells64(r->stream, r->m, r->n, r->g, d));
    std::vector<int32_t> all(r->n);
    MPIC_CHECK(hipMemcpyAsync(all.data(), d, sizeof(int32_t) * r->n, hipMemcpyDeviceToHost,
                              r->stream));
    MPIC_CHECK(hipStreamSynchronize(r->stream));
    MPIC_CHECK(hipFreeAsync(d, r->stream));
    long o = 0;
    for (long q = 0; q < r->n; ++q) {
        if (all[q] < 0) continue;
        if (o >= *n) return MPIC_ERR_CAPACITY;
        cells[o++] = all[q];
    }
    *n = o;
    return MPIC_OK;
}

extern "C" int mpic_sort(mpic_run_t *r, int32_t *counts)
{
    (void)counts;
    if (!r) return MPIC_ERR_ARG;
    mpic_set_error("mpic_sort: not available in this build");
    return MPIC_ERR_STATE;
}
