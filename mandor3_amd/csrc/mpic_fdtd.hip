/*
 * mpic_fdtd.hip -- Yee field update, boundary caps and walls (FP64).
 *
 * Built with -ffp-contract=off: every update keeps the reference's
 * association (source/core/em.c:55-222, source/core/em_caps.c:405-809), so
 * given the same J the fields are bit-identical to the CPU path.
 *
 * The reference advances a 3-node boundary layer ("caps") one half step
 * ahead of the interior so that halo messages overlap with the push; on one
 * GPU there is no halo, but the cap arithmetic (full-step H^{n+1/2}, the
 * H^n = (H^{n-1/2} + H^{n+1/2})/2 average, the wall order and Mur buffers)
 * decides the boundary values, so it is kept: Hc/Ec hold the cap values of
 * the whole layer in one array (the per-cap copies of the reference hold
 * identical values on their overlaps).
 *
 * Roofline: all kernels here are HBM-bound stencils.  Interior H update
 * reads E (3 comps) and H (3) and writes H (3): 72 B/node; interior E
 * update likewise 72 B/node; E -= tau*J reads J (3) and E (3), writes E
 * (3): 72 B/node.
 */
#include "mpic_internal.h"
#include "mpic_launch.h"

#define TPB 256

static inline unsigned nblk(long n) { return (unsigned)((n + TPB - 1) / TPB); }

/* Volume kernels, vector form: a thread takes V consecutive nodes of a row
 * (k + 2 = 0 mod V) and moves each field as one 16-byte access (V = 4 FP32
 * or 2 FP64 nodes); the row pitch is a multiple of 4 nodes, so every such
 * access is aligned.  Per node the arithmetic is the scalar update's.  A
 * run of workgroups covers one i plane (threads over the plane's rows and
 * groups); planes are dealt to XCDs in contiguous ranges, so the planes an
 * update reads beside its own are mostly in the same L2. */
template <class T> struct Vec;
template <> struct Vec<float> {
    typedef float v __attribute__((ext_vector_type(4)));
    enum { N = 4 };
};
template <> struct Vec<double> {
    typedef double v __attribute__((ext_vector_type(2)));
    enum { N = 2 };
};

/* Groups of V along k covering nodes [klo, khi], rows j in [jlo, jhi]. */
struct Span {
    int ilo, jlo, nj, glo, ng;      /* first plane, first row, rows, first group, groups */
};

template <int V>
static inline Span span_of(int ilo, int jlo, int jhi, int klo, int khi)
{
    Span sp;
    sp.ilo = ilo;
    sp.jlo = jlo;
    sp.nj = jhi - jlo + 1;
    sp.glo = (klo + MPIC_GHOST) / V;
    sp.ng = (khi + MPIC_GHOST) / V - sp.glo + 1;
    return sp;
}

static inline unsigned grid_span(const Span &sp, int planes)
{
    const long per = (long)sp.nj * sp.ng;
    return (unsigned)((per + TPB - 1) / TPB) * (unsigned)planes;
}

/* Plane i, row j and first node k0 of this thread's group. */
__device__ static inline bool span_group(const Span &sp, int V, const VBlock &b,
                                         const VBlock &gd, int &i, int &j, int &k0)
{
    const unsigned per = (unsigned)sp.nj * (unsigned)sp.ng;
    const unsigned bpp = (per + TPB - 1) / TPB;              /* workgroups per plane */
    const unsigned lb = xcd_block(b.x, gd.x);
    const unsigned ip = lb / bpp;
    const unsigned t = (lb - ip * bpp) * TPB + threadIdx.x;
    if (t >= per) return false;
    const unsigned jj = t / (unsigned)sp.ng;
    i = sp.ilo + (int)ip;
    j = sp.jlo + (int)jj;
    k0 = (sp.glo + (int)(t - jj * (unsigned)sp.ng)) * V - MPIC_GHOST;
    return true;
}

__device__ static inline bool box_decode(const BoxSet &bs, long t, int &i, int &j, int &k)
{
    if (t >= bs.off[bs.n]) return false;
    int b = 0;
    while (t >= bs.off[b + 1]) ++b;
    /* boxes hold fewer than 2^32 nodes: 32-bit division */
    unsigned l = (unsigned)(t - bs.off[b]);
    const Box &B = bs.b[b];
    const unsigned dz = B.hi[2] - B.lo[2] + 1, dy = B.hi[1] - B.lo[1] + 1;
    const unsigned l2 = l / dz;
    k = B.lo[2] + (int)(l - l2 * dz);
    const unsigned l3 = l2 / dy;
    j = B.lo[1] + (int)(l2 - l3 * dy);
    i = B.lo[0] + (int)l3;
    return true;
}

/* Cap H^{n+1/2} = H + c3*(dEy/dz) - c2*(dEz/dy)..., source/core/em_caps.c:418-439. */
template <class T> struct CapArgs {
    FieldsT<T> f;
    Grid       g;
    BoxSet     bs;
    T          c1, c2, c3;
    int        E;                  /* 0: cap H, 1: cap E */
};

template <class T>
__device__ static inline void cap_H(const FieldsT<T> &f, const Grid &g, const BoxSet &bs, T c1,
                                    T c2, T c3, long t)
{
    int i, j, k;
    if (!box_decode(bs, t, i, j, k)) return;
    long o = gidx(g, i, j, k);
    long oi = gidx(g, i - 1, j, k), oj = gidx(g, i, j - 1, k), ok = gidx(g, i, j, k - 1);
    const T *Ex = f.E[0], *Ey = f.E[1], *Ez = f.E[2];
    /* all loads first (the stores go to Hc, which the loads never read) */
    const T h0 = f.H[0][o], h1 = f.H[1][o], h2 = f.H[2][o];
    const T ex = Ex[o], exj = Ex[oj], exk = Ex[ok], ey = Ey[o], eyi = Ey[oi], eyk = Ey[ok];
    const T ez = Ez[o], ezi = Ez[oi], ezj = Ez[oj];
    f.Hc[0][o] = h0 + c3 * (ey - eyk) - c2 * (ez - ezj);
    f.Hc[1][o] = h1 + c1 * (ez - ezi) - c3 * (ex - exk);
    f.Hc[2][o] = h2 + c2 * (ex - exj) - c1 * (ey - eyi);
}

/* Cap E^{n+1} from the cap H, source/core/em_caps.c:580-599. */
template <class T>
__device__ static inline void cap_E(const FieldsT<T> &f, const Grid &g, const BoxSet &bs, T c1,
                                    T c2, T c3, long t)
{
    int i, j, k;
    if (!box_decode(bs, t, i, j, k)) return;
    long o = gidx(g, i, j, k);
    long oi = gidx(g, i + 1, j, k), oj = gidx(g, i, j + 1, k), ok = gidx(g, i, j, k + 1);
    const T *Hx = f.Hc[0], *Hy = f.Hc[1], *Hz = f.Hc[2];
    /* all loads first (the stores go to Ec, which the loads never read) */
    const T e0 = f.E[0][o], e1 = f.E[1][o], e2 = f.E[2][o];
    const T hx = Hx[o], hxj = Hx[oj], hxk = Hx[ok], hy = Hy[o], hyi = Hy[oi], hyk = Hy[ok];
    const T hz = Hz[o], hzi = Hz[oi], hzj = Hz[oj];
    f.Ec[0][o] = e0 - c3 * (hyk - hy) + c2 * (hzj - hz);
    f.Ec[1][o] = e1 - c1 * (hzi - hz) + c3 * (hxk - hx);
    f.Ec[2][o] = e2 - c2 * (hxj - hx) + c1 * (hyi - hy);
}

template <class T> struct KernelDef<CapArgs<T>> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const CapArgs<T> &a, VBlock b, VBlock)
    {
        const long t = (long)b.x * TPB + threadIdx.x;
        if (a.E)
            cap_E<T>(a.f, a.g, a.bs, a.c1, a.c2, a.c3, t);
        else
            cap_H<T>(a.f, a.g, a.bs, a.c1, a.c2, a.c3, t);
    }
};

template <class T>
static int launch_cap(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &bs, int E)
{
    const long n = bs.off[bs.n];
    if (n <= 0) return MPIC_OK;
    CapArgs<T> a = {f, g, bs, (T)(g.tau / g.h[0]), (T)(g.tau / g.h[1]), (T)(g.tau / g.h[2]), E};
    return launch_def(s, a, nblk(n));
}

template <class T>
int launch_cap_H(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &bs)
{
    return launch_cap(s, f, g, bs, 0);
}

template <class T>
int launch_cap_E(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &bs)
{
    return launch_cap(s, f, g, bs, 1);
}

/* Walls of one face: periodic, mirror and first-order Mur
 * (source/core/emCap_periodic.c:12-68, emCap_mirror.c:17-72,
 * emCap_Mur.c:55-175). */
template <class T>
__device__ static inline void face_apply(const Grid &g, const FaceOp &op, long t)
{
    T *X[3] = {(T *)op.X[0], (T *)op.X[1], (T *)op.X[2]};
    const T alpha = (T)op.alpha;
    const unsigned np = op.phi - op.plo + 1, nq = op.qhi - op.qlo + 1;
    const long tot = (long)np * nq;
    if (t >= tot) return;
    /* consecutive threads take consecutive nodes along the higher tangential
     * axis (k for x and y walls, j for z walls), so that a wavefront's
     * accesses are contiguous; the Mur buffer is indexed by t in the same
     * order by every kind */
    int p, q;
    if (op.ea == 1) {
        const unsigned qq = (unsigned)t / np;
        p = op.plo + (int)((unsigned)t - qq * np);
        q = op.qlo + (int)qq;
    } else {
        const unsigned pq = (unsigned)t / nq;
        p = op.plo + (int)pq;
        q = op.qlo + (int)((unsigned)t - pq * nq);
    }
    int ea = op.ea, ep = (ea + 1) % 3, eq = (ea + 2) % 3;
    int r[3];
    r[ep] = p;
    r[eq] = q;
#define AT(a) (r[ea] = (a), gidx(g, r[0], r[1], r[2]))
    const int N = op.N, top = op.top;
    const int pl = top ? N : 0;
    /* in every case all values are loaded before any is stored: the nodes a
     * line reads and writes are distinct (per component), so this is the
     * reference's sequence with one memory latency */
    switch (op.kind) {
    case FACE_PERIODIC_H: {
        long o0 = AT(0), oN = AT(N), o1 = AT(1), oN1 = AT(N + 1);
        T vN[3], v1[3];
        for (int c = 0; c < 3; ++c) {
            vN[c] = X[c][oN];
            v1[c] = X[c][o1];
        }
        for (int c = 0; c < 3; ++c) {
            X[c][o0] = vN[c];
            X[c][oN1] = v1[c];
        }
        break;
    }
    case FACE_PERIODIC_E: {
        long o0 = AT(0), oN = AT(N), o1 = AT(1), oN1 = AT(N + 1);
        T v0[3], v1[3];
        for (int c = 0; c < 3; ++c) {
            v0[c] = X[c][o0];
            v1[c] = X[c][o1];
        }
        for (int c = 0; c < 3; ++c) {
            X[c][oN] = v0[c];
            X[c][oN1] = v1[c];
        }
        break;
    }
    case FACE_WALL_H: {
        long opl = AT(pl), oin = AT(pl + 1 - top), oex = AT(pl + top);
        const T vp = X[ep][oin], vq = X[eq][oin];
        if (op.zeroN) X[ea][opl] = 0;
        X[ep][oex] = vp;
        X[eq][oex] = vq;
        break;
    }
    case FACE_MIRROR_E: {
        long opl = AT(pl), oin = AT(pl + 1 - top), oex = AT(pl + top);
        const T va = X[ea][oin];
        X[ep][opl] = 0;
        X[eq][opl] = 0;
        X[ea][oex] = va;
        break;
    }
    case FACE_MUR_E: {
        long opl = AT(pl), os = AT(pl + 1 - 2 * top), oin = AT(pl + 1 - top), oex = AT(pl + top);
        T *b = (T *)op.buf + 4 * t;
        const T va = X[ea][oin], sp = X[ep][os], sq = X[eq][os];
        const T b0 = b[0], b1 = b[1], b2 = b[2], b3 = b[3];
        const T vp = b2 + alpha * (sp - b0);
        const T vq = b3 + alpha * (sq - b1);
        X[ea][oex] = va;
        X[ep][opl] = vp;
        X[eq][opl] = vq;
        b[0] = vp;
        b[1] = vq;
        b[2] = sp;
        b[3] = sq;
        break;
    }
    case FACE_MUR_INIT: {
        long opl = AT(pl), os = AT(pl + 1 - 2 * top), oin = AT(pl + 1 - top), oex = AT(pl + top);
        T *b = (T *)op.buf + 4 * t;
        const T pp = X[ep][opl], pq = X[eq][opl], sp = X[ep][os], sq = X[eq][os], va = X[ea][oin];
        b[0] = pp;
        b[1] = pq;
        b[2] = sp;
        b[3] = sq;
        X[ea][oex] = va;
        break;
    }
    }
#undef AT
}

/* One wall, or the two walls of one axis (block y picks the wall). */
template <class T> struct FaceArgs {
    Grid   g;
    FaceOp op[2];
};

template <class T> struct KernelDef<FaceArgs<T>> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const FaceArgs<T> &a, VBlock b, VBlock)
    {
        face_apply<T>(a.g, a.op[b.y], (long)b.x * TPB + threadIdx.x);
    }
};

int launch_face(hipStream_t s, const Grid &g, const FaceOp &op)
{
    long n = (long)(op.phi - op.plo + 1) * (op.qhi - op.qlo + 1);
    if (n <= 0) return MPIC_OK;
    if (op.fp32) {
        FaceArgs<float> a = {g, {op, op}};
        return launch_def(s, a, nblk(n));
    }
    FaceArgs<double> a = {g, {op, op}};
    return launch_def(s, a, nblk(n));
}

/* The y and z wall ops of a list, in list order, one workgroup per x plane:
 * every line of a y or z wall keeps i fixed, so a plane's lines depend only
 * on earlier ops' lines of the same plane, and a workgroup barrier between
 * consecutive ops gives the list's sequence exactly (edges included).  The
 * x walls, whose lines cross planes, are launched before. */
template <class T> struct FacePlanesArgs {
    Grid   g;
    FaceOp op[4];
    int    phase[4];      /* ops of one phase touch disjoint nodes; a barrier between phases */
    int    nop, ilo;
};

/* Lines of op on plane i: [first, first + count). */
__device__ static inline void plane_lines(const FaceOp &op, int i, long &first, long &count)
{
    const long np = op.phi - op.plo + 1, nq = op.qhi - op.qlo + 1;
    first = 0;
    count = 0;
    if (op.ea == 1) {                 /* eq = 0: q is i, consecutive lines along p */
        if (i < op.qlo || i > op.qhi) return;
        first = (long)(i - op.qlo) * np;
        count = np;
    } else {                          /* ea == 2, ep = 0: p is i */
        if (i < op.plo || i > op.phi) return;
        first = (long)(i - op.plo) * nq;
        count = nq;
    }
}

template <class T> struct KernelDef<FacePlanesArgs<T>> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const FacePlanesArgs<T> &a, VBlock b, VBlock)
    {
        const int i = a.ilo + (int)b.x;
        for (int q = 0; q < a.nop; ++q) {
            if (q && a.phase[q] != a.phase[q - 1])
                __syncthreads();          /* the plane's lines of the earlier phase are done */
            long first, count;
            plane_lines(a.op[q], i, first, count);
            for (long u = threadIdx.x; u < count; u += TPB) face_apply<T>(a.g, a.op[q], first + u);
        }
    }
};

int launch_face_planes(hipStream_t s, const Grid &g, const FaceOp *ops, int nop)
{
    if (nop <= 0) return MPIC_OK;
    int ilo = 1 << 30, ihi = -(1 << 30);
    for (int q = 0; q < nop; ++q) {
        const FaceOp &op = ops[q];
        const int lo = op.ea == 1 ? op.qlo : op.plo, hi = op.ea == 1 ? op.qhi : op.phi;
        ilo = lo < ilo ? lo : ilo;
        ihi = hi > ihi ? hi : ihi;
    }
    if (ihi < ilo) return MPIC_OK;
    /* the min and max wall of an axis with at least 8 cells touch disjoint
     * node layers (as launch_face_pair uses): one phase */
    int phase[4], ph = 0;
    for (int q = 0; q < nop; ++q) {
        const bool pair = q > 0 && ops[q].ea == ops[q - 1].ea && ops[q - 1].top == 0 &&
                          ops[q].top == 1 && ops[q].kind == ops[q - 1].kind &&
                          ops[q].kind != FACE_PERIODIC_H && ops[q].kind != FACE_PERIODIC_E &&
                          ops[q].N >= 8 && (q < 2 || phase[q - 1] != phase[q - 2]);
        if (q > 0 && !pair) ++ph;
        phase[q] = ph;
    }
    if (ops[0].fp32) {
        FacePlanesArgs<float> a;
        a.g = g;
        for (int q = 0; q < nop; ++q) {
            a.op[q] = ops[q];
            a.phase[q] = phase[q];
        }
        a.nop = nop;
        a.ilo = ilo;
        return launch_def(s, a, (unsigned)(ihi - ilo + 1));
    }
    FacePlanesArgs<double> a;
    a.g = g;
    for (int q = 0; q < nop; ++q) {
        a.op[q] = ops[q];
        a.phase[q] = phase[q];
    }
    a.nop = nop;
    a.ilo = ilo;
    return launch_def(s, a, (unsigned)(ihi - ilo + 1));
}

/* a and b: the min and max wall of one axis with at least 8 cells, so that
 * the node layers they touch ([0, 1] and [N-1, N+1]) are disjoint and one
 * launch gives the result of the two in sequence. */
int launch_face_pair(hipStream_t s, const Grid &g, const FaceOp &a, const FaceOp &b)
{
    const long na = (long)(a.phi - a.plo + 1) * (a.qhi - a.qlo + 1);
    const long nb = (long)(b.phi - b.plo + 1) * (b.qhi - b.qlo + 1);
    const long n = na > nb ? na : nb;
    if (n <= 0) return MPIC_OK;
    if (a.fp32) {
        FaceArgs<float> fa = {g, {a, b}};
        return launch_def(s, fa, nblk(n), 2);
    }
    FaceArgs<double> fa = {g, {a, b}};
    return launch_def(s, fa, nblk(n), 2);
}

/* Cap-flush nodes of a vector group (row i, j, nodes k0 .. k0+V-1): bit e
 * is set when node k0+e is not an interior node ('interior': bits of the
 * nodes the interior formula takes), lies in [0, n+1] along k and in a box
 * of bs.  The boxes' i, j ranges are tested once per group.  Only groups at
 * the edges of the volume have any bit set; they also load the cap values
 * of the whole group together with the fields, so that no wavefront waits
 * twice on memory. */
__device__ static inline unsigned flush_mask(const BoxSet &bs, int i, int j, int k0, int V,
                                             int nk, unsigned interior)
{
    unsigned m = 0;
    if (bs.shell >= 0) {                     /* every non-interior node of [0, n+1]^3 */
        for (int e = 0; e < V; ++e)
            if (k0 + e >= 0 && k0 + e <= nk + 1) m |= 1u << e;
        return m & ~interior;
    }
    for (int b = 0; b < bs.n; ++b) {
        const Box &B = bs.b[b];
        if (i < B.lo[0] || i > B.hi[0] || j < B.lo[1] || j > B.hi[1]) continue;
        for (int e = 0; e < V; ++e) {
            const int k = k0 + e;
            if (k >= B.lo[2] && k <= B.hi[2] && k >= 0 && k <= nk + 1) m |= 1u << e;
        }
    }
    return m & ~interior;
}

/* Bits of the nodes k0 .. k0+V-1 inside [lo, hi]. */
__device__ static inline unsigned k_bits(int k0, int V, int lo, int hi)
{
    unsigned m = 0;
    for (int e = 0; e < V; ++e)
        if (k0 + e >= lo && k0 + e <= hi) m |= 1u << e;
    return m;
}

/* One-plane laser soft source, source/scpic/em_mirror.c:667-693.  The time
 * factors (envelope, cos, sin of the carrier) are evaluated in FP64 on the
 * host once per step and passed in. */
template <class T>
__device__ static inline void laser_node(const LaserPlane &lp, int j, int k, double A, double cs,
                                         double sn, T &fy, T &fz)
{
    const long cnt = (long)(j - lp.JMIN) * (lp.KMAX - lp.KMIN + 1) + k - lp.KMIN;
    const double *w = lp.map + 4 * cnt;
    fy = (T)(fy + A * (w[2] * cs + w[3] * sn));
    fz = (T)(fz + A * (w[0] * cs + w[1] * sn));
}

/* The source added to a vector group (nodes k0 .. k0+V-1 of row (i, j))
 * whose new Fy/Fz values are in registers: the same arithmetic on the
 * same values as the separate launch after the update stored them. */
template <class T, class VT>
__device__ static inline void laser_group(const LaserPlane &lp, int i, int j, int k0, int V,
                                          VT &fy, VT &fz)
{
    if (!lp.on || i != lp.pos || j < lp.jmin || j > lp.jmax) return;
    if (k0 + V - 1 < lp.kmin || k0 > lp.kmax) return;
    const double *e = lp.tab + 3 * (*lp.step % lp.cap);
    const double A = e[0], cs = e[1], sn = e[2];
    for (int q = 0; q < V; ++q) {
        const int k = k0 + q;
        if (k < lp.kmin || k > lp.kmax) continue;
        T y = fy[q], z = fz[q];
        laser_node<T>(lp, j, k, A, cs, sn, y, z);
        fy[q] = y;
        fz[q] = z;
    }
}

/* Interior H half step (source/core/em.c:55-77) fused with the flush of the
 * cap layer: H = (H + Hc)/2 after the first half step (cap_flushHalfH,
 * em_caps.c:730-749) or H = Hc after the second (cap_flushH, :754-770).
 * Nodes [0, n+1]^3; interior and flush regions are disjoint. */
template <class T> struct HUpdateArgs {
    FieldsT<T> f;
    Grid       g;
    BoxSet     fl;
    int        full;
    Span       sp;
    T          c1, c2, c3;
    double    *part;       /* non-null: em_energy partial sums, 2 per workgroup */
    LaserPlane lp;         /* a laser's H plane added after the update (on = 0: none) */
};

/* Sum over the workgroup of (we, wb), stored as part[0..1]; every thread of
 * the workgroup calls it. */
__device__ static inline void block_sum2(double we, double wb, double *part)
{
    __shared__ double se[TPB / 64], sb[TPB / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int d = 32; d > 0; d >>= 1) {
        we += __shfl_down(we, d, 64);
        wb += __shfl_down(wb, d, 64);
    }
    if (lane == 0) {
        se[wv] = we;
        sb[wv] = wb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, c = 0;
        for (int q = 0; q < TPB / 64; ++q) {
            a += se[q];
            c += sb[q];
        }
        part[0] = a;
        part[1] = c;
    }
}

/* em_energy's terms of node (i,j,k) (source/core/em.c:227-252): each
 * component squared where the node is that component's point of a cell of
 * [0,n)^3 -- Ex at (i+1,j,k), Ey at (i,j+1,k), Ez at (i,j,k+1), Hx at
 * (i,j+1,k+1), Hy at (i+1,j,k+1), Hz at (i+1,j+1,k) of cell (i,j,k).
 * A thread's partial over its few nodes is kept in the field precision
 * (FP32 form: ~1e-7 relative); the workgroup sums are FP64. */
template <class T>
__device__ static inline void energy_terms(const Grid &g, int i, int j, int k, T ex, T ey, T ez,
                                           T hx, T hy, T hz, T &we, T &wb)
{
    const bool i0 = i >= 0 && i < g.n[0], i1 = i >= 1 && i <= g.n[0];
    const bool j0 = j >= 0 && j < g.n[1], j1 = j >= 1 && j <= g.n[1];
    const bool k0 = k >= 0 && k < g.n[2], k1 = k >= 1 && k <= g.n[2];
    if (i1 && j0 && k0) we += ex * ex;
    if (i0 && j1 && k0) we += ey * ey;
    if (i0 && j0 && k1) we += ez * ez;
    if (i0 && j1 && k1) wb += hx * hx;
    if (i1 && j0 && k1) wb += hy * hy;
    if (i1 && j1 && k0) wb += hz * hz;
}

template <class T>
__device__ static inline void h_update_group(const FieldsT<T> &f, const Grid &g, const BoxSet &fl,
                                             int full, T c1, T c2, T c3, bool energy, int i,
                                             int j, int k0, T &we, T &wb,
                                             const LaserPlane &lp)
{
    typedef typename Vec<T>::v VT;
    constexpr int V = Vec<T>::N;
    const long o = gidx(g, i, j, k0);
    const long si = (long)g.GY * g.GZ, sj = g.GZ;
    const T *Ex = f.E[0], *Ey = f.E[1], *Ez = f.E[2];
    const VT ex = *(const VT *)(Ex + o), ey = *(const VT *)(Ey + o), ez = *(const VT *)(Ez + o);
    const VT exj = *(const VT *)(Ex + o - sj), ezj = *(const VT *)(Ez + o - sj);
    const VT eyi = *(const VT *)(Ey + o - si), ezi = *(const VT *)(Ez + o - si);
    const T exk0 = Ex[o - 1], eyk0 = Ey[o - 1];          /* k - 1 of the first node */
    VT hx = *(const VT *)(f.H[0] + o), hy = *(const VT *)(f.H[1] + o), hz = *(const VT *)(f.H[2] + o);
    const bool ij_in = i >= 3 && i <= g.n[0] - 2 && j >= 3 && j <= g.n[1] - 2;
    const unsigned inner = ij_in ? k_bits(k0, V, 3, g.n[2] - 2) : 0u;
    unsigned fm = 0;
    VT cx = {}, cy = {}, cz = {};
    if (inner != (1u << V) - 1) {
        fm = flush_mask(fl, i, j, k0, V, g.n[2], inner);
        if (fm) {
            cx = *(const VT *)(f.Hc[0] + o);
            cy = *(const VT *)(f.Hc[1] + o);
            cz = *(const VT *)(f.Hc[2] + o);
        }
    }
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const int k = k0 + e;
        const T exk = e ? ex[e - 1] : exk0, eyk = e ? ey[e - 1] : eyk0;
        if (inner >> e & 1) {
            hx[e] += c3 * (ey[e] - eyk) - c2 * (ez[e] - ezj[e]);
            hy[e] += c1 * (ez[e] - ezi[e]) - c3 * (ex[e] - exk);
            hz[e] += c2 * (ex[e] - exj[e]) - c1 * (ey[e] - eyi[e]);
        } else if (fm >> e & 1) {
            if (full) {
                hx[e] = cx[e];
                hy[e] = cy[e];
                hz[e] = cz[e];
            } else {
                hx[e] = (T)0.5 * (hx[e] + cx[e]);
                hy[e] = (T)0.5 * (hy[e] + cy[e]);
                hz[e] = (T)0.5 * (hz[e] + cz[e]);
            }
        }
        if (energy) energy_terms<T>(g, i, j, k, ex[e], ey[e], ez[e], hx[e], hy[e], hz[e], we, wb);
    }
    laser_group<T>(lp, i, j, k0, V, hy, hz);
    *(VT *)(f.H[0] + o) = hx;
    *(VT *)(f.H[1] + o) = hy;
    *(VT *)(f.H[2] + o) = hz;
}

template <class T>
__device__ static inline void h_update(const FieldsT<T> &f, const Grid &g, const BoxSet &fl,
                                       int full, const Span &sp, T c1, T c2, T c3, double *part,
                                       const LaserPlane &lp, const VBlock &b, const VBlock &gd)
{
    int i, j, k0;
    const bool on = span_group(sp, Vec<T>::N, b, gd, i, j, k0);
    T we = 0, wb = 0;
    if (on) h_update_group<T>(f, g, fl, full, c1, c2, c3, part != nullptr, i, j, k0, we, wb, lp);
    /* the energy of H after this update (the half step: H^n) and of E */
    if (part) block_sum2((double)we, (double)wb, part + 2 * (long)b.x);
}

template <class T> struct KernelDef<HUpdateArgs<T>> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const HUpdateArgs<T> &a, VBlock b, VBlock gd)
    {
        h_update<T>(a.f, a.g, a.fl, a.full, a.sp, a.c1, a.c2, a.c3, a.part, a.lp, b, gd);
    }
};

static Span h_span(const Grid &g, int V)
{
    return V == 4 ? span_of<4>(0, 0, g.n[1] + 1, 0, g.n[2] + 1)
                  : span_of<2>(0, 0, g.n[1] + 1, 0, g.n[2] + 1);
}

/* Workgroups of the H update: the energy partial sums it writes. */
int h_update_blocks(const Grid &g, int fp32)
{
    return (int)grid_span(h_span(g, fp32 ? Vec<float>::N : Vec<double>::N), g.n[0] + 2);
}

template <class T>
static HUpdateArgs<T> h_args(const FieldsT<T> &f, const Grid &g, const BoxSet &flush, int full,
                             double *part, const LaserPlane &lp)
{
    const Span sp = h_span(g, Vec<T>::N);
    HUpdateArgs<T> a = {f, g, flush, full, sp, (T)(0.5 * g.tau / g.h[0]),
                        (T)(0.5 * g.tau / g.h[1]), (T)(0.5 * g.tau / g.h[2]), part, lp};
    return a;
}

template <class T>
int launch_h_update(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &flush,
                    int full, double *part, const LaserPlane &lp)
{
    const HUpdateArgs<T> a = h_args(f, g, flush, full, part, lp);
    return launch_def(s, a, grid_span(a.sp, g.n[0] + 2));
}

/* Interior E step (source/core/em.c:134-161) and/or the flush of the cap
 * layer E = Ec (cap_flushE, em_caps.c:775-794). */
template <class T> struct EUpdateArgs {
    FieldsT<T> f;
    Grid       g;
    BoxSet     fl;
    int        interior, do_flush;
    Span       sp;
    T          c1, c2, c3;
};

template <class T>
__device__ static inline void e_update(const FieldsT<T> &f, const Grid &g, const BoxSet &fl,
                                       int interior, int do_flush, const Span &sp, T c1, T c2,
                                       T c3, const VBlock &b, const VBlock &gd)
{
    typedef typename Vec<T>::v VT;
    constexpr int V = Vec<T>::N;
    int i, j, k0;
    if (!span_group(sp, V, b, gd, i, j, k0)) return;
    const long o = gidx(g, i, j, k0);
    const long si = (long)g.GY * g.GZ, sj = g.GZ;
    const T *Hx = f.H[0], *Hy = f.H[1], *Hz = f.H[2];
    const VT hx = *(const VT *)(Hx + o), hy = *(const VT *)(Hy + o), hz = *(const VT *)(Hz + o);
    const VT hxj = *(const VT *)(Hx + o + sj), hzj = *(const VT *)(Hz + o + sj);
    const VT hyi = *(const VT *)(Hy + o + si), hzi = *(const VT *)(Hz + o + si);
    const T hxkV = Hx[o + V], hykV = Hy[o + V];          /* k + 1 of the last node */
    VT ex = *(const VT *)(f.E[0] + o), ey = *(const VT *)(f.E[1] + o), ez = *(const VT *)(f.E[2] + o);
    const bool ij_in = i >= 2 && i <= g.n[0] - 2 && j >= 2 && j <= g.n[1] - 2;
    const unsigned inner = ij_in ? k_bits(k0, V, 2, g.n[2] - 2) : 0u;
    unsigned fm = 0;
    VT cx = {}, cy = {}, cz = {};
    if (do_flush && inner != (1u << V) - 1) {
        fm = flush_mask(fl, i, j, k0, V, g.n[2], inner);
        if (fm) {
            cx = *(const VT *)(f.Ec[0] + o);
            cy = *(const VT *)(f.Ec[1] + o);
            cz = *(const VT *)(f.Ec[2] + o);
        }
    }
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const T hxk = e < V - 1 ? hx[e + 1] : hxkV, hyk = e < V - 1 ? hy[e + 1] : hykV;
        if (inner >> e & 1) {
            if (interior) {
                ex[e] -= c3 * (hyk - hy[e]) - c2 * (hzj[e] - hz[e]);
                ey[e] -= c1 * (hzi[e] - hz[e]) - c3 * (hxk - hx[e]);
                ez[e] -= c2 * (hxj[e] - hx[e]) - c1 * (hyi[e] - hy[e]);
            }
        } else if (fm >> e & 1) {
            ex[e] = cx[e];
            ey[e] = cy[e];
            ez[e] = cz[e];
        }
    }
    *(VT *)(f.E[0] + o) = ex;
    *(VT *)(f.E[1] + o) = ey;
    *(VT *)(f.E[2] + o) = ez;
}

template <class T> struct KernelDef<EUpdateArgs<T>> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const EUpdateArgs<T> &a, VBlock b, VBlock gd)
    {
        e_update<T>(a.f, a.g, a.fl, a.interior, a.do_flush, a.sp, a.c1, a.c2, a.c3, b, gd);
    }
};

template <class T>
int launch_e_update(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &flush,
                    int interior, int do_flush)
{
    const Span sp = span_of<Vec<T>::N>(0, 0, g.n[1] + 1, 0, g.n[2] + 1);
    EUpdateArgs<T> a = {f, g, flush, interior, do_flush, sp, (T)(g.tau / g.h[0]),
                        (T)(g.tau / g.h[1]), (T)(g.tau / g.h[2])};
    return launch_def(s, a, grid_span(sp, g.n[0] + 2));
}

/* E -= tau*J over nodes [-1, n+1]^3, source/core/em.c:166-184.  FP32
 * fields: the FP64 J sum is applied in FP64 and rounded once. */
template <class T> struct ETauJArgs {
    FieldsT<T> f;
    Grid       g;
    Span       sp;
    int       *tick;
    const int *jr;         /* x planes [jr[0], jr[1]] of nodes where J may be non-zero */
};

template <class T>
__device__ static inline void e_tauJ(const FieldsT<T> &f, const Grid &g, const Span &sp,
                                     int *tick, const int *jr, const VBlock &b, const VBlock &gd)
{
    typedef typename Vec<T>::v VT;
    typedef double VD __attribute__((ext_vector_type(Vec<T>::N)));
    constexpr int V = Vec<T>::N;
    /* the run's step counter, when this is the step's last kernel */
    if (tick && b.x == 0 && threadIdx.x == 0) *tick += 1;
    int i, j, k0;
    if (!span_group(sp, V, b, gd, i, j, k0)) return;
    if (i < jr[0] || i > jr[1]) return;                  /* J = 0: E - tau*0 is E */
    const long o = gidx(g, i, j, k0);
    const double tau = g.tau;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        VT e = *(const VT *)(f.E[c] + o);
        const VD jv = *(const VD *)(f.J[c] + o);
#pragma unroll
        for (int q = 0; q < V; ++q) {
            const int k = k0 + q;
            if (k >= -1 && k <= g.n[2] + 1) e[q] = (T)(e[q] - tau * jv[q]);
        }
        *(VT *)(f.E[c] + o) = e;
    }
}

template <class T> struct KernelDef<ETauJArgs<T>> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const ETauJArgs<T> &a, VBlock b, VBlock gd)
    {
        e_tauJ<T>(a.f, a.g, a.sp, a.tick, a.jr, b, gd);
    }
};

template <class T>
int launch_e_tauJ(hipStream_t s, const FieldsT<T> &f, const Grid &g, int *tick, const int *jr)
{
    const Span sp = span_of<Vec<T>::N>(-1, -1, g.n[1] + 1, -1, g.n[2] + 1);
    ETauJArgs<T> a = {f, g, sp, tick, jr};
    return launch_def(s, a, grid_span(sp, g.n[0] + 3));
}

/* The E step of the loop body in one pass: the interior update of
 * em_EStep_start, the cap flush E = Ec and E -= tau J of em_EStep_finish,
 * node by node in that order, over nodes [-1, n+1]^3 (the J fold has run
 * before).  Per node the arithmetic is that of e_update then e_tauJ; the
 * loop body uses it where no source adds to E between the two (no laser, no
 * TFSF interface) in the FP64 form, and always in the FP32 split form,
 * whose sources then add after E -= tau J (a different rounding order). */
template <class T> struct EFusedArgs {
    FieldsT<T> f;
    Grid       g;
    BoxSet     fl;
    Span       sp;
    T          c1, c2, c3;
    const int *jr;         /* x planes [jr[0], jr[1]] of nodes where J may be non-zero */
    LaserPlane lp;         /* a laser's E plane, added between the update and E -= tau J
                            * as the reference's sequence does (on = 0: none) */
};

template <class T>
__device__ static inline void e_fused(const FieldsT<T> &f, const Grid &g, const BoxSet &fl,
                                      const Span &sp, T c1, T c2, T c3, const int *jr,
                                      const LaserPlane &lp, const VBlock &b, const VBlock &gd)
{
    typedef typename Vec<T>::v VT;
    typedef double VD __attribute__((ext_vector_type(Vec<T>::N)));
    constexpr int V = Vec<T>::N;
    int i, j, k0;
    if (!span_group(sp, V, b, gd, i, j, k0)) return;
    const long o = gidx(g, i, j, k0);
    const long si = (long)g.GY * g.GZ, sj = g.GZ;
    const T *Hx = f.H[0], *Hy = f.H[1], *Hz = f.H[2];
    const VT hx = *(const VT *)(Hx + o), hy = *(const VT *)(Hy + o), hz = *(const VT *)(Hz + o);
    const VT hxj = *(const VT *)(Hx + o + sj), hzj = *(const VT *)(Hz + o + sj);
    const VT hyi = *(const VT *)(Hy + o + si), hzi = *(const VT *)(Hz + o + si);
    const T hxkV = Hx[o + V], hykV = Hy[o + V];          /* k + 1 of the last node */
    VT ex = *(const VT *)(f.E[0] + o), ey = *(const VT *)(f.E[1] + o), ez = *(const VT *)(f.E[2] + o);
    /* J is read only on the planes where it may be non-zero (elsewhere
     * E - tau*0 is E, the same value) */
    VD jx = {}, jy = {}, jz = {};
    if (i >= jr[0] && i <= jr[1]) {
        jx = *(const VD *)(f.J[0] + o);
        jy = *(const VD *)(f.J[1] + o);
        jz = *(const VD *)(f.J[2] + o);
    }
    const bool ij_in = i >= 2 && i <= g.n[0] - 2 && j >= 2 && j <= g.n[1] - 2;
    const bool ij_fl = i >= 0 && j >= 0;                 /* e_update's nodes start at 0 */
    const unsigned inner = ij_in ? k_bits(k0, V, 2, g.n[2] - 2) : 0u;
    unsigned fm = 0;
    VT cx = {}, cy = {}, cz = {};
    if (ij_fl && inner != (1u << V) - 1) {
        fm = flush_mask(fl, i, j, k0, V, g.n[2], inner);
        if (fm) {
            cx = *(const VT *)(f.Ec[0] + o);
            cy = *(const VT *)(f.Ec[1] + o);
            cz = *(const VT *)(f.Ec[2] + o);
        }
    }
    const double tau = g.tau;
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const int k = k0 + e;
        const T hxk = e < V - 1 ? hx[e + 1] : hxkV, hyk = e < V - 1 ? hy[e + 1] : hykV;
        if (inner >> e & 1) {
            ex[e] -= c3 * (hyk - hy[e]) - c2 * (hzj[e] - hz[e]);
            ey[e] -= c1 * (hzi[e] - hz[e]) - c3 * (hxk - hx[e]);
            ez[e] -= c2 * (hxj[e] - hx[e]) - c1 * (hyi[e] - hy[e]);
        } else if (fm >> e & 1) {
            ex[e] = cx[e];
            ey[e] = cy[e];
            ez[e] = cz[e];
        }
    }
    /* em_EStep_start, the laser's plane, then em_EStep_finish: per node the
     * operations of the separate launches, in their order */
    laser_group<T>(lp, i, j, k0, V, ey, ez);
#pragma unroll
    for (int e = 0; e < V; ++e) {
        const int k = k0 + e;
        if (k >= -1 && k <= g.n[2] + 1) {
            ex[e] = (T)(ex[e] - tau * jx[e]);
            ey[e] = (T)(ey[e] - tau * jy[e]);
            ez[e] = (T)(ez[e] - tau * jz[e]);
        }
    }
    *(VT *)(f.E[0] + o) = ex;
    *(VT *)(f.E[1] + o) = ey;
    *(VT *)(f.E[2] + o) = ez;
}

template <class T> struct KernelDef<EFusedArgs<T>> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const EFusedArgs<T> &a, VBlock b, VBlock gd)
    {
        e_fused<T>(a.f, a.g, a.fl, a.sp, a.c1, a.c2, a.c3, a.jr, a.lp, b, gd);
    }
};

template <class T>
int launch_e_fused(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &flush,
                   const int *jr, const LaserPlane &lp)
{
    const Span sp = span_of<Vec<T>::N>(-1, -1, g.n[1] + 1, -1, g.n[2] + 1);
    EFusedArgs<T> a = {f, g, flush, sp, (T)(g.tau / g.h[0]), (T)(g.tau / g.h[1]),
                       (T)(g.tau / g.h[2]), jr, lp};
    return launch_def(s, a, grid_span(sp, g.n[0] + 3));
}

/* The laser plane as a launch of its own. */
template <class T> struct LaserArgs {
    T         *Fy, *Fz;
    Grid       g;
    LaserPlane lp;
};

template <class T>
__device__ static inline void laser(T *Fy, T *Fz, const Grid &g, const LaserPlane &lp, long t)
{
    int nk = lp.kmax - lp.kmin + 1;
    if (t >= (long)(lp.jmax - lp.jmin + 1) * nk) return;
    /* {A, cos, sin} of this step, evaluated on the host */
    const double *e = lp.tab + 3 * (*lp.step % lp.cap);
    const double A = e[0], cs = e[1], sn = e[2];
    const unsigned jq = (unsigned)t / (unsigned)nk;
    int j = lp.jmin + (int)jq, k = lp.kmin + (int)((unsigned)t - jq * (unsigned)nk);
    long o = gidx(g, lp.pos, j, k);
    T y = Fy[o], z = Fz[o];
    laser_node<T>(lp, j, k, A, cs, sn, y, z);
    Fy[o] = y;
    Fz[o] = z;
}

template <class T> struct KernelDef<LaserArgs<T>> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const LaserArgs<T> &a, VBlock b, VBlock)
    {
        laser<T>(a.Fy, a.Fz, a.g, a.lp, (long)b.x * TPB + threadIdx.x);
    }
};

template <class T>
int launch_laser(hipStream_t s, T *Fy, T *Fz, const Grid &g, const LaserPlane &lp)
{
    if (!lp.on || lp.jmax < lp.jmin || lp.kmax < lp.kmin) return MPIC_OK;
    long n = (long)(lp.jmax - lp.jmin + 1) * (lp.kmax - lp.kmin + 1);
    LaserArgs<T> a = {Fy, Fz, g, lp};
    return launch_def(s, a, nblk(n));
}

/* Hard Gaussian spot (source/core/em_gaussSpot.c:129-156): Ey, Ez on plane
 * i over j, k in [-1, n+1] set to the step's amplitudes times the spatial
 * profiles (both evaluated on the host). */
template <class T> struct SpotArgs {
    T            *Ey, *Ez;
    Grid          g;
    int           i;
    const double *py, *pz, *tab;
    const int    *step;
    int           cap;
};

template <class T>
__device__ static inline void spot(T *Ey, T *Ez, const Grid &g, int i, const double *py,
                                   const double *pz, const double *tab, const int *step, int cap,
                                   long t)
{
    const int nk = g.n[2] + 3;
    if (t >= (long)(g.n[1] + 3) * nk) return;
    const unsigned jq = (unsigned)t / (unsigned)nk;
    const int j = -1 + (int)jq, k = -1 + (int)((unsigned)t - jq * (unsigned)nk);
    const double *e = tab + 2 * (*step % cap);
    const long o = gidx(g, i, j, k);
    Ey[o] = (T)(e[0] * py[t]);
    Ez[o] = (T)(e[1] * pz[t]);
}

template <class T> struct KernelDef<SpotArgs<T>> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const SpotArgs<T> &a, VBlock b, VBlock)
    {
        spot<T>(a.Ey, a.Ez, a.g, a.i, a.py, a.pz, a.tab, a.step, a.cap,
                (long)b.x * TPB + threadIdx.x);
    }
};

template <class T>
int launch_spot(hipStream_t s, T *Ey, T *Ez, const Grid &g, int i, const double *py,
                const double *pz, const double *tab, const int *step, int cap)
{
    const long n = (long)(g.n[1] + 3) * (g.n[2] + 3);
    SpotArgs<T> a = {Ey, Ez, g, i, py, pz, tab, step, cap};
    return launch_def(s, a, nblk(n));
}

/* TFSF_addSource / TFSF_subSource (source/core/em_TFSF.c:666-772) for all
 * twelve faces of one kind in one launch: block y is the face, a thread
 * one node of it.  Faces of the same component meet along the interface's
 * edges, and the reference adds them one face after the other, so a node
 * shared by several faces is updated by the thread of the first of them,
 * which adds every face's value in face order (bit-exact in FP64).  The
 * frame played this step comes from the host-filled table (slot 0: H
 * faces, 1: E faces; -1: interface off). */
__device__ static inline bool tfsf_has(const TfsfFace &f, const int r[3])
{
    return f.off >= 0 && r[0] >= f.lo[0] && r[0] <= f.hi[0] && r[1] >= f.lo[1] &&
           r[1] <= f.hi[1] && r[2] >= f.lo[2] && r[2] <= f.hi[2];
}

template <class T> struct TfsfArgs {
    T            *F0, *F1, *F2;
    Grid          g;
    TfsfSet       fs;
    const double *data;
    const int    *tab, *step;
    int           cap, slot;
    double        sign;
};

template <class T>
__device__ static inline void tfsf(T *F0, T *F1, T *F2, const Grid &g, const TfsfSet &fs,
                                   const double *data, const int *tab, const int *step, int cap,
                                   int slot, double sign, int f, long t)
{
    const TfsfFace &fc = fs.f[f];
    if (fc.off < 0) return;
    if (t >= fc.vol) return;
    const int frame = tab[2 * (*step % cap) + slot];
    if (frame < 0) return;
    int r[3];
    r[fc.ea] = fc.lo[fc.ea];
    const unsigned pq = (unsigned)t / (unsigned)fc.dQ;
    r[fc.ep] = fc.lo[fc.ep] + (int)pq;
    r[fc.eq] = fc.lo[fc.eq] + (int)((unsigned)t - pq * (unsigned)fc.dQ);
    for (int a = 0; a < 3; ++a)
        if (r[a] < fs.lo_in || r[a] > g.n[a] - 2) return;
    for (int q = 0; q < f; ++q)
        if (fs.f[q].comp == fc.comp && tfsf_has(fs.f[q], r)) return;   /* an earlier face owns it */
    T *F = fc.comp == 0 ? F0 : (fc.comp == 1 ? F1 : F2);
    const long o = gidx(g, r[0], r[1], r[2]);
    double v = F[o];
    for (int q = f; q < 12; ++q) {
        const TfsfFace &h = fs.f[q];
        if (h.comp != fc.comp || !tfsf_has(h, r)) continue;
        const double d = data[h.off + (long)frame * h.vol + (long)(r[h.ep] - h.lo[h.ep]) * h.dQ +
                              (r[h.eq] - h.lo[h.eq])];
        if (sign > 0)
            v += d;
        else
            v -= d;
    }
    F[o] = (T)v;
}

template <class T> struct KernelDef<TfsfArgs<T>> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const TfsfArgs<T> &a, VBlock b, VBlock)
    {
        tfsf<T>(a.F0, a.F1, a.F2, a.g, a.fs, a.data, a.tab, a.step, a.cap, a.slot, a.sign,
                (int)b.y, (long)b.x * TPB + threadIdx.x);
    }
};

template <class T>
int launch_tfsf(hipStream_t s, T *const F[3], const Grid &g, const TfsfSet &fs, long maxvol,
                const double *data, const int *tab, const int *step, int cap, int slot,
                double sign)
{
    if (maxvol <= 0) return MPIC_OK;
    TfsfArgs<T> a = {F[0], F[1], F[2], g, fs, data, tab, step, cap, slot, sign};
    return launch_def(s, a, nblk(maxvol), 12);
}

template <class T>
struct Ptr3 {
    T *p[3];
};

/* TFSF recording (TFSF_preHStep / TFSF_writeMesh, source/core/em_TFSF.c:
 * 573-624): at the start of a step, E on the 12 E faces and H on the 12 H
 * faces (block y = slot, E faces first), each value times its face's
 * factor, into frame (*step - step0) of the face's frames (ep outer, eq
 * inner).  Frames past 'cap' are not stored. */
template <class T> struct TfsfRecArgs {
    Ptr3<T>       E, H;
    Grid          g;
    TfsfSet       fe, fh;
    const double *factor;
    double       *out;
    const int    *step;
    int           step0, cap;
};

template <class T>
__device__ static inline void tfsf_record(const Ptr3<T> &E, const Ptr3<T> &H, const Grid &g,
                                          const TfsfSet &fe, const TfsfSet &fh,
                                          const double *factor, double *out, const int *step,
                                          int step0, int cap, int slot, long t)
{
    const int f = slot % 12;
    const TfsfFace &fc = slot < 12 ? fe.f[f] : fh.f[f];
    if (t >= fc.vol) return;
    const int frame = *step - step0;
    if (frame < 0 || frame >= cap) return;
    int r[3];
    r[fc.ea] = fc.lo[fc.ea];
    const unsigned pq = (unsigned)t / (unsigned)fc.dQ;
    r[fc.ep] = fc.lo[fc.ep] + (int)pq;
    r[fc.eq] = fc.lo[fc.eq] + (int)((unsigned)t - pq * (unsigned)fc.dQ);
    const T *F = slot < 12 ? E.p[fc.comp] : H.p[fc.comp];
    out[fc.off + (long)frame * fc.vol + t] = (double)F[gidx(g, r[0], r[1], r[2])] * factor[f];
}

template <class T> struct KernelDef<TfsfRecArgs<T>> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const TfsfRecArgs<T> &a, VBlock b, VBlock)
    {
        tfsf_record<T>(a.E, a.H, a.g, a.fe, a.fh, a.factor, a.out, a.step, a.step0, a.cap,
                       (int)b.y, (long)b.x * TPB + threadIdx.x);
    }
};

template <class T>
int launch_tfsf_record(hipStream_t s, T *const E[3], T *const H[3], const Grid &g,
                       const TfsfSet &fe, const TfsfSet &fh, long maxvol, const double *factor,
                       double *out, const int *step, int step0, int cap)
{
    if (maxvol <= 0) return MPIC_OK;
    Ptr3<T> Pe = {{E[0], E[1], E[2]}}, Ph = {{H[0], H[1], H[2]}};
    TfsfRecArgs<T> a = {Pe, Ph, g, fe, fh, factor, out, step, step0, cap};
    return launch_def(s, a, nblk(maxvol), 24);
}

/* The run's step counter, advanced once per step on the stream. */
struct TickArgs {
    int *step;
};
template <> struct KernelDef<TickArgs> {
    static constexpr int kBlock = 64;
    __device__ __forceinline__ static void body(const TickArgs &a, VBlock b, VBlock)
    {
        if (threadIdx.x == 0 && b.x == 0) *a.step += 1;
    }
};

int launch_tick(hipStream_t s, int *step)
{
    TickArgs a = {step};
    return launch_def(s, a, 1);
}

/* J boundary fold of one face b (order x-min, x-max, y-min, y-max, z-min,
 * z-max), source/core/plasma_currentWalls.c:132-227. */
struct JFoldArgs {
    JPtrs      f;
    Grid       g;
    int        ea, bc0, bc1, top0;
    const int *jr;         /* x planes [jr[0], jr[1]] of nodes where J may be non-zero */
};

/* A line whose nodes all lie on x planes where J is zero is left alone: the
 * fold of zeros writes zeros.  For y and z walls that is a line on a plane
 * outside jr; an x wall is skipped whole when its layers miss jr (a
 * periodic x fold reads both walls' layers). */
/* Line t of the fold of wall 'top' (0 min, 1 max) of axis ea, boundary
 * code bc. */
__device__ static inline void jfold_line(const JPtrs &f, const Grid &g, int ea, int bc, int top,
                                         const int *jr, long t)
{
    const int ep = (ea == 0) ? 1 : 0, eq = (ea == 2) ? 1 : 2;
    const int np = g.n[ep] + 5, nq = g.n[eq] + 5;
    const int N = g.n[ea];
    if (t >= (long)np * nq) return;
    int r[3];
    const unsigned pq = (unsigned)t / (unsigned)nq;
    r[ep] = -2 + (int)pq;
    r[eq] = -2 + (int)((unsigned)t - pq * (unsigned)nq);
    if (ea != 0 && (r[0] < jr[0] || r[0] > jr[1])) return;
#define AT(a) (r[ea] = (a), gidx(g, r[0], r[1], r[2]))
    /* every value is loaded before any is stored: the nodes of one line are
     * distinct, so this is the reference's sequence with one memory latency */
    if (bc == MPIC_BC_PERIODIC && N < 5) {          /* the layers overlap: in order */
        for (int a = -2; a <= 2; ++a) {
            const long lo = AT(a), hi = AT(a + N);
            for (int c = 0; c < 3; ++c) {
                f.J[c][hi] += f.J[c][lo];
                f.J[c][lo] = f.J[c][hi];
            }
        }
    } else if (bc == MPIC_BC_PERIODIC) {
        long lo[5], hi[5];
        double vl[3][5], vh[3][5];
#pragma unroll
        for (int a = 0; a < 5; ++a) {
            lo[a] = AT(a - 2);
            hi[a] = AT(a - 2 + N);
        }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int a = 0; a < 5; ++a) {
                vl[c][a] = f.J[c][lo[a]];
                vh[c][a] = f.J[c][hi[a]];
            }
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int a = 0; a < 5; ++a) {
                const double v = vh[c][a] + vl[c][a];
                f.J[c][hi[a]] = v;
                f.J[c][lo[a]] = v;
            }
    } else {
        const int pl = top ? N : 0, so = top ? 1 : -1, in_ = 1 - top, ex_ = top;
        double *Ja = f.J[ea], *Jp = f.J[ep], *Jq = f.J[eq];
        long oin = AT(pl + in_), oex = AT(pl + ex_), om1 = AT(pl - so), op1 = AT(pl + so);
        long om2 = AT(pl - 2 * so), op2 = AT(pl + 2 * so), oin2 = AT(pl + in_ - so),
             oex2 = AT(pl + ex_ + so), o0 = AT(pl);
        const double a_in = Ja[oin], a_ex = Ja[oex], a_in2 = Ja[oin2], a_ex2 = Ja[oex2];
        const double p_m1 = Jp[om1], p_p1 = Jp[op1], p_m2 = Jp[om2], p_p2 = Jp[op2];
        const double q_m1 = Jq[om1], q_p1 = Jq[op1], q_m2 = Jq[om2], q_p2 = Jq[op2];
        const double A1 = a_in - a_ex, A2 = a_in2 - a_ex2;
        const double P1 = p_m1 + p_p1, P2 = p_m2 + p_p2, Q1 = q_m1 + q_p1, Q2 = q_m2 + q_p2;
        Ja[oin] = A1;
        Ja[oex] = A1;
        Ja[oin2] = A2;
        Ja[oex2] = A2;
        Jp[om1] = P1;
        Jp[op1] = -P1;
        Jq[om1] = Q1;
        Jq[op1] = -Q1;
        Jp[om2] = P2;
        Jp[op2] = -P2;
        Jq[om2] = Q2;
        Jq[op2] = -Q2;
        Jp[o0] = 0;
        Jq[o0] = 0;
    }
#undef AT
}

__device__ static inline void jfold(const JPtrs &f, const Grid &g, int ea, int bc0, int bc1,
                                    int top0, const int *jr, const VBlock &b)
{
    /* wall top0 + block y of the axis: 0 min, 1 max */
    const int top = top0 + (int)b.y, bc = top ? bc1 : bc0;
    const int N = g.n[ea];
    if (ea == 0) {
        const bool lo_wall = jr[0] <= 2 && jr[1] >= -2, hi_wall = jr[0] <= N + 2 && jr[1] >= N - 2;
        if (bc == MPIC_BC_PERIODIC ? !(lo_wall || hi_wall) : !(top ? hi_wall : lo_wall)) return;
    }
    jfold_line(f, g, ea, bc, top, jr, (long)b.x * TPB + threadIdx.x);
}

/* The J fold of both walls of axis ea (min first).  A periodic axis is one
 * fold; two non-periodic walls of an axis with at least 8 cells touch the
 * disjoint layers [-2, 2] and [N-2, N+2] and share one launch. */
template <> struct KernelDef<JFoldArgs> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const JFoldArgs &a, VBlock b, VBlock)
    {
        jfold(a.f, a.g, a.ea, a.bc0, a.bc1, a.top0, a.jr, b);
    }
};

/* The y and z folds (min then max wall of y, then of z) of x plane
 * b.x - 2, one workgroup per plane, a barrier between consecutive walls: a
 * y or z fold line keeps i fixed (see the wall planes above).  Planes
 * outside jr hold no current and are left alone. */
struct JFoldPlanesArgs {
    JPtrs      f;
    Grid       g;
    int        bc[6];
    const int *jr;
};

template <> struct KernelDef<JFoldPlanesArgs> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const JFoldPlanesArgs &a, VBlock b, VBlock)
    {
        const int i = (int)b.x - 2;
        if (i < a.jr[0] || i > a.jr[1]) return;          /* the whole workgroup */
        bool first = true;
        for (int ea = 1; ea < 3; ++ea) {
            const int eq = (ea == 2) ? 1 : 2;            /* ep = 0: a line's p is its plane */
            const int nq = a.g.n[eq] + 5;
            const int bc0 = a.bc[ea], bc1 = a.bc[3 + ea];
            const int walls = bc0 == MPIC_BC_PERIODIC ? 1 : 2;   /* a periodic axis folds once */
            /* the two walls of an axis with at least 8 cells touch disjoint
             * layers ([-2, 2] and [N-2, N+2]): one phase */
            const bool together = walls == 2 && a.g.n[ea] >= 8;
            for (int top = 0; top < walls; ++top) {
                if (!first && !(together && top == 1)) __syncthreads();
                first = false;
                const long t0 = (long)(i + 2) * nq;
                for (int u = threadIdx.x; u < nq; u += TPB)
                    jfold_line(a.f, a.g, ea, top ? bc1 : bc0, top, a.jr, t0 + u);
            }
        }
    }
};

int launch_jfold_planes(hipStream_t s, const JPtrs &f, const Grid &g, const int bc[6],
                        const int *jr)
{
    JFoldPlanesArgs a;
    a.f = f;
    a.g = g;
    for (int q = 0; q < 6; ++q) a.bc[q] = bc[q];
    a.jr = jr;
    return launch_def(s, a, (unsigned)(g.n[0] + 5));
}

#define TRY_LAUNCH(x)                   \
    do {                                \
        const int rc_ = (x);            \
        if (rc_ != MPIC_OK) return rc_; \
    } while (0)

int launch_jfold(hipStream_t s, const JPtrs &f, const Grid &g, int ea, int bc0, int bc1,
                 const int *jr)
{
    const int ep = (ea == 0) ? 1 : 0, eq = (ea == 2) ? 1 : 2;
    const long n = (long)(g.n[ep] + 5) * (g.n[eq] + 5);
    const unsigned nb = (unsigned)nblk(n);
    JFoldArgs a = {f, g, ea, bc0, bc1, 0, jr};
    if (bc0 == MPIC_BC_PERIODIC) return launch_def(s, a, nb, 1);
    if (g.n[ea] >= 8) return launch_def(s, a, nb, 2);
    TRY_LAUNCH(launch_def(s, a, nb, 1));
    a.top0 = 1;
    return launch_def(s, a, nb, 1);
}

/* Field energy sums, source/core/em.c:226-252 (FP64; block partial sums
 * combined with FP64 atomics, so the order differs from the CPU loop). */
template <class T>
__global__ void k_energy(FieldsT<T> f, Grid g, double *acc)
{
    __shared__ double se[TPB], sb[TPB];
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tot = (long)g.n[0] * g.n[1] * g.n[2];
    double we = 0, wb = 0;
    for (long c = t; c < tot; c += (long)gridDim.x * blockDim.x) {
        const unsigned cu = (unsigned)c, r = cu / (unsigned)g.n[2], r2 = r / (unsigned)g.n[1];
        const int k = (int)(cu - r * (unsigned)g.n[2]), j = (int)(r - r2 * (unsigned)g.n[1]),
                  i = (int)r2;
        double ex = f.E[0][gidx(g, i + 1, j, k)], ey = f.E[1][gidx(g, i, j + 1, k)],
               ez = f.E[2][gidx(g, i, j, k + 1)], hx = f.H[0][gidx(g, i, j + 1, k + 1)],
               hy = f.H[1][gidx(g, i + 1, j, k + 1)], hz = f.H[2][gidx(g, i + 1, j + 1, k)];
        we += ex * ex + ey * ey + ez * ez;
        wb += hx * hx + hy * hy + hz * hz;
    }
    se[threadIdx.x] = we;
    sb[threadIdx.x] = wb;
    __syncthreads();
    for (int w = TPB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            se[threadIdx.x] += se[threadIdx.x + w];
            sb[threadIdx.x] += sb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(acc + 0, se[0]);
        atomicAdd(acc + 1, sb[0]);
    }
}

template <class T>
int launch_energy(hipStream_t s, const FieldsT<T> &f, const Grid &g, double *acc)
{
    long tot = (long)g.n[0] * g.n[1] * g.n[2];
    long nb = (tot + TPB - 1) / TPB;
    if (nb > 2048) nb = 2048;
    k_energy<T><<<(unsigned)nb, TPB, 0, s>>>(f, g, acc);
    MPIC_CHECK(hipGetLastError());
    return MPIC_OK;
}

/* VSP_vectorSweep / VSP_scalarSweep (source/core/plasma_VSP.c:32-176): the
 * (beta, alpha, beta) convolution along axis 'ax' of every line of the
 * padded grid (nodes [-2, n+2] on every axis), in place; one thread per
 * line keeps the untouched previous value in a register, as the reference
 * sweep does, so each new value takes the old values of its neighbours and
 * the arithmetic is the reference's expression for expression.  Lines are
 * dealt so that consecutive threads take consecutive k (ax = 0, 1). */
struct VspArgs {
    double *F0, *F1, *F2;
    int     ncomp;
    Grid    g;
    int     ax;
    double  alpha, beta;
};

__device__ static inline void vsp(double *F0, double *F1, double *F2, int ncomp, const Grid &g,
                                  int ax, double alpha, double beta, long t)
{
    const int p = ax == 0 ? 1 : 0, q = ax == 2 ? 1 : 2;
    const int np = g.n[p] + 5, nq = g.n[q] + 5;
    if (t >= (long)np * nq) return;
    const int a = (int)(t / nq) - 2, b = (int)(t % nq) - 2;
    int r[3];
    r[p] = a;
    r[q] = b;
    r[ax] = -2;
    const long base = gidx(g, r[0], r[1], r[2]);
    const long st = ax == 0 ? (long)g.GY * g.GZ : (ax == 1 ? (long)g.GZ : 1L);
    const int len = g.n[ax] + 5;
    double *F[3] = {F0, F1, F2};
    for (int c = 0; c < ncomp; ++c) {
        double *f = F[c] + base;
        double un = f[0];
        f[0] = beta * f[st] + alpha * f[0];
        double cur = f[st];
        for (int k = 1; k < len - 1; ++k) {
            const double nx = f[(k + 1) * st];
            const double v = alpha * (cur) + beta * (un + nx);
            un = cur;
            f[k * st] = v;
            cur = nx;
        }
        f[(len - 1) * st] = beta * un + alpha * cur;
    }
}

template <> struct KernelDef<VspArgs> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const VspArgs &a, VBlock b, VBlock)
    {
        vsp(a.F0, a.F1, a.F2, a.ncomp, a.g, a.ax, a.alpha, a.beta, (long)b.x * TPB + threadIdx.x);
    }
};

int launch_vsp(hipStream_t s, double *const F[3], int ncomp, const Grid &g, double alpha)
{
    const double beta = 0.5 * (1.0 - alpha);
    for (int ax = 2; ax >= 0; --ax) {
        const int p = ax == 0 ? 1 : 0, q = ax == 2 ? 1 : 2;
        const long lines = (long)(g.n[p] + 5) * (g.n[q] + 5);
        VspArgs a = {F[0], ncomp > 1 ? F[1] : nullptr, ncomp > 2 ? F[2] : nullptr, ncomp, g, ax,
                     alpha, beta};
        TRY_LAUNCH(launch_def(s, a, nblk(lines)));
    }
    return MPIC_OK;
}

/* em_energy of the loop body (source/core/em.c:226-252) without atomics:
 * one wavefront per cell row (i, j), lanes along k, each workgroup's partial
 * sums stored in part[2 * block + {0, 1}] and combined in fixed order by
 * k_diag_ring. */
/* Row of the diagnostics ring of the current step (device step counter
 * *step): the two energy sums of k_energy_part's nparts partials (fixed
 * order) and the three kinetic sums wk[0..2] of the push. */
struct DiagRingArgs {
    const double *part;
    int           nparts;
    const double *wk;
    double       *ring;
    int          *step;
    int           cap;
    int           tick;    /* advance the step counter after the row (the step's last kernel) */
};

#define RING_TPB 1024

__device__ static inline void diag_ring(const double *part, int nparts, const double *wk,
                                        double *ring, int *step, int cap, int tick)
{
    __shared__ double se[RING_TPB], sb[RING_TPB];
    double a = 0, c = 0;
    for (int q = threadIdx.x; q < nparts; q += RING_TPB) {
        a += part[2 * q];
        c += part[2 * q + 1];
    }
    se[threadIdx.x] = a;
    sb[threadIdx.x] = c;
    __syncthreads();
    for (int w = RING_TPB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            se[threadIdx.x] += se[threadIdx.x + w];
            sb[threadIdx.x] += sb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double *row = ring + 5 * (long)(*step % cap);
        row[0] = se[0];
        row[1] = sb[0];
        row[2] = wk[0];
        row[3] = wk[1];
        row[4] = wk[2];
        if (tick) *step += 1;
    }
}

template <> struct KernelDef<DiagRingArgs> {
    static constexpr int kBlock = RING_TPB;
    __device__ __forceinline__ static void body(const DiagRingArgs &a, VBlock, VBlock)
    {
        diag_ring(a.part, a.nparts, a.wk, a.ring, a.step, a.cap, a.tick);
    }
};

int launch_diag_ring(hipStream_t s, const double *part, int nparts, const double *wk,
                     double *ring, int *step, int cap, int tick)
{
    DiagRingArgs a = {part, nparts, wk, ring, step, cap, tick};
    return launch_def(s, a, 1);
}

__global__ void k_zero(double *p, long n)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    for (; t < n; t += (long)gridDim.x * blockDim.x) p[t] = 0.0;
}

int launch_zero(hipStream_t s, double *p, long n)
{
    long nb = (n + TPB - 1) / TPB;
    if (nb > 65536) nb = 65536;
    k_zero<<<(unsigned)nb, TPB, 0, s>>>(p, n);
    MPIC_CHECK(hipGetLastError());
    return MPIC_OK;
}

/* J cleared before the push, only on the x planes the previous step's
 * deposit can have written, and the three temperature sums.
 *
 * A tiled push deposits the markers of the active tiles only (the list is
 * in tile order, tiles i-major), each into nodes [4u-1, 4u+5] along x of
 * its tile u; the fold and the VSP filter leave zero where J is zero except
 * near the x walls (a periodic fold moves current across the box), and the
 * filter spreads it by one node.  So the planes written in a step are those
 * of [4 u_first - 3, 4 u_last + 7] for the step's first and last active
 * tiles, or the whole grid when that range comes within 4 planes of an x
 * wall.  Each launch clears the range recorded by the previous step's
 * launch (slot step % 2) and records the range of the list it sees, which
 * the coming push uses (slot (step + 1) % 2), also as range[4..5], where
 * the E updates of the same step find the planes on which J may be non-zero.
 * 'full' (an untiled push) clears everything and records the whole grid. */
struct JClearArgs {
    double       *J;          /* 3 components, ncell apart */
    long          ncell, plane;
    int           n0, tpp;    /* cells along x, tiles per x layer of tiles */
    double       *acc;        /* 3 sums */
    int          *range;      /* [2][2] plane ranges by step parity, [2] this step's */
    const int    *step, *active, *nactive;
    int           full;
};

__device__ static inline void jclear(const JClearArgs &a, const VBlock &b, const VBlock &gd)
{
    const int s = *a.step;
    int lo = -2, hi = a.n0 + 2;
    if (!a.full) {
        lo = a.range[2 * (s & 1)];
        hi = a.range[2 * (s & 1) + 1];
    }
    if (b.x == 0 && threadIdx.x == 0) {
        a.acc[0] = 0.0;
        a.acc[1] = 0.0;
        a.acc[2] = 0.0;
        int nlo = -2, nhi = a.n0 + 2;
        if (!a.full) {
            const int na = *a.nactive;
            if (na <= 0) {
                nlo = 0;
                nhi = -1;                                    /* nothing written */
            } else {
                nlo = 4 * (a.active[0] / a.tpp) - 3;
                nhi = 4 * (a.active[na - 1] / a.tpp) + 7;
                if (nlo <= 3 || nhi >= a.n0 - 3) {
                    nlo = -2;
                    nhi = a.n0 + 2;
                }
            }
        }
        a.range[2 * ((s + 1) & 1)] = nlo;
        a.range[2 * ((s + 1) & 1) + 1] = nhi;
        a.range[4] = nlo;                                    /* this step's, for E -= tau J */
        a.range[5] = nhi;
    }
    if (hi < lo) return;
    typedef double D2 __attribute__((ext_vector_type(2)));
    const long base = (long)(lo + MPIC_GHOST) * a.plane;
    const long cnt2 = (long)(hi - lo + 1) * a.plane / 2;   /* planes hold an even count */
    const long stride = (long)gd.x * TPB;
    const D2 z = {0.0, 0.0};
    for (int c = 0; c < 3; ++c) {
        D2 *p = (D2 *)(a.J + c * a.ncell + base);
        for (long u = (long)b.x * TPB + threadIdx.x; u < cnt2; u += stride) p[u] = z;
    }
}

template <> struct KernelDef<JClearArgs> {
    static constexpr int kBlock = TPB;
    __device__ __forceinline__ static void body(const JClearArgs &a, VBlock b, VBlock gd)
    {
        jclear(a, b, gd);
    }
};

static JClearArgs jclear_args(double *J, const Grid &g, double *acc, int *range, const int *step,
                              const int *active, const int *nactive, int tpp, int full)
{
    JClearArgs a = {J, g.ncell, (long)g.GY * g.GZ, g.n[0], tpp, acc, range, step, active,
                    nactive, full};
    return a;
}

/* 16-byte stores; enough workgroups to stream the whole grid at full rate,
 * few enough that a narrow range costs little more than a launch. */
static unsigned jclear_blocks(const Grid &g)
{
    const long cnt2 = (long)(g.n[0] + 5) * g.GY * g.GZ / 2;
    long nb = (cnt2 + 4 * TPB - 1) / (4 * TPB);
    if (nb > 4096) nb = 4096;
    if (nb < 1) nb = 1;
    return (unsigned)nb;
}

int launch_jclear(hipStream_t s, double *J, const Grid &g, double *acc, int *range,
                  const int *step, const int *active, const int *nactive, int tpp, int full)
{
    return launch_def(s, jclear_args(J, g, acc, range, step, active, nactive, tpp, full),
                      jclear_blocks(g));
}

/* CIC rho fold of one face, source/core/plasma_rho.c:23-96. */
__global__ void k_rho_fold(double *rho, Grid g, int b, int bc)
{
    const int ea = b >> 1, top = b & 1;
    const int ep = (ea == 0) ? 1 : 0, eq = (ea == 2) ? 1 : 2;
    const int np = g.n[ep] + 5, nq = g.n[eq] + 5;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)np * nq) return;
    int r[3];
    const unsigned pq = (unsigned)t / (unsigned)nq;
    r[ep] = -2 + (int)pq;
    r[eq] = -2 + (int)((unsigned)t - pq * (unsigned)nq);
    const int N = g.n[ea];
#define AT(a) (r[ea] = (a), gidx(g, r[0], r[1], r[2]))
    if (bc == MPIC_BC_PERIODIC) {
        for (int a = -2; a <= 2; ++a) {
            long lo = AT(a), hi = AT(a + N);
            rho[hi] += rho[lo];
            rho[lo] = rho[hi];
        }
    } else {
        const int pl = top ? N : 0, so = top ? 1 : -1;
        long om1 = AT(pl - so), op1 = AT(pl + so), om2 = AT(pl - 2 * so), op2 = AT(pl + 2 * so),
             o0 = AT(pl);
        rho[om1] += rho[op1];
        rho[om2] += rho[op2];
        rho[o0] = 0;
        rho[op1] = -rho[om1];
        rho[op2] = -rho[om2];
    }
#undef AT
}

int launch_rho_fold(hipStream_t s, double *rho, const Grid &g, int b, int bc)
{
    const int ea = b >> 1;
    const int ep = (ea == 0) ? 1 : 0, eq = (ea == 2) ? 1 : 2;
    if (bc == MPIC_BC_PERIODIC && (b & 1)) return MPIC_OK;
    long n = (long)(g.n[ep] + 5) * (g.n[eq] + 5);
    k_rho_fold<<<nblk(n), TPB, 0, s>>>(rho, g, b, bc);
    MPIC_CHECK(hipGetLastError());
    return MPIC_OK;
}

/* Gauss / continuity residuals over nodes [0, n]^3,
 * source/core/plasma_gaussTest.c:110-149.  acc = {max r_J, max r_E} as
 * unsigned bit patterns of non-negative doubles (ordered like the values). */
template <class T>
__global__ void k_gauss(FieldsT<T> f, const double *rho1, const double *rho2, Grid g,
                        unsigned long long *acc)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int NX = g.n[0] + 1, NY = g.n[1] + 1, NZ = g.n[2] + 1;
    if (t >= (long)NX * NY * NZ) return;
    const unsigned tu = (unsigned)t, r = tu / (unsigned)NZ, r2 = r / (unsigned)NY;
    const int k = (int)(tu - r * (unsigned)NZ), j = (int)(r - r2 * (unsigned)NY), i = (int)r2;
    long o = gidx(g, i, j, k);
    const double h1 = g.h[0], h2 = g.h[1], h3 = g.h[2];
    double t0 = (rho2[o] - rho1[o]) / g.tau + (f.J[0][gidx(g, i + 1, j, k)] - f.J[0][o]) / h1 +
                (f.J[1][gidx(g, i, j + 1, k)] - f.J[1][o]) / h2 +
                (f.J[2][gidx(g, i, j, k + 1)] - f.J[2][o]) / h3;
    double t1 = ((double)f.E[0][gidx(g, i + 1, j, k)] - (double)f.E[0][o]) / h1 +
                ((double)f.E[1][gidx(g, i, j + 1, k)] - (double)f.E[1][o]) / h2 +
                ((double)f.E[2][gidx(g, i, j, k + 1)] - (double)f.E[2][o]) / h3 - rho2[o];
    t0 = fabs(t0);
    t1 = fabs(t1);
    atomicMax(acc + 0, (unsigned long long)__double_as_longlong(t0));
    atomicMax(acc + 1, (unsigned long long)__double_as_longlong(t1));
}

template <class T>
int launch_gauss(hipStream_t s, const FieldsT<T> &f, const double *rho1, const double *rho2,
                 const Grid &g, double *acc)
{
    long n = (long)(g.n[0] + 1) * (g.n[1] + 1) * (g.n[2] + 1);
    k_gauss<T><<<nblk(n), TPB, 0, s>>>(f, rho1, rho2, g, (unsigned long long *)acc);
    MPIC_CHECK(hipGetLastError());
    return MPIC_OK;
}

/* Node values of the Tecplot records and the probes (source/IO_tecplot.c:
 * 131-153, source/core/diag_probe.c:361-383), in FP64: kind 0 (E, J) is the
 * mean of the two values along the component's own axis, kind 1 (H) the mean
 * of the four around the node in the plane normal to it, summed in the
 * reference's order. */
template <class T>
__device__ static inline void node_values(T *const F[3], const Grid &g, int kind, int i, int j,
                                          int k, double v[3])
{
    if (kind == 0) {
        v[0] = 0.5 * ((double)F[0][gidx(g, i, j, k)] + (double)F[0][gidx(g, i + 1, j, k)]);
        v[1] = 0.5 * ((double)F[1][gidx(g, i, j, k)] + (double)F[1][gidx(g, i, j + 1, k)]);
        v[2] = 0.5 * ((double)F[2][gidx(g, i, j, k)] + (double)F[2][gidx(g, i, j, k + 1)]);
    } else {
        v[0] = 0.25 * ((double)F[0][gidx(g, i, j, k)] + (double)F[0][gidx(g, i, j + 1, k)] +
                       (double)F[0][gidx(g, i, j, k + 1)] + (double)F[0][gidx(g, i, j + 1, k + 1)]);
        v[1] = 0.25 * ((double)F[1][gidx(g, i, j, k)] + (double)F[1][gidx(g, i + 1, j, k)] +
                       (double)F[1][gidx(g, i, j, k + 1)] + (double)F[1][gidx(g, i + 1, j, k + 1)]);
        v[2] = 0.25 * ((double)F[2][gidx(g, i, j, k)] + (double)F[2][gidx(g, i + 1, j, k)] +
                       (double)F[2][gidx(g, i, j + 1, k)] + (double)F[2][gidx(g, i + 1, j + 1, k)]);
    }
}

/* nodes [0, n] (k fastest), three interleaved doubles per node */
template <class T>
__global__ void k_node_fields(Ptr3<T> F, Grid g, int kind, double *out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned nz = (unsigned)g.n[2] + 1, ny = (unsigned)g.n[1] + 1;
    if (t >= (long)(g.n[0] + 1) * ny * nz) return;
    const unsigned r = (unsigned)t / nz, i = r / ny;
    const int k = (int)((unsigned)t - r * nz), j = (int)(r - i * ny);
    double v[3];
    node_values(F.p, g, kind, (int)i, j, k, v);
    out[3 * t] = v[0];
    out[3 * t + 1] = v[1];
    out[3 * t + 2] = v[2];
}

template <class T>
int launch_node_fields(hipStream_t s, T *const F[3], const Grid &g, int kind, double *out)
{
    Ptr3<T> P = {{F[0], F[1], F[2]}};
    const long n = (long)(g.n[0] + 1) * (g.n[1] + 1) * (g.n[2] + 1);
    k_node_fields<T><<<nblk(n), TPB, 0, s>>>(P, g, kind, out);
    MPIC_CHECK(hipGetLastError());
    return MPIC_OK;
}

/* probe p at node ijk[3p..3p+2] (checked on the host to lie in [0, n-1]):
 * out[6p..6p+5] = Ex Ey Ez Hx Hy Hz */
template <class T>
__global__ void k_probes(Ptr3<T> E, Ptr3<T> H, Grid g, const int *ijk, int np, double *out)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= np) return;
    const int i = ijk[3 * p], j = ijk[3 * p + 1], k = ijk[3 * p + 2];
    double v[3];
    node_values(E.p, g, 0, i, j, k, v);
    out[6 * p] = v[0];
    out[6 * p + 1] = v[1];
    out[6 * p + 2] = v[2];
    node_values(H.p, g, 1, i, j, k, v);
    out[6 * p + 3] = v[0];
    out[6 * p + 4] = v[1];
    out[6 * p + 5] = v[2];
}

template <class T>
int launch_probes(hipStream_t s, T *const E[3], T *const H[3], const Grid &g, const int *ijk,
                  int np, double *out)
{
    if (np <= 0) return MPIC_OK;
    Ptr3<T> Pe = {{E[0], E[1], E[2]}}, Ph = {{H[0], H[1], H[2]}};
    k_probes<T><<<(np + TPB - 1) / TPB, TPB, 0, s>>>(Pe, Ph, g, ijk, np, out);
    MPIC_CHECK(hipGetLastError());
    return MPIC_OK;
}

/* The H update paired with the J clear (the H half step of the loop body:
 * the update never touches J) or with the y and z J folds (the H step: the
 * x folds have run before). */
template <class T>
int launch_h_update_jclear(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &flush,
                           int full, double *part, double *J, double *acc, int *range,
                           const int *step, const int *active, const int *nactive, int tpp,
                           int jfull)
{
    const HUpdateArgs<T> a = h_args(f, g, flush, full, part, LaserPlane());
    return launch_pair(s, a, dim3(grid_span(a.sp, g.n[0] + 2)),
                       jclear_args(J, g, acc, range, step, active, nactive, tpp, jfull),
                       dim3(jclear_blocks(g)));
}

template <class T>
int launch_h_update_jfold(hipStream_t s, const FieldsT<T> &f, const Grid &g, const BoxSet &flush,
                          int full, const LaserPlane &lp, const JPtrs &jp, const int bc[6],
                          const int *jr)
{
    const HUpdateArgs<T> a = h_args(f, g, flush, full, nullptr, lp);
    JFoldPlanesArgs b;
    b.f = jp;
    b.g = g;
    for (int q = 0; q < 6; ++q) b.bc[q] = bc[q];
    b.jr = jr;
    return launch_pair(s, a, dim3(grid_span(a.sp, g.n[0] + 2)), b, dim3(g.n[0] + 5));
}

/* explicit instantiations for the two field precisions */
#define INST(T)                                                                               \
    template int launch_cap_H<T>(hipStream_t, const FieldsT<T> &, const Grid &, const BoxSet &); \
    template int launch_cap_E<T>(hipStream_t, const FieldsT<T> &, const Grid &, const BoxSet &); \
    template int launch_h_update<T>(hipStream_t, const FieldsT<T> &, const Grid &,                \
                                    const BoxSet &, int, double *, const LaserPlane &);         \
    template int launch_h_update_jclear<T>(hipStream_t, const FieldsT<T> &, const Grid &,         \
                                           const BoxSet &, int, double *, double *, double *,   \
                                           int *, const int *, const int *, const int *, int,   \
                                           int);                                                \
    template int launch_h_update_jfold<T>(hipStream_t, const FieldsT<T> &, const Grid &,          \
                                          const BoxSet &, int, const LaserPlane &,              \
                                          const JPtrs &, const int *, const int *);             \
    template int launch_e_fused<T>(hipStream_t, const FieldsT<T> &, const Grid &,                 \
                                   const BoxSet &, const int *, const LaserPlane &);            \
    template int launch_e_update<T>(hipStream_t, const FieldsT<T> &, const Grid &,                \
                                    const BoxSet &, int, int);                                  \
    template int launch_e_tauJ<T>(hipStream_t, const FieldsT<T> &, const Grid &, int *,         \
                                  const int *);         \
    template int launch_laser<T>(hipStream_t, T *, T *, const Grid &, const LaserPlane &);       \
    template int launch_spot<T>(hipStream_t, T *, T *, const Grid &, int, const double *,        \
                                const double *, const double *, const int *, int);              \
    template int launch_energy<T>(hipStream_t, const FieldsT<T> &, const Grid &, double *);      \
    template int launch_tfsf<T>(hipStream_t, T *const[3], const Grid &, const TfsfSet &, long,   \
                                const double *, const int *, const int *, int, int, double);   \
    template int launch_gauss<T>(hipStream_t, const FieldsT<T> &, const double *,                \
                                 const double *, const Grid &, double *);                       \
    template int launch_node_fields<T>(hipStream_t, T *const[3], const Grid &, int, double *);   \
    template int launch_tfsf_record<T>(hipStream_t, T *const[3], T *const[3], const Grid &,      \
                                       const TfsfSet &, const TfsfSet &, long, const double *,  \
                                       double *, const int *, int, int);                        \
    template int launch_probes<T>(hipStream_t, T *const[3], T *const[3], const Grid &,           \
                                  const int *, int, double *);
INST(double)
INST(float)
