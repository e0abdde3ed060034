/*
 * opic.c -- CPU oracle for the Mandor PIC time step.  TEST INFRASTRUCTURE.
 *
 * Restates, for one MPI rank owning the whole 3D domain, the reference path
 * of cvarin/mandor3 (file:line citations refer to that repository):
 *
 *   push / gather / Boris      source/core/plasma.c:95-263, source/misc_PIC.h
 *   current deposit            source/core/plasma_currentKernel.c:20-279
 *   marker bookkeeping         source/core/plasma.c:99-106,239-253,343-414;
 *                              source/core/plasma_IO.c:124-160,227-287
 *   particle walls             source/core/plasma_walls.c:27-104
 *   J boundary fold            source/core/plasma_currentWalls.c:27-121
 *   Yee interior steps         source/core/em.c:55-222
 *   boundary caps              source/core/em_caps.c:68-809
 *   periodic / mirror / Mur    source/core/emCap_{periodic,mirror,Mur}.c
 *   laser soft source          source/scpic/em_mirror.c:667-717
 *   Gaussian hard spot         source/core/em_gaussSpot.c:42-156
 *   energies                   source/core/em.c:227-252, plasma.c:140-143
 *   CIC rho + Gauss residuals  source/core/plasma_rho.c:23-173,
 *                              source/core/plasma_gaussTest.c:110-149
 *
 * Compiled with -O2 -ffp-contract=off (no FMA), IEEE '/' and sqrt: every
 * floating expression below keeps the reference's association, so the
 * oracle reproduces the reference binaries bit for bit.
 *
 * How each part is pinned (tests/test_oracle_golden.py,
 * tests/test_oracle_math.py):
 *   - push, deposit, walls, J fold, Yee, caps, periodic/mirror/Mur, laser
 *     injection, marker order: per-step SHA-256 of E, H and the marker array
 *     equal to the reference build's records on golden fixtures (periodic
 *     box; all-Mur box with laser; mixed periodic/mirror/Mur box).
 *   - energies: against the reference diag.dat (printed to 7 digits) and
 *     by recomputation from the pinned fields.
 *   - rho / Gauss residuals: against the reference MC_GAUSS_TEST build's
 *     gauss_test.dat, and the discrete continuity identity (~1e-14).
 *   - independently of the fixtures: continuity at round-off for single
 *     markers crossing 0-3 faces (deposit), p linear in t under uniform E
 *     and the exact 2 atan|t| rotation under uniform H (Boris), and the
 *     exact discrete standing-wave solution of the Yee leapfrog (field
 *     update order of the main loop).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include "opic.h"

/* ------------------------------------------------------------------ */
/* Meshes: box of nodes [lo, hi] per axis, nc doubles per node, k fastest. */
typedef struct { int lo[3], hi[3]; } box3;

typedef struct {
    box3    b;
    int     nc;
    long    si, sj;
    double *d;
} mesh_t;

static inline double *AT(const mesh_t *m, int i, int j, int k)
{
    return m->d + (long)m->nc * ((long)(i - m->b.lo[0]) * m->si +
                                 (long)(j - m->b.lo[1]) * m->sj +
                                 (long)(k - m->b.lo[2]));
}
static inline double *ATv(const mesh_t *m, const int r[3])
{
    return AT(m, r[0], r[1], r[2]);
}

static inline double *jat(const mesh_t *m, int r[3], int ea, int v)
{
    r[ea] = v;
    return ATv(m, r);
}

static void mesh_alloc(mesh_t *m, box3 b, int nc)
{
    m->b = b;
    m->nc = nc;
    m->sj = b.hi[2] - b.lo[2] + 1;
    m->si = m->sj * (b.hi[1] - b.lo[1] + 1);
    long n = m->si * (b.hi[0] - b.lo[0] + 1) * nc;
    m->d = (double *)malloc(sizeof(double) * (size_t)n);
    /* NaN fill, as the reference with MC_FILL_MESH_WITH_NAN
     * (source/type_mesh.c:30-38). */
    memset(m->d, 0xff, sizeof(double) * (size_t)n);
}
static void mesh_zero(mesh_t *m)
{
    long n = m->si * (m->b.hi[0] - m->b.lo[0] + 1) * m->nc;
    memset(m->d, 0, sizeof(double) * (size_t)n);
}
static void mesh_free(mesh_t *m) { free(m->d); m->d = NULL; }

static box3 box_make(int a0, int a1, int a2, int b0, int b1, int b2)
{
    box3 r = {{a0, a1, a2}, {b0, b1, b2}};
    return r;
}

/* ------------------------------------------------------------------ */
typedef struct {           /* reference marker_t, source/type_marker.h:12-25 */
    double x, y, z, vx, vy, vz;
    float  rho, qDivM;
    int    id;
    int    pad;
} mk_t;

typedef struct {           /* Mur buffer of one boundary on one cap,
                              source/core/em_caps.c:42-49 */
    double *buf;
    double  alpha;
    int     ea, top, on;
} mur_t;

typedef struct {           /* source/core/em_caps.c:54-61 */
    mesh_t m;
    box3   upd, fl;
    int    ncopy;
    box3   copy[6];
    int    copy_src[6];
    mur_t  mur[6];
} cap_t;

typedef struct {           /* mirror_t subset used by core.out */
    double  A, T0, delay, dt, phi0;
    int     forward, plane;
    double *mapE, *mapH;
} laser_t;

/* Hard Gaussian spot source, source/core/em_gaussSpot.c:20-30. */
typedef struct {
    int    on, i;
    double ampl, omega, x, y, z, width, duration, front, plateau, Ey, Ez;
} spot_t;

/* One TFSF face (tfsfFace_t, source/core/em_TFSF.c:64-76): the played
 * frames (nframes x volume, ep outer / eq inner) replace the file stream. */
typedef struct {
    int     comp, ea, ep, eq;
    int     lo[3], hi[3];
    long    vol;
    double *data;          /* NULL: face opened or absent */
} tface_t;

/* TFSF playback state (source/core/em_TFSF.c:124-133). */
typedef struct {
    int     on, mode, last, frame;
    tface_t E[12], H[12];
} tfsf_t;

struct opic_s {
    int    n[3];
    double h[3], L[3], tau, time;
    int    bc[6];          /* xmin ymin zmin xmax ymax zmax */
    mesh_t E, H, J, rho1, rho2;
    cap_t  capE[6], capH[6];
    /* markers: plasma[0, nShell) shell, [nCore, nAll) core */
    mk_t  *pl;
    long   nAll, nCore, nShell;
    double core1[3], core2[3], shell1[3], shell2[3];
    double Wx, Wy, Wz;
    int    nlaser;
    laser_t laser[4];
    spot_t spot;
    tfsf_t tfsf;
    double vsp;            /* VSP convolution weight alpha (1: identity) */
    int    inited;
};

/* ------------------------------------------------------------------ */
opic_t *opic_create(const int n[3], const double h[3], const double L[3],
                    double tau, const int bc[6], double time)
{
    opic_t *s = (opic_t *)calloc(1, sizeof(opic_t));
    for (int a = 0; a < 3; ++a) {
        s->n[a] = n[a];
        s->h[a] = h[a];
        s->L[a] = L[a];
    }
    for (int b = 0; b < 6; ++b) s->bc[b] = bc[b];
    s->tau = tau;
    s->time = time;
    /* Ghost extents, source/core/main.c:109-115. */
    mesh_alloc(&s->E, box_make(-1, -1, -1, n[0] + 1, n[1] + 1, n[2] + 1), 3);
    mesh_alloc(&s->H, box_make(-1, -1, -1, n[0] + 1, n[1] + 1, n[2] + 1), 3);
    mesh_alloc(&s->J, box_make(-2, -2, -2, n[0] + 2, n[1] + 2, n[2] + 2), 3);
    mesh_alloc(&s->rho1, box_make(-2, -2, -2, n[0] + 2, n[1] + 2, n[2] + 2), 1);
    mesh_alloc(&s->rho2, box_make(-2, -2, -2, n[0] + 2, n[1] + 2, n[2] + 2), 1);
    mesh_zero(&s->J);
    s->vsp = 1.0;
    return s;
}

/* VSP_configure, source/core/plasma_VSP.c:18-27. */
int opic_set_vsp(opic_t *s, double alpha)
{
    if (!(alpha > -0.1 && alpha <= 1.1)) return -1;
    s->vsp = alpha;
    return 0;
}

/* VSP_vectorSweep / VSP_scalarSweep, source/core/plasma_VSP.c:32-176: the
 * (beta, alpha, beta) convolution, beta = (1 - alpha)/2, along z, then y,
 * then x, over every line of the mesh (nodes [-2, n+2]); each new value
 * takes the old values of its neighbours (the sweep keeps the untouched
 * previous value), and the end nodes take their one inner neighbour.  A
 * no-op for |alpha - 1| < 1e-10. */
static void vsp_sweep(const opic_t *s, mesh_t *m)
{
    if (fabs(s->vsp - 1.0) < 1e-10) return;
    const double alpha = s->vsp, beta = 0.5 * (1.0 - alpha);
    double un[3];
    for (int ax = 2; ax >= 0; --ax) {
        const int p = ax == 0 ? 1 : 0, q = ax == 2 ? 1 : 2;
        const int lo = -2, hi = s->n[ax] + 2;
        int r[3];
        for (int a = m->b.lo[p]; a <= m->b.hi[p]; ++a)
            for (int b = m->b.lo[q]; b <= m->b.hi[q]; ++b) {
                r[p] = a;
                r[q] = b;
                r[ax] = lo;
                double *f0 = ATv(m, r);
                r[ax] = lo + 1;
                double *f1 = ATv(m, r);
                for (int c = 0; c < m->nc; ++c) {
                    un[c] = f0[c];
                    f0[c] = beta * f1[c] + alpha * f0[c];
                }
                for (int k = lo + 1; k <= hi - 1; ++k) {
                    r[ax] = k;
                    double *f = ATv(m, r);
                    r[ax] = k + 1;
                    double *fn = ATv(m, r);
                    for (int c = 0; c < m->nc; ++c) {
                        const double v = alpha * (f[c]) + beta * (un[c] + fn[c]);
                        un[c] = f[c];
                        f[c] = v;
                    }
                }
                r[ax] = hi;
                double *fe = ATv(m, r);
                for (int c = 0; c < m->nc; ++c) fe[c] = beta * un[c] + alpha * fe[c];
            }
    }
}

void opic_destroy(opic_t *s)
{
    if (!s) return;
    mesh_free(&s->E); mesh_free(&s->H); mesh_free(&s->J);
    mesh_free(&s->rho1); mesh_free(&s->rho2);
    for (int b = 0; b < 6; ++b) {
        if (s->capE[b].m.d) mesh_free(&s->capE[b].m);
        if (s->capH[b].m.d) mesh_free(&s->capH[b].m);
        for (int q = 0; q < 6; ++q) free(s->capE[b].mur[q].buf);
    }
    for (int l = 0; l < s->nlaser; ++l) {
        free(s->laser[l].mapE);
        free(s->laser[l].mapH);
    }
    for (int f = 0; f < 12; ++f) {
        free(s->tfsf.E[f].data);
        free(s->tfsf.H[f].data);
    }
    free(s->pl);
    free(s);
}

double opic_time(const opic_t *s) { return s->time; }

void opic_set_fields(opic_t *s, const double *E, const double *H)
{
    long c = 0;
    for (int i = 0; i <= s->n[0]; ++i)
        for (int j = 0; j <= s->n[1]; ++j)
            for (int k = 0; k <= s->n[2]; ++k, c += 3) {
                memcpy(AT(&s->E, i, j, k), E + c, 3 * sizeof(double));
                memcpy(AT(&s->H, i, j, k), H + c, 3 * sizeof(double));
            }
}

void opic_get_fields(const opic_t *s, double *E, double *H)
{
    long c = 0;
    for (int i = 0; i <= s->n[0]; ++i)
        for (int j = 0; j <= s->n[1]; ++j)
            for (int k = 0; k <= s->n[2]; ++k, c += 3) {
                if (E) memcpy(E + c, AT(&s->E, i, j, k), 3 * sizeof(double));
                if (H) memcpy(H + c, AT(&s->H, i, j, k), 3 * sizeof(double));
            }
}

void opic_get_J(const opic_t *s, double *J)
{
    long n = s->J.si * (s->J.b.hi[0] - s->J.b.lo[0] + 1) * 3;
    memcpy(J, s->J.d, sizeof(double) * (size_t)n);
}

/* Loading: plasma_load with one rank, source/core/plasma_IO.c:104-160,
 * 227-267.  Each chunk of up to 2^18 markers passes through _3_sort, whose
 * swap-from-tail loop leaves a chunk [m0 m1 .. mN-1] as [m1 .. mN-1 m0]
 * when every marker belongs to the single host; the chunk is then appended
 * to the shell.  All markers start in the shell; the core is empty. */
int opic_set_markers(opic_t *s, const void *rec, long n, int recsize)
{
    const long CHUNK = 1L << 18;
    long cap = n + n / 8 + 1024;
    free(s->pl);
    s->pl = (mk_t *)calloc((size_t)cap, sizeof(mk_t));
    if (!s->pl) return -1;
    const char *src = (const char *)rec;
    long out = 0;
    for (long c0 = 0; c0 < n; c0 += CHUNK) {
        long m = (n - c0 < CHUNK) ? n - c0 : CHUNK;
        for (long q = 0; q < m; ++q) {
            long from = c0 + ((q + 1) % m);
            mk_t *d = s->pl + out++;
            memset(d, 0, sizeof(*d));
            memcpy(d, src + from * recsize, (size_t)(recsize < 60 ? recsize : 60));
        }
    }
    s->nAll = cap;
    s->nShell = n;
    s->nCore = cap;
    return 0;
}

long opic_marker_count(const opic_t *s)
{
    return s->nShell + (s->nAll - s->nCore);
}

/* plasma_save order, source/core/plasma_IO.c:273-287. */
void opic_get_markers(const opic_t *s, void *rec, int recsize)
{
    char *dst = (char *)rec;
    long o = 0;
    int nb = recsize < 60 ? recsize : 60;
    for (long q = 0; q < s->nShell; ++q, ++o) {
        memset(dst + o * recsize, 0, (size_t)recsize);
        memcpy(dst + o * recsize, s->pl + q, (size_t)nb);
    }
    for (long q = s->nCore; q < s->nAll; ++q, ++o) {
        memset(dst + o * recsize, 0, (size_t)recsize);
        memcpy(dst + o * recsize, s->pl + q, (size_t)nb);
    }
}

int opic_add_laser(opic_t *s, double A, double T0, double delay, double dt,
                   double phi0, int forward, int plane, const double *map)
{
    if (s->nlaser >= 4) return -1;
    laser_t *l = s->laser + s->nlaser++;
    long nodes = (long)(s->n[1] + 1 - 10) * (s->n[2] + 1 - 10);
    l->A = A; l->T0 = T0; l->delay = delay; l->dt = dt; l->phi0 = phi0;
    l->forward = forward; l->plane = plane;
    l->mapE = (double *)malloc(sizeof(double) * 4 * (size_t)nodes);
    l->mapH = (double *)malloc(sizeof(double) * 4 * (size_t)nodes);
    memcpy(l->mapE, map, sizeof(double) * 4 * (size_t)nodes);
    memcpy(l->mapH, map + 4 * nodes, sizeof(double) * 4 * (size_t)nodes);
    return 0;
}

/* ================================================================== */
/* Frame helpers: ea = normal axis, ep/eq = the two others in cyclic
 * order (capAxisFrameFinder, source/core/em_caps.c:66). */
static inline int EP(int ea) { return (ea + 1) % 3; }
static inline int EQ(int ea) { return (ea + 2) % 3; }

/* ---- periodic field BC, source/core/emCap_periodic.c:12-68 ---- */
static void bc_periodic(const opic_t *s, mesh_t *mn, mesh_t *mx, int ea,
                        int isE)
{
    int ep = EP(ea), eq = EQ(ea);
    int r0[3], r1[3];
    for (int p = mn->b.lo[ep]; p <= mn->b.hi[ep]; ++p)
        for (int q = mn->b.lo[eq]; q <= mn->b.hi[eq]; ++q) {
            r0[ep] = p; r0[eq] = q;
            r1[ep] = p; r1[eq] = q;
            if (!isE) {
                /* minH(0) = maxH(N); maxH(N+1) = minH(1) */
                r0[ea] = 0; r1[ea] = s->n[ea];
                memcpy(ATv(mn, r0), ATv(mx, r1), 3 * sizeof(double));
                r0[ea] = 1; r1[ea] = s->n[ea] + 1;
                memcpy(ATv(mx, r1), ATv(mn, r0), 3 * sizeof(double));
            } else {
                /* maxE(N) = minE(0); maxE(N+1) = minE(1) */
                r0[ea] = 0; r1[ea] = s->n[ea];
                memcpy(ATv(mx, r1), ATv(mn, r0), 3 * sizeof(double));
                r0[ea] = 1; r1[ea] = s->n[ea] + 1;
                memcpy(ATv(mx, r1), ATv(mn, r0), 3 * sizeof(double));
            }
        }
}

/* ---- mirror H / E and Mur H, source/core/emCap_mirror.c:17-72,
 *      source/core/emCap_Mur.c:55-78 ---- */
static void bc_wall_H(const opic_t *s, mesh_t *m, int ea, int top, int zeroN)
{
    int ep = EP(ea), eq = EQ(ea);
    int pl = top ? s->n[ea] : 0;
    int ri[3], rx[3];
    for (int p = m->b.lo[ep]; p <= m->b.hi[ep]; ++p)
        for (int q = m->b.lo[eq]; q <= m->b.hi[eq]; ++q) {
            ri[ep] = rx[ep] = p;
            ri[eq] = rx[eq] = q;
            if (zeroN) {
                ri[ea] = pl;
                ATv(m, ri)[ea] = 0;
            }
            ri[ea] = pl + (1 - top);   /* intl  */
            rx[ea] = pl + top;         /* extnl */
            double *I = ATv(m, ri), *X = ATv(m, rx);
            X[ep] = I[ep];
            X[eq] = I[eq];
        }
}

static void bc_mirror_E(const opic_t *s, mesh_t *m, int ea, int top)
{
    int ep = EP(ea), eq = EQ(ea);
    int pl = top ? s->n[ea] : 0;
    int r[3], ri[3], rx[3];
    for (int p = m->b.lo[ep]; p <= m->b.hi[ep]; ++p)
        for (int q = m->b.lo[eq]; q <= m->b.hi[eq]; ++q) {
            r[ep] = ri[ep] = rx[ep] = p;
            r[eq] = ri[eq] = rx[eq] = q;
            r[ea] = pl;
            ATv(m, r)[ep] = 0;
            ATv(m, r)[eq] = 0;
            ri[ea] = pl + (1 - top);
            rx[ea] = pl + top;
            ATv(m, rx)[ea] = ATv(m, ri)[ea];
        }
}

/* ---- Mur E, source/core/emCap_Mur.c:92-175.  Buffer record per face node:
 *      {Ep0, Eq0, Ep1, Eq1} = old tangential E on the plane / one inside. */
static void mur_init(opic_t *s, mur_t *pk, const mesh_t *capm, int ea, int top)
{
    int ep = EP(ea), eq = EQ(ea);
    int pl = top ? s->n[ea] : 0;
    int sin = 1 - 2 * top;
    long np = capm->b.hi[ep] - capm->b.lo[ep] + 1;
    long nq = capm->b.hi[eq] - capm->b.lo[eq] + 1;
    pk->buf = (double *)malloc(sizeof(double) * 4 * (size_t)(np * nq));
    pk->alpha = (s->tau - s->h[ea]) / (s->tau + s->h[ea]);
    pk->ea = ea;
    pk->top = top;
    pk->on = 1;
    long o = 0;
    int r[3], rs[3], ri[3], rx[3];
    for (int p = capm->b.lo[ep]; p <= capm->b.hi[ep]; ++p)
        for (int q = capm->b.lo[eq]; q <= capm->b.hi[eq]; ++q, o += 4) {
            r[ep] = rs[ep] = ri[ep] = rx[ep] = p;
            r[eq] = rs[eq] = ri[eq] = rx[eq] = q;
            r[ea] = pl;
            rs[ea] = pl + sin;
            /* reads/writes the main E mesh (not the cap) */
            pk->buf[o + 0] = ATv(&s->E, r)[ep];
            pk->buf[o + 1] = ATv(&s->E, r)[eq];
            pk->buf[o + 2] = ATv(&s->E, rs)[ep];
            pk->buf[o + 3] = ATv(&s->E, rs)[eq];
            ri[ea] = pl + (1 - top);
            rx[ea] = pl + top;
            ATv(&s->E, rx)[ea] = ATv(&s->E, ri)[ea];
        }
}

static void mur_E(const opic_t *s, mur_t *pk, mesh_t *m)
{
    int ea = pk->ea, top = pk->top, ep = EP(ea), eq = EQ(ea);
    int pl = top ? s->n[ea] : 0;
    int sin = 1 - 2 * top;
    double alpha = pk->alpha, *buf = pk->buf;
    long o = 0;
    int r[3], rs[3], ri[3], rx[3];
    for (int p = m->b.lo[ep]; p <= m->b.hi[ep]; ++p)
        for (int q = m->b.lo[eq]; q <= m->b.hi[eq]; ++q, o += 4) {
            r[ep] = rs[ep] = ri[ep] = rx[ep] = p;
            r[eq] = rs[eq] = ri[eq] = rx[eq] = q;
            r[ea] = pl;
            rs[ea] = pl + sin;
            ri[ea] = pl + (1 - top);
            rx[ea] = pl + top;
            ATv(m, rx)[ea] = ATv(m, ri)[ea];
            double *P = ATv(m, r), *S = ATv(m, rs);
            P[ep] = buf[o + 2] + alpha * (S[ep] - buf[o + 0]);
            P[eq] = buf[o + 3] + alpha * (S[eq] - buf[o + 1]);
            buf[o + 0] = P[ep];
            buf[o + 1] = P[eq];
            buf[o + 2] = S[ep];
            buf[o + 3] = S[eq];
        }
}

/* ---- cap tables, source/core/em_caps.c:68-113; index [bc][top][lo/hi] ---- */
static const int capUpdateE[3][2][2] = {{{0, 1}, {-1, -1}}, {{1, 1}, {-1, 0}}, {{1, 1}, {-1, 0}}};
static const int capFlushE[3][2][2]  = {{{0, 1}, {-1, +1}}, {{0, 1}, {-1, +1}}, {{0, 1}, {-1, +1}}};
static const int capUpdateH[3][2][2] = {{{0, 2}, {-1, 0}}, {{1, 2}, {-1, 0}}, {{0, 2}, {-1, 0}}};
static const int capFlushH[3][2][2]  = {{{0, 2}, {-1, +1}}, {{0, 2}, {-1, +1}}, {{0, 2}, {-1, +1}}};

/* cap_1_initCaps, source/core/em_caps.c:140-186 */
static void caps_init(opic_t *s, const int (*upd)[2][2], const int (*fl)[2][2],
                      cap_t *caps)
{
    const int *bmin = s->bc, *bmax = s->bc + 3;
    int side[2][3] = {{0, 0, 0}, {s->n[0], s->n[1], s->n[2]}};
    for (int b = 0; b < 6; ++b) {
        int ea = b >> 1, ep = EP(ea), eq = EQ(ea), top = b & 1;
        int bct = top ? bmax[ea] : bmin[ea];
        cap_t *c = caps + b;
        c->fl.lo[ea] = side[top][ea] + fl[bct][top][0];
        c->fl.hi[ea] = side[top][ea] + fl[bct][top][1];
        c->upd.lo[ea] = side[top][ea] + upd[bct][top][0];
        c->upd.hi[ea] = side[top][ea] + upd[bct][top][1];
        int ax[2] = {ep, eq};
        for (int t = 0; t < 2; ++t) {
            int a = ax[t];
            c->fl.lo[a] = side[0][a] + fl[bmin[a]][0][0];
            c->fl.hi[a] = side[1][a] + fl[bmax[a]][1][1];
            c->upd.lo[a] = side[0][a] + upd[bmin[a]][0][0];
            c->upd.hi[a] = side[1][a] + upd[bmax[a]][1][1];
        }
        mesh_alloc(&c->m, c->fl, 3);
    }
}

/* cap_2_shrink, source/core/em_caps.c:191-259 */
static void caps_shrink(cap_t *caps)
{
    for (int b = 2; b < 6; ++b) {
        cap_t *c = caps + b;
        c->ncopy = 0;
        int ea = b >> 1, ax[2] = {EP(ea), EQ(ea)};
        for (int rd = 0; rd < b; ++rd) {
            if ((rd >> 1) == (b >> 1)) continue;
            const cap_t *k = caps + rd;
            box3 cm = c->upd;
            for (int a = 0; a < 3; ++a) {
                if (cm.lo[a] < k->upd.lo[a]) cm.lo[a] = k->upd.lo[a];
                if (cm.hi[a] > k->upd.hi[a]) cm.hi[a] = k->upd.hi[a];
            }
            c->copy[c->ncopy] = cm;
            c->copy_src[c->ncopy++] = rd;
            for (int t = 0; t < 2; ++t) {
                int a = ax[t];
                if (c->upd.hi[a] == k->upd.hi[a] && c->upd.lo[a] < k->upd.lo[a])
                    c->upd.hi[a] = k->upd.lo[a] - 1;
                if (c->upd.lo[a] == k->upd.lo[a] && c->upd.hi[a] > k->upd.hi[a])
                    c->upd.lo[a] = k->upd.hi[a] + 1;
            }
            for (int t = 0; t < 2; ++t) {
                int a = ax[t];
                if (c->fl.hi[a] == k->fl.hi[a] && c->fl.lo[a] < k->fl.lo[a])
                    c->fl.hi[a] = k->fl.lo[a] - 1;
                if (c->fl.lo[a] == k->fl.lo[a] && c->fl.hi[a] > k->fl.hi[a])
                    c->fl.lo[a] = k->fl.hi[a] + 1;
            }
        }
    }
}

/* cap_resetGhosts, source/core/em_caps.c:266-308 (main meshes). */
static void reset_ghosts(opic_t *s)
{
    for (int ax = 0; ax < 3; ++ax) {
        if (s->bc[ax] == OPIC_BC_PERIODIC) {
            bc_periodic(s, &s->H, &s->H, ax, 0);
            bc_periodic(s, &s->E, &s->E, ax, 1);
        } else if (s->bc[ax] == OPIC_BC_MIRROR) {
            bc_wall_H(s, &s->H, ax, 0, 1);
            bc_mirror_E(s, &s->E, ax, 0);
        }
        if (s->bc[3 + ax] == OPIC_BC_MIRROR) {
            bc_wall_H(s, &s->H, ax, 1, 1);
            bc_mirror_E(s, &s->E, ax, 1);
        }
    }
}

/* Laser soft source add_source, source/scpic/em_mirror.c:667-693. */
#define LOFFSET 5
static void laser_add(const opic_t *s, mesh_t *m, const double *src, int pos,
                      double ampl, double phase, const box3 *tu)
{
    if (pos < tu->lo[0] || pos > tu->hi[0]) return;
    int JMIN = (int)(fmax(0 - 1, LOFFSET) + 0.1);
    int JMAX = (int)(fmin(s->n[1] + 1, s->n[1] - LOFFSET) + 0.1);
    int KMIN = (int)(fmax(0 - 1, LOFFSET) + 0.1);
    int KMAX = (int)(fmin(s->n[2] + 1, s->n[2] - LOFFSET) + 0.1);
    int jmin = (int)(fmax(tu->lo[1], JMIN) + 0.1);
    int jmax = (int)(fmin(tu->hi[1], JMAX) + 0.1);
    int kmin = (int)(fmax(tu->lo[2], KMIN) + 0.1);
    int kmax = (int)(fmin(tu->hi[2], KMAX) + 0.1);
    double cs = cos(phase), sn = sin(phase), A = ampl * s->tau / s->h[0];
    for (int j = jmin; j <= jmax; j++)
        for (int k = kmin; k <= kmax; k++) {
            long cnt = (long)(j - JMIN) * (KMAX - KMIN + 1) + k - KMIN;
            const double *w = src + 4 * cnt;   /* y_cos y_sin z_cos z_sin */
            double *f = AT(m, pos, j, k);
            f[1] += A * (w[2] * cs + w[3] * sn);
            f[2] += A * (w[0] * cs + w[1] * sn);
        }
}

static void add_H_sources(const opic_t *s, mesh_t *H, double time, const box3 *tu)
{
    for (int l = 0; l < s->nlaser; ++l) {
        const laser_t *L = s->laser + l;
        double t = time - L->T0 + L->delay;
        double ampl = exp(-pow(2 * t / L->dt, 2));
        int I = L->forward ? L->plane : s->n[0] - L->plane + 1;
        laser_add(s, H, L->mapE, I, ampl, 2 * M_PI * t + L->phi0, tu);
    }
}

static void add_E_sources(const opic_t *s, mesh_t *E, double time, const box3 *tu)
{
    for (int l = 0; l < s->nlaser; ++l) {
        const laser_t *L = s->laser + l;
        double t = time + 0.5 * s->tau - L->T0 + L->delay;
        double ampl = exp(-pow(2 * t / L->dt, 2));
        int I = L->forward ? L->plane - 1 : s->n[0] - L->plane + 1;
        laser_add(s, E, L->mapH, I, ampl, 2 * M_PI * t + L->phi0, tu);
    }
}

/* Applies the local field BCs to all caps in the reference order,
 * source/core/em_caps.c:470-553 (H) and :631-721 (E). */
static void caps_apply_bc(opic_t *s, cap_t *caps, int isE)
{
    for (int ea = 0; ea < 3; ++ea) {
        int ep = EP(ea), eq = EQ(ea);
        int list[5] = {(ea << 1) | 0, (ep << 1) | 0, (ep << 1) | 1,
                       (eq << 1) | 0, (eq << 1) | 1};
        for (int top = 0; top < 2; ++top) {
            int bcv = s->bc[3 * top + ea];
            if (bcv == OPIC_BC_PERIODIC && top == 0) {
                bc_periodic(s, &caps[(ea << 1) | 0].m, &caps[(ea << 1) | 1].m, ea, isE);
                for (int t = 1; t < 5; ++t)
                    bc_periodic(s, &caps[list[t]].m, &caps[list[t]].m, ea, isE);
            } else if (bcv == OPIC_BC_MIRROR) {
                int l2[5] = {(ea << 1) | top, list[1], list[2], list[3], list[4]};
                for (int t = 0; t < 5; ++t) {
                    if (isE) bc_mirror_E(s, &caps[l2[t]].m, ea, top);
                    else     bc_wall_H(s, &caps[l2[t]].m, ea, top, 1);
                }
            } else if (bcv == OPIC_BC_OPEN) {
                int l2[5] = {(ea << 1) | top, list[1], list[2], list[3], list[4]};
                int bound = (ea << 1) | top;
                for (int t = 0; t < 5; ++t) {
                    if (isE) mur_E(s, &caps[l2[t]].mur[bound], &caps[l2[t]].m);
                    else     bc_wall_H(s, &caps[l2[t]].m, ea, top, 0);
                }
            }
        }
    }
}

static void caps_copy(cap_t *caps)
{
    for (int b = 0; b < 6; ++b) {
        cap_t *c = caps + b;
        for (int t = 0; t < c->ncopy; ++t) {
            const box3 *r = c->copy + t;
            const mesh_t *src = &caps[c->copy_src[t]].m;
            for (int i = r->lo[0]; i <= r->hi[0]; ++i)
                for (int j = r->lo[1]; j <= r->hi[1]; ++j)
                    for (int k = r->lo[2]; k <= r->hi[2]; ++k)
                        memcpy(AT(&c->m, i, j, k), AT(src, i, j, k), 3 * sizeof(double));
        }
    }
}

/* cap_cacheEH, source/core/em_caps.c:405-725. */
static void cap_cacheEH(opic_t *s)
{
    const double c1 = s->tau / s->h[0], c2 = s->tau / s->h[1], c3 = s->tau / s->h[2];
    const mesh_t *E = &s->E, *H = &s->H;
    for (int b = 0; b < 6; ++b) {
        mesh_t *cH = &s->capH[b].m;
        const box3 *r = &s->capH[b].upd;
        for (int i = r->lo[0]; i <= r->hi[0]; ++i)
            for (int j = r->lo[1]; j <= r->hi[1]; ++j)
                for (int k = r->lo[2]; k <= r->hi[2]; ++k) {
                    double *o = AT(cH, i, j, k);
                    const double *h = AT(H, i, j, k), *e = AT(E, i, j, k);
                    const double *em_i = AT(E, i - 1, j, k), *em_j = AT(E, i, j - 1, k),
                                 *em_k = AT(E, i, j, k - 1);
                    o[0] = h[0] + c3 * (e[1] - em_k[1]) - c2 * (e[2] - em_j[2]);
                    o[1] = h[1] + c1 * (e[2] - em_i[2]) - c3 * (e[0] - em_k[0]);
                    o[2] = h[2] + c2 * (e[0] - em_j[0]) - c1 * (e[1] - em_i[1]);
                }
    }
    for (int b = 0; b < 6; ++b)
        add_H_sources(s, &s->capH[b].m, s->time, &s->capH[b].fl);
    caps_copy(s->capH);
    caps_apply_bc(s, s->capH, 0);

    for (int b = 0; b < 6; ++b) {
        mesh_t *cE = &s->capE[b].m;
        const mesh_t *cH = &s->capH[b].m;
        const box3 *r = &s->capE[b].upd;
        for (int i = r->lo[0]; i <= r->hi[0]; ++i)
            for (int j = r->lo[1]; j <= r->hi[1]; ++j)
                for (int k = r->lo[2]; k <= r->hi[2]; ++k) {
                    double *o = AT(cE, i, j, k);
                    const double *e = AT(E, i, j, k), *h = AT(cH, i, j, k);
                    const double *hp_i = AT(cH, i + 1, j, k), *hp_j = AT(cH, i, j + 1, k),
                                 *hp_k = AT(cH, i, j, k + 1);
                    o[0] = e[0] - c3 * (hp_k[1] - h[1]) + c2 * (hp_j[2] - h[2]);
                    o[1] = e[1] - c1 * (hp_i[2] - h[2]) + c3 * (hp_k[0] - h[0]);
                    o[2] = e[2] - c2 * (hp_j[0] - h[0]) + c1 * (hp_i[1] - h[1]);
                }
    }
    caps_copy(s->capE);
    caps_apply_bc(s, s->capE, 1);
}

/* em_HSmallStep, source/core/em.c:54-77 */
static void em_HSmallStep(opic_t *s)
{
    const double c1 = 0.5 * s->tau / s->h[0], c2 = 0.5 * s->tau / s->h[1],
                 c3 = 0.5 * s->tau / s->h[2];
    for (int i = 3; i <= s->n[0] - 2; i++)
        for (int j = 3; j <= s->n[1] - 2; j++)
            for (int k = 3; k <= s->n[2] - 2; k++) {
                double *h = AT(&s->H, i, j, k);
                const double *e = AT(&s->E, i, j, k);
                const double *em_i = AT(&s->E, i - 1, j, k), *em_j = AT(&s->E, i, j - 1, k),
                             *em_k = AT(&s->E, i, j, k - 1);
                h[0] += c3 * (e[1] - em_k[1]) - c2 * (e[2] - em_j[2]);
                h[1] += c1 * (e[2] - em_i[2]) - c3 * (e[0] - em_k[0]);
                h[2] += c2 * (e[0] - em_j[0]) - c1 * (e[1] - em_i[1]);
            }
}

static void em_HHalfStep(opic_t *s)
{
    cap_cacheEH(s);
    em_HSmallStep(s);
    /* cap_flushHalfH, source/core/em_caps.c:730-749 */
    for (int b = 0; b < 6; ++b) {
        const box3 *r = &s->capH[b].fl;
        for (int i = r->lo[0]; i <= r->hi[0]; ++i)
            for (int j = r->lo[1]; j <= r->hi[1]; ++j)
                for (int k = r->lo[2]; k <= r->hi[2]; ++k) {
                    double *h = AT(&s->H, i, j, k);
                    const double *c = AT(&s->capH[b].m, i, j, k);
                    h[0] = 0.5 * (h[0] + c[0]);
                    h[1] = 0.5 * (h[1] + c[1]);
                    h[2] = 0.5 * (h[2] + c[2]);
                }
    }
}

static void flush_full(mesh_t *dst, cap_t *caps)
{
    for (int b = 0; b < 6; ++b) {
        const box3 *r = &caps[b].fl;
        for (int i = r->lo[0]; i <= r->hi[0]; ++i)
            for (int j = r->lo[1]; j <= r->hi[1]; ++j)
                for (int k = r->lo[2]; k <= r->hi[2]; ++k)
                    memcpy(AT(dst, i, j, k), AT(&caps[b].m, i, j, k), 3 * sizeof(double));
    }
}

/* TFSF_addSource / TFSF_subSource (source/core/em_TFSF.c:666-772): the
 * cached frame of every active face, in face order, added (sign +1) or
 * subtracted (sign -1) on the face's nodes inside the region tu. */
static void tfsf_apply(mesh_t *F, const tface_t *faces, long frame, double sign, const box3 *tu)
{
    for (int f = 0; f < 12; ++f) {
        const tface_t *fc = faces + f;
        if (!fc->data) continue;
        const int ea = fc->ea, ep = fc->ep, eq = fc->eq;
        if (fc->lo[ea] > tu->hi[ea] || fc->hi[ea] < tu->lo[ea]) continue;
        const int p1 = fc->lo[ep] < tu->lo[ep] ? tu->lo[ep] : fc->lo[ep];
        const int p2 = fc->hi[ep] > tu->hi[ep] ? tu->hi[ep] : fc->hi[ep];
        const int q1 = fc->lo[eq] < tu->lo[eq] ? tu->lo[eq] : fc->lo[eq];
        const int q2 = fc->hi[eq] > tu->hi[eq] ? tu->hi[eq] : fc->hi[eq];
        const long dQ = fc->hi[eq] - fc->lo[eq] + 1;
        const double *buf = fc->data + frame * fc->vol;
        int r[3];
        r[ea] = fc->lo[ea];
        for (int p = p1; p <= p2; ++p)
            for (int q = q1; q <= q2; ++q) {
                r[ep] = p;
                r[eq] = q;
                const double v = buf[(q - fc->lo[eq]) + (p - fc->lo[ep]) * dQ];
                double *e = AT(F, r[0], r[1], r[2]) + fc->comp;
                if (sign > 0)
                    *e += v;
                else
                    *e -= v;
            }
    }
}

/* The interface is live while the frame counter F satisfies F + 1 < last - 1:
 * at the start of every step cap_cacheEH calls TFSF_postEStep on the cap
 * meshes (source/core/em_caps.c:607), which switches the interface off once
 * F + 1 >= last - 1 (source/core/em_TFSF.c:838-841) before any source of that
 * step is applied. */
static int tfsf_live(const tfsf_t *t) { return t->on && t->frame + 1 < t->last - 1; }

/* TFSF_postHStep on the inner mesh (source/core/em_TFSF.c:812-825, called
 * from em_HStep, source/core/em.c:125): H faces play frame F (forward) or
 * last - 1 - F (backward), always added. */
static void tfsf_H(opic_t *s, const box3 *inner)
{
    const tfsf_t *t = &s->tfsf;
    if (!tfsf_live(t)) return;
    const long fr = t->mode == 0 ? t->frame : t->last - 1 - t->frame;
    tfsf_apply(&s->H, t->H, fr, 1.0, inner);
}

/* TFSF_postEStep on the inner mesh (source/core/em_TFSF.c:832-850, called
 * from em_EStep_finish, source/core/em.c:200): E faces play frame F + 1
 * (forward, added) or last - 1 - F (backward, subtracted). */
static void tfsf_E(opic_t *s, const box3 *inner)
{
    const tfsf_t *t = &s->tfsf;
    if (!tfsf_live(t)) return;
    if (t->mode == 0)
        tfsf_apply(&s->E, t->E, t->frame + 1, 1.0, inner);
    else
        tfsf_apply(&s->E, t->E, t->last - 1 - t->frame, -1.0, inner);
}

int opic_add_tfsf(opic_t *s, const int reg[6], int mode, int nframes,
                  const double *const dataE[12], const double *const dataH[12])
{
    /* TFSF_init, source/core/em_TFSF.c:481-545 (all axes active) */
    tfsf_t *t = &s->tfsf;
    if (t->on || (mode != 0 && mode != 2) || nframes < 1) return -1;
    const int side[2][3] = {{reg[0], reg[2], reg[4]}, {reg[1], reg[3], reg[5]}};
    for (int ea = 0; ea < 3; ++ea) {
        const int ep = (ea + 1) % 3, eq = (ea + 2) % 3;
        for (int top = 0; top < 2; ++top)
            for (int comp = 0; comp < 2; ++comp) {
                const int f = ea * 4 + top * 2 + comp;
                tface_t *fe = t->E + f, *fh = t->H + f;
                for (int a = 0; a < 3; ++a) {
                    fe->lo[a] = side[0][a];
                    fe->hi[a] = side[1][a];
                }
                fe->lo[ea] = fe->hi[ea] = side[top][ea];
                fe->comp = (ea + 1 + comp) % 3;
                fe->lo[fe->comp] += 1;               /* E nodes of that component */
                fe->ea = ea;
                fe->ep = ep;
                fe->eq = eq;
                *fh = *fe;
                fh->lo[ea] += top;                   /* H nodes of the face */
                fh->hi[ea] += top;
                fh->comp = 3 - fe->comp - ea;
                fe->vol = (long)(fe->hi[ep] - fe->lo[ep] + 1) * (fe->hi[eq] - fe->lo[eq] + 1);
                fh->vol = fe->vol;
                fe->data = fh->data = NULL;
                if (dataE && dataE[f]) {
                    fe->data = (double *)malloc(sizeof(double) * fe->vol * (size_t)nframes);
                    memcpy(fe->data, dataE[f], sizeof(double) * fe->vol * (size_t)nframes);
                }
                if (dataH && dataH[f]) {
                    fh->data = (double *)malloc(sizeof(double) * fh->vol * (size_t)nframes);
                    memcpy(fh->data, dataH[f], sizeof(double) * fh->vol * (size_t)nframes);
                }
            }
    }
    t->mode = mode;
    t->last = nframes;
    t->frame = (int)(s->time / s->tau + 0.1);        /* TFSF_turnPlay, :466 */
    t->on = 1;
    return 0;
}

static void em_HStep(opic_t *s)
{
    em_HSmallStep(s);
    flush_full(&s->H, s->capH);                     /* cap_flushH */
    box3 inner = box_make(3, 3, 3, s->n[0] - 2, s->n[1] - 2, s->n[2] - 2);
    tfsf_H(s, &inner);
    add_H_sources(s, &s->H, s->time, &inner);
}

/* em_EStep_start, source/core/em.c:134-161 */
static void em_EStep_start(opic_t *s)
{
    const double c1 = s->tau / s->h[0], c2 = s->tau / s->h[1], c3 = s->tau / s->h[2];
    for (int i = 2; i <= s->n[0] - 2; i++)
        for (int j = 2; j <= s->n[1] - 2; j++)
            for (int k = 2; k <= s->n[2] - 2; k++) {
                double *e = AT(&s->E, i, j, k);
                const double *h = AT(&s->H, i, j, k);
                const double *hp_i = AT(&s->H, i + 1, j, k), *hp_j = AT(&s->H, i, j + 1, k),
                             *hp_k = AT(&s->H, i, j, k + 1);
                e[0] -= c3 * (hp_k[1] - h[1]) - c2 * (hp_j[2] - h[2]);
                e[1] -= c1 * (hp_i[2] - h[2]) - c3 * (hp_k[0] - h[0]);
                e[2] -= c2 * (hp_j[0] - h[0]) - c1 * (hp_i[1] - h[1]);
            }
}

/* gaussSpot_forceE, source/core/em_gaussSpot.c:129-156: sets Ey, Ez on the
 * plane i over the whole E mesh to the Gaussian profile times the
 * oscillation and the time envelope (hard source). */
static void spot_force(opic_t *s, double t)
{
    const spot_t *g = &s->spot;
    if (!g->on || g->i < -1 || g->i > s->n[0] + 1) return;
    double tf = 1;
    if (t < g->front) {
        const double l = (t - g->front) / g->duration;
        tf = exp(-l * l);
    }
    if (t > g->front + g->plateau) {
        const double l = (t - g->front - g->plateau) / g->duration;
        tf = exp(-l * l);
    }
    for (int j = -1; j <= s->n[1] + 1; j++) {
        const double dy_ = ((j - 0.5) * s->h[1] - g->y) / g->width,
                     dy = (j * s->h[1] - g->y) / g->width;
        for (int k = -1; k <= s->n[2] + 1; k++) {
            const double dz = (k * s->h[2] - g->z) / g->width,
                         dz_ = ((k - 0.5) * s->h[2] - g->z) / g->width;
            double *e = AT(&s->E, g->i, j, k);
            e[1] = g->Ey * sin(g->omega * t) * tf * exp(-dy_ * dy_ - dz * dz);
            e[2] = g->Ez * cos(g->omega * t) * tf * exp(-dy * dy - dz_ * dz_);
        }
    }
}

int opic_add_gauss_spot(opic_t *s, const double p[11])
{
    /* gaussSpot_init, source/core/em_gaussSpot.c:42-120: the values of
     * .gaussSpot.cfg in file order */
    spot_t *g = &s->spot;
    g->ampl = p[0];
    g->omega = 2 * M_PI * p[1];
    g->x = p[2];
    g->y = p[3];
    g->z = p[4];
    g->width = p[5];
    g->duration = p[6];
    g->front = p[7];
    g->plateau = p[8];
    const int ey = (int)p[9], ez = (int)p[10];
    g->Ey = (ey > 0) - (ey < 0);
    g->Ez = (ez > 0) - (ez < 0);
    g->i = (int)(g->x / s->h[0] + 0.5);
    g->Ey *= g->ampl;
    g->Ez *= g->ampl;
    g->on = 1;
    return 0;
}

/* em_EStep_finish, source/core/em.c:188-221 */
static void em_EStep_finish(opic_t *s)
{
    flush_full(&s->E, s->capE);                     /* cap_flushE */
    box3 innerE = box_make(2, 2, 2, s->n[0] - 2, s->n[1] - 2, s->n[2] - 2);
    tfsf_E(s, &innerE);
    box3 dom = box_make(-1, -1, -1, s->n[0] + 1, s->n[1] + 1, s->n[2] + 1);
    add_E_sources(s, &s->E, s->time, &dom);
    for (int i = -1; i <= s->n[0] + 1; i++)
        for (int j = -1; j <= s->n[1] + 1; j++)
            for (int k = -1; k <= s->n[2] + 1; k++) {
                double *e = AT(&s->E, i, j, k);
                const double *jj = AT(&s->J, i, j, k);
                e[0] -= s->tau * jj[0];
                e[1] -= s->tau * jj[1];
                e[2] -= s->tau * jj[2];
            }
    spot_force(s, s->time + s->tau);                /* source/core/em.c:216 */
    if (s->tfsf.on) ++s->tfsf.frame;                 /* TFSF_completeFrame, :857-860 */
}

/* em_energy, source/core/em.c:226-252 */
static void em_energy(const opic_t *s, double *WE, double *WM)
{
    double we = 0, wb = 0;
    for (int i = 0; i < s->n[0]; ++i)
        for (int j = 0; j < s->n[1]; ++j)
            for (int k = 0; k < s->n[2]; ++k) {
                double ex = AT(&s->E, i + 1, j, k)[0], ey = AT(&s->E, i, j + 1, k)[1],
                       ez = AT(&s->E, i, j, k + 1)[2], hx = AT(&s->H, i, j + 1, k + 1)[0],
                       hy = AT(&s->H, i + 1, j, k + 1)[1], hz = AT(&s->H, i + 1, j + 1, k)[2];
                we += ex * ex + ey * ey + ez * ez;
                wb += hx * hx + hy * hy + hz * hz;
            }
    *WE = 0.5 * we * s->h[0] * s->h[1] * s->h[2];
    *WM = 0.5 * wb * s->h[0] * s->h[1] * s->h[2];
}

/* ================================================================== */
/* Particles. */
#define DBL_TO_INT(x) ((int)((x) + 128.0) - 128)   /* source/misc_PIC.h:18 */

/* Trilinear gather in the association of mf_PIC_general_read,
 * source/misc_PIC.h:69-126. */
static inline double gather(const mesh_t *m, int c, int i, double sx, int j,
                            double sy, int k, double sz)
{
    return AT(m, i, j, k)[c] * (1 - sx) * ((1 - sy) * (1 - sz)) +
           AT(m, i + 1, j, k)[c] * sx * ((1 - sy) * (1 - sz)) +
           AT(m, i, j + 1, k)[c] * (1 - sx) * (sy * (1 - sz)) +
           AT(m, i + 1, j + 1, k)[c] * sx * (sy * (1 - sz)) +
           AT(m, i, j, k + 1)[c] * (1 - sx) * ((1 - sy) * sz) +
           AT(m, i + 1, j, k + 1)[c] * sx * ((1 - sy) * sz) +
           AT(m, i, j + 1, k + 1)[c] * (1 - sx) * (sy * sz) +
           AT(m, i + 1, j + 1, k + 1)[c] * sx * (sy * sz);
}

/* plasmaCurrents_kernel, source/core/plasma_currentKernel.c:20-124. */
static void deposit_seg(opic_t *s, int i, int j, int k, double x1, double x2,
                        double y1, double y2, double z1, double z2, double rp)
{
    const double h1 = s->h[0], h2 = s->h[1], h3 = s->h[2], tau = s->tau;
    const double dX = x2 - x1, dY = y2 - y1, dZ = z2 - z1;
    const double tX = 0.5 * (x1 + x2) - i * h1;
    const double tY = 0.5 * (y1 + y2) - j * h2;
    const double tZ = 0.5 * (z1 + z2) - k * h3;
    mesh_t *J = &s->J;
    const double yz = dY * dZ / 12, xz = dX * dZ / 12, xy = dX * dY / 12;
    const double ax = rp * dX / (tau * h2 * h3);
    AT(J, i + 1, j, k)[0] += ax * ((h2 - tY) * (h3 - tZ) + yz);
    AT(J, i + 1, j + 1, k)[0] += ax * (tY * (h3 - tZ) - yz);
    AT(J, i + 1, j, k + 1)[0] += ax * ((h2 - tY) * tZ - yz);
    AT(J, i + 1, j + 1, k + 1)[0] += ax * (tY * tZ + yz);
    const double ay = rp * dY / (tau * h1 * h3);
    AT(J, i, j + 1, k)[1] += ay * ((h1 - tX) * (h3 - tZ) + xz);
    AT(J, i + 1, j + 1, k)[1] += ay * (tX * (h3 - tZ) - xz);
    AT(J, i, j + 1, k + 1)[1] += ay * ((h1 - tX) * tZ - xz);
    AT(J, i + 1, j + 1, k + 1)[1] += ay * (tX * tZ + xz);
    const double az = rp * dZ / (tau * h1 * h2);
    AT(J, i, j, k + 1)[2] += az * ((h1 - tX) * (h2 - tY) + xy);
    AT(J, i + 1, j, k + 1)[2] += az * (tX * (h2 - tY) - xy);
    AT(J, i, j + 1, k + 1)[2] += az * ((h1 - tX) * tY - xy);
    AT(J, i + 1, j + 1, k + 1)[2] += az * (tX * tY + xy);
}

/* plasmaCurrents_multiCellStep, source/core/plasma_currentKernel.c:156-279. */
static void deposit_multi(opic_t *s, int i1, int j1, int k1, int i2, int j2,
                          int k2, double x1, double y1, double z1, double x2,
                          double y2, double z2, double rp)
{
    const double h1 = s->h[0], h2 = s->h[1], h3 = s->h[2];
    const int di = abs(i2 - i1), dj = abs(j2 - j1), dk = abs(k2 - k1);
    if (di + dj + dk == 2) {
        int i = i1, j = j1, k = k1;
        const double sgX = di * ((i1 + i2 + 1) * 0.5 * h1 - x1) / (x2 - x1 + ((1 - di) << 16));
        const double sgY = dj * ((j1 + j2 + 1) * 0.5 * h2 - y1) / (y2 - y1 + ((1 - dj) << 16));
        const double sgZ = dk * ((k1 + k2 + 1) * 0.5 * h3 - z1) / (z2 - z1 + ((1 - dk) << 16));
        double k1s = sgX + (1 - di) * sgY;
        double k2s = sgZ + (1 - dk) * sgY;
        int dI[2][3] = {{0, 0, 0}, {0, 0, 0}};
        int st = 0;
        dI[0][0] = di * (i2 - i1);
        dI[0][1] = (1 - di) * (j2 - j1);
        dI[1][1] = (1 - dk) * (j2 - j1);
        dI[1][2] = dk * (k2 - k1);
        if (k1s > k2s) {
            double t = k1s; k1s = k2s; k2s = t;
            st = 1;
        }
        const double xA = x1 + k1s * (x2 - x1), yA = y1 + k1s * (y2 - y1), zA = z1 + k1s * (z2 - z1);
        const double xB = x1 + k2s * (x2 - x1), yB = y1 + k2s * (y2 - y1), zB = z1 + k2s * (z2 - z1);
        deposit_seg(s, i, j, k, x1, xA, y1, yA, z1, zA, rp);
        i += dI[st][0]; j += dI[st][1]; k += dI[st][2];
        deposit_seg(s, i, j, k, xA, xB, yA, yB, zA, zB, rp);
        i += dI[1 - st][0]; j += dI[1 - st][1]; k += dI[1 - st][2];
        deposit_seg(s, i, j, k, xB, x2, yB, y2, zB, z2, rp);
        return;
    }
    const double xC = (i1 + i2 + 1) * 0.5 * h1, yC = (j1 + j2 + 1) * 0.5 * h2,
                 zC = (k1 + k2 + 1) * 0.5 * h3;
    double sg1 = (xC - x1) / (x2 - x1), sg2 = (yC - y1) / (y2 - y1), sg3 = (zC - z1) / (z2 - z1);
    double rc[3][3];
    int dI[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    int s1 = 0, s2 = 1, s3 = 2, i = i1, j = j1, k = k1;
    rc[0][0] = xC; rc[0][1] = y1 + sg1 * (y2 - y1); rc[0][2] = z1 + sg1 * (z2 - z1);
    dI[0][0] = i2 - i1;
    rc[1][0] = x1 + sg2 * (x2 - x1); rc[1][1] = yC; rc[1][2] = z1 + sg2 * (z2 - z1);
    dI[1][1] = j2 - j1;
    rc[2][0] = x1 + sg3 * (x2 - x1); rc[2][1] = y1 + sg3 * (y2 - y1); rc[2][2] = zC;
    dI[2][2] = k2 - k1;
    if (sg1 > sg2) { int t = s1; s1 = s2; s2 = t; double d = sg1; sg1 = sg2; sg2 = d; }
    if (sg1 > sg3) { int t = s1; s1 = s3; s3 = t; double d = sg1; sg1 = sg3; sg3 = d; }
    if (sg2 > sg3) { int t = s2; s2 = s3; s3 = t; double d = sg2; sg2 = sg3; sg3 = d; }
    deposit_seg(s, i, j, k, x1, rc[s1][0], y1, rc[s1][1], z1, rc[s1][2], rp);
    i += dI[s1][0]; j += dI[s1][1]; k += dI[s1][2];
    deposit_seg(s, i, j, k, rc[s1][0], rc[s2][0], rc[s1][1], rc[s2][1], rc[s1][2], rc[s2][2], rp);
    i += dI[s2][0]; j += dI[s2][1]; k += dI[s2][2];
    deposit_seg(s, i, j, k, rc[s2][0], rc[s3][0], rc[s2][1], rc[s3][1], rc[s2][2], rc[s3][2], rp);
    deposit_seg(s, i2, j2, k2, rc[s3][0], x2, rc[s3][1], y2, rc[s3][2], z2, rp);
}

/* One marker: gather, Boris, deposit, move; source/core/plasma.c:115-238. */
static void push_one(opic_t *s, mk_t *p)
{
    const double h1 = s->h[0], h2 = s->h[1], h3 = s->h[2], tau = s->tau;
    int i, j, k, i_, j_, k_, i2, j2, k2, di, dj, dk;
    const double x1 = p->x, y1 = p->y, z1 = p->z;
    double sx, sy, sz, sx_, sy_, sz_;
    sx = x1 / h1; sx_ = sx + 0.5; i = DBL_TO_INT(sx); i_ = DBL_TO_INT(sx_); sx -= i; sx_ -= i_;
    sy = y1 / h2; sy_ = sy + 0.5; j = DBL_TO_INT(sy); j_ = DBL_TO_INT(sy_); sy -= j; sy_ -= j_;
    sz = z1 / h3; sz_ = sz + 0.5; k = DBL_TO_INT(sz); k_ = DBL_TO_INT(sz_); sz -= k; sz_ -= k_;

    const double Ex = gather(&s->E, 0, i_, sx_, j, sy, k, sz),
                 Ey = gather(&s->E, 1, i, sx, j_, sy_, k, sz),
                 Ez = gather(&s->E, 2, i, sx, j, sy, k_, sz_),
                 Hx = gather(&s->H, 0, i, sx, j_, sy_, k_, sz_),
                 Hy = gather(&s->H, 1, i_, sx_, j, sy, k_, sz_),
                 Hz = gather(&s->H, 2, i_, sx_, j_, sy_, k, sz);

    const double tauMass = 0.5 * tau * p->qDivM,
                 px = p->vx + tauMass * Ex, py = p->vy + tauMass * Ey, pz = p->vz + tauMass * Ez,
                 gamma = sqrt(1 + px * px + py * py + pz * pz);
    const double W = (gamma - 1.0) * p->rho / (px * px + py * py + pz * pz + 1e-19) * h1 * h2 * h3 / p->qDivM;
    s->Wx += W * px * px;
    s->Wy += W * py * py;
    s->Wz += W * pz * pz;

    const double tauGamma = tauMass / gamma;
    const double aHx = tauGamma * Hx, aHy = tauGamma * Hy, aHz = tauGamma * Hz;
    const double u = px + py * aHz - pz * aHy, v = py + pz * aHx - px * aHz,
                 w = pz + px * aHy - py * aHx;
    const double bxx = aHx * aHx, byy = aHy * aHy, bzz = aHz * aHz, bxy = aHx * aHy,
                 byz = aHy * aHz, bxz = aHz * aHx;
    const double delta = 1.0 / (1 + bxx + byy + bzz);
    p->vx = tauMass * Ex + ((1 + bxx) * u + (bxy + aHz) * v + (bxz - aHy) * w) * delta;
    p->vy = tauMass * Ey + ((bxy - aHz) * u + (1 + byy) * v + (byz + aHx) * w) * delta;
    p->vz = tauMass * Ez + ((bxz + aHy) * u + (byz - aHx) * v + (1 + bzz) * w) * delta;

    const double tg = tau / sqrt(1 + p->vx * p->vx + p->vy * p->vy + p->vz * p->vz);
    const double x2 = p->x + p->vx * tg, y2 = p->y + p->vy * tg, z2 = p->z + p->vz * tg;

    i2 = DBL_TO_INT(x2 / h1);
    j2 = DBL_TO_INT(y2 / h2);
    k2 = DBL_TO_INT(z2 / h3);
    di = abs(i2 - i);
    dj = abs(j2 - j);
    dk = abs(k2 - k);
    if (fabs(x2 - x1) < 1e-10 * h1) { di = 0; i2 = i; }
    if (fabs(y2 - y1) < 1e-10 * h2) { dj = 0; j2 = j; }
    if (fabs(z2 - z1) < 1e-10 * h3) { dk = 0; k2 = k; }

    const double rp = p->rho;
    if (di + dj + dk == 0) {
        deposit_seg(s, i, j, k, x1, x2, y1, y2, z1, z2, rp);
    } else if (di + dj + dk == 1) {
        const double alpha = dk * ((k + k2 + 1) * 0.5 * h3 - z1) / (z2 - z1 + ((1 - dk) << 16))
                           + dj * ((j + j2 + 1) * 0.5 * h2 - y1) / (y2 - y1 + ((1 - dj) << 16))
                           + di * ((i + i2 + 1) * 0.5 * h1 - x1) / (x2 - x1 + ((1 - di) << 16));
        const double xC = x1 + alpha * (x2 - x1), yC = y1 + alpha * (y2 - y1),
                     zC = z1 + alpha * (z2 - z1);
        deposit_seg(s, i, j, k, x1, xC, y1, yC, z1, zC, rp);
        deposit_seg(s, i2, j2, k2, xC, x2, yC, y2, zC, z2, rp);
    } else {
        deposit_multi(s, i, j, k, i2, j2, k2, x1, y1, z1, x2, y2, z2, rp);
    }
    p->x = x2;
    p->y = y2;
    p->z = z2;
}

static inline int in_core(const opic_t *s, const mk_t *p)
{
    return 1 * (p->x - s->core1[0]) * (s->core2[0] - p->x) >= 0 &&
           1 * (p->y - s->core1[1]) * (s->core2[1] - p->y) >= 0 &&
           1 * (p->z - s->core1[2]) * (s->core2[2] - p->z) >= 0;
}

/* plasma_timestep, source/core/plasma.c:94-263 (single rank: the
 * inbox test is empty). */
static void plasma_timestep(opic_t *s, int shell_stage, long N)
{
    mk_t *p = s->pl + (shell_stage ? 0 : s->nAll - N);
    const long FREQ = 2000;
    while (N > 0) {
        mk_t *sh = s->pl + s->nShell, *co = s->pl + s->nCore;
        long page = (N < FREQ || shell_stage) ? N : FREQ;
        N -= page;
        for (; page > 0; --page, ++p) {
            push_one(s, p);
            if (in_core(s, p)) {
                if (shell_stage) {
                    *(--co) = *p;
                    *p = *(--sh);
                    --p;
                }
            } else if (!shell_stage) {
                *(sh++) = *p;
                *p = *(co++);
            }
        }
        s->nShell = sh - s->pl;
        s->nCore = co - s->pl;
    }
}

/* Particle walls on the shell, source/core/plasma_walls.c:27-104, applied
 * in the order xMin yMin zMin xMax yMax zMax (plasma.c:286-287). */
static void plasma_walls(opic_t *s)
{
    for (int b = 0; b < 6; ++b) {
        int ax = b % 3, top = b / 3;
        int bcv = s->bc[b];
        double bnd = top ? s->shell2[ax] : s->shell1[ax];
        if (bcv == OPIC_BC_PERIODIC) {
            double shift = top ? -s->L[ax] : s->L[ax];
            for (mk_t *p = s->pl, *e = p + s->nShell; p < e; ++p) {
                double *x = &p->x + ax;
                if (top ? (*x >= bnd) : (*x <= bnd)) *x += shift;
            }
        } else if (bcv == OPIC_BC_MIRROR) {
            for (mk_t *p = s->pl, *e = p + s->nShell; p < e; ++p) {
                double *x = &p->x + ax, *v = &p->vx + ax;
                if (top ? (*x >= bnd) : (*x <= bnd)) {
                    *x = 2 * bnd - *x;
                    *v *= -1;
                }
            }
        } else if (bcv == OPIC_BC_OPEN) {
            double lim = top ? bnd - s->h[ax] : bnd + s->h[ax];
            mk_t *e = s->pl + s->nShell;
            for (mk_t *p = s->pl; p < e; ++p) {
                double x = (&p->x)[ax];
                if (top ? (x >= lim) : (x <= lim)) {
                    *p = *(--e);
                    --p;
                }
            }
            s->nShell = e - s->pl;
        }
    }
}

/* J boundary fold, source/core/plasma_currentWalls.c:132-227. */
static void jbc_fold(opic_t *s)
{
    static const int axes[6][3] = {{0, 1, 2}, {0, 1, 2}, {1, 0, 2}, {1, 0, 2}, {2, 0, 1}, {2, 0, 1}};
    const int BCv[6] = {s->bc[0], s->bc[3], s->bc[1], s->bc[4], s->bc[2], s->bc[5]};
    mesh_t *J = &s->J;
    for (int b = 0; b < 6; ++b) {
        int ea = axes[b][0], ep = axes[b][1], eq = axes[b][2], top = b & 1;
        int N = s->n[ea];
        int r[3];
        if (BCv[b] == OPIC_BC_PERIODIC) {
            if (top) continue;
            for (int p = -2; p <= s->n[ep] + 2; ++p)
                for (int q = -2; q <= s->n[eq] + 2; ++q)
                    for (int a = -2; a <= 2; ++a) {
                        int ra[3], rb[3];
                        ra[ep] = rb[ep] = p;
                        ra[eq] = rb[eq] = q;
                        ra[ea] = a;
                        rb[ea] = a + N;
                        double *lo = ATv(J, ra), *hi = ATv(J, rb);
                        hi[0] += lo[0];
                        hi[1] += lo[1];
                        hi[2] += lo[2];
                        memcpy(lo, hi, 3 * sizeof(double));
                    }
        } else if (BCv[b] == OPIC_BC_MIRROR || BCv[b] == OPIC_BC_OPEN) {
            int pl = top ? N : 0, so = top ? 1 : -1;
            int in_ = 1 - top, ex_ = top;
            for (int p = -2; p <= s->n[ep] + 2; ++p)
                for (int q = -2; q <= s->n[eq] + 2; ++q) {
                    r[ep] = p; r[eq] = q;
double *Jin = jat(J, r, ea, pl + in_), *Jex = jat(J, r, ea, pl + ex_);
                    double *Jm1 = jat(J, r, ea, pl - so), *Jp1 = jat(J, r, ea, pl + so);
                    double *Jm2 = jat(J, r, ea, pl - 2 * so), *Jp2 = jat(J, r, ea, pl + 2 * so);
                    double *Jin2 = jat(J, r, ea, pl + in_ - so), *Jex2 = jat(J, r, ea, pl + ex_ + so);
                    double *J0 = jat(J, r, ea, pl);
                    Jin[ea] -= Jex[ea];
                    Jm1[ep] += Jp1[ep];
                    Jm1[eq] += Jp1[eq];
                    Jin2[ea] -= Jex2[ea];
                    Jm2[ep] += Jp2[ep];
                    Jm2[eq] += Jp2[eq];
                    J0[ep] = 0;
                    J0[eq] = 0;
                    Jex[ea] = Jin[ea];
                    Jp1[ep] = -Jm1[ep];
                    Jp1[eq] = -Jm1[eq];
                    Jex2[ea] = Jin2[ea];
                    Jp2[ep] = -Jm2[ep];
                    Jp2[eq] = -Jm2[eq];
                }
        }
    }
    vsp_sweep(s, &s->J);                        /* jbc_finish, plasma_currentWalls.c:204 */
}

/* CIC charge density with boundary fold, source/core/plasma_rho.c:23-96,
 * 160-173, then the VSP sweep (:153). */
static void rho_compute(opic_t *s, mesh_t *rho)
{
    mesh_zero(rho);
    const double h1 = s->h[0], h2 = s->h[1], h3 = s->h[2];
    for (int part = 0; part < 2; ++part) {
        long a = part ? s->nCore : 0, e = part ? s->nAll : s->nShell;
        for (long q = a; q < e; ++q) {
            const mk_t *p = s->pl + q;
            int i, j, k;
            double sx, sy, sz;
            sx = p->x / h1; i = DBL_TO_INT(sx); sx -= i;
            sy = p->y / h2; j = DBL_TO_INT(sy); sy -= j;
            sz = p->z / h3; k = DBL_TO_INT(sz); sz -= k;
            double val = p->rho;
            /* mf_PIC_distribute association, source/misc_PIC.h:170-201 */
            *AT(rho, i, j, k) += (1 - sx) * ((1 - sy) * ((1 - sz) * val));
            *AT(rho, i + 1, j, k) += sx * ((1 - sy) * ((1 - sz) * val));
            *AT(rho, i, j + 1, k) += (1 - sx) * (sy * ((1 - sz) * val));
            *AT(rho, i + 1, j + 1, k) += sx * (sy * ((1 - sz) * val));
            *AT(rho, i, j, k + 1) += (1 - sx) * ((1 - sy) * (sz * val));
            *AT(rho, i + 1, j, k + 1) += sx * ((1 - sy) * (sz * val));
            *AT(rho, i, j + 1, k + 1) += (1 - sx) * (sy * (sz * val));
            *AT(rho, i + 1, j + 1, k + 1) += sx * (sy * (sz * val));
        }
    }
    static const int axes[6][3] = {{0, 1, 2}, {0, 1, 2}, {1, 0, 2}, {1, 0, 2}, {2, 0, 1}, {2, 0, 1}};
    const int BCv[6] = {s->bc[0], s->bc[3], s->bc[1], s->bc[4], s->bc[2], s->bc[5]};
    for (int b = 0; b < 6; ++b) {
        int ea = axes[b][0], ep = axes[b][1], eq = axes[b][2], top = b & 1;
        int N = s->n[ea];
        int r[3];
        if (BCv[b] == OPIC_BC_PERIODIC) {
            if (top) continue;
            for (int p = -2; p <= s->n[ep] + 2; ++p)
                for (int q = -2; q <= s->n[eq] + 2; ++q)
                    for (int a = -2; a <= 2; ++a) {
                        int ra[3], rb[3];
                        ra[ep] = rb[ep] = p; ra[eq] = rb[eq] = q;
                        ra[ea] = a; rb[ea] = a + N;
                        *ATv(rho, rb) += *ATv(rho, ra);
                        *ATv(rho, ra) = *ATv(rho, rb);
                    }
        } else {
            int pl = top ? N : 0, so = top ? 1 : -1;
            for (int p = -2; p <= s->n[ep] + 2; ++p)
                for (int q = -2; q <= s->n[eq] + 2; ++q) {
                    r[ep] = p; r[eq] = q;
double *Rm1 = jat(rho, r, ea, pl - so), *Rp1 = jat(rho, r, ea, pl + so);
                    double *Rm2 = jat(rho, r, ea, pl - 2 * so), *Rp2 = jat(rho, r, ea, pl + 2 * so);
                    *Rm1 += *Rp1;
                    *Rm2 += *Rp2;
                    *jat(rho, r, ea, pl) = 0;
                    *Rp1 = -*Rm1;
                    *Rp2 = -*Rm2;
                }
        }
    }
    vsp_sweep(s, rho);                          /* source/core/plasma_rho.c:153 */
}

void opic_rho(opic_t *s, double *rho)
{
    rho_compute(s, &s->rho2);
    long n = s->rho2.si * (s->rho2.b.hi[0] - s->rho2.b.lo[0] + 1);
    memcpy(rho, s->rho2.d, sizeof(double) * (size_t)n);
}

/* ================================================================== */
/* em_init + plasma_init, source/core/em.c:39-49, em_caps.c:334-397,
 * plasma.c:342-418. */
int opic_init(opic_t *s)
{
    if (!(1.0 / (s->tau * s->tau) > 1.0 / (s->h[0] * s->h[0]) + 1.0 / (s->h[1] * s->h[1]) +
                                         1.0 / (s->h[2] * s->h[2])))
        return -1;                                   /* "Yee-solver is unstable" */
    for (int a = 0; a < 3; ++a)
        if (s->n[a] < 7) return -2;                  /* "core region is too small" */
    reset_ghosts(s);
    caps_init(s, capUpdateE, capFlushE, s->capE);
    caps_init(s, capUpdateH, capFlushH, s->capH);
    caps_shrink(s->capE);
    caps_shrink(s->capH);
    for (int ea = 0; ea < 3; ++ea) {
        int ep = EP(ea), eq = EQ(ea);
        for (int top = 0; top < 2; ++top) {
            if (s->bc[3 * top + ea] != OPIC_BC_OPEN) continue;
            int bound = (ea << 1) | top;
            int list[5] = {bound, (ep << 1) | 0, (ep << 1) | 1, (eq << 1) | 0, (eq << 1) | 1};
            for (int t = 0; t < 5; ++t)
                mur_init(s, &s->capE[list[t]].mur[bound], &s->capE[list[t]].m, ea, top);
        }
    }
    for (int a = 0; a < 3; ++a) {
        s->core1[a] = (0 + 3) * s->h[a] * 1;
        s->core2[a] = (s->n[a] - 3) * s->h[a] * 1;
        s->shell1[a] = 0 * s->h[a] * 1;
        s->shell2[a] = s->n[a] * s->h[a] * 1;
    }
    plasma_walls(s);
    s->inited = 1;
    return 0;
}

static void gauss_residuals(opic_t *s, double *res)
{
    double d0 = -1, d1 = -1;
    const double h1 = s->h[0], h2 = s->h[1], h3 = s->h[2];
    for (int i = 0; i <= s->n[0]; ++i)
        for (int j = 0; j <= s->n[1]; ++j)
            for (int k = 0; k <= s->n[2]; ++k) {
                const double *J0 = AT(&s->J, i, j, k), *E0 = AT(&s->E, i, j, k);
                double t0 = (*AT(&s->rho2, i, j, k) - *AT(&s->rho1, i, j, k)) / s->tau +
                            (AT(&s->J, i + 1, j, k)[0] - J0[0]) / h1 +
                            (AT(&s->J, i, j + 1, k)[1] - J0[1]) / h2 +
                            (AT(&s->J, i, j, k + 1)[2] - J0[2]) / h3;
                double t1 = (AT(&s->E, i + 1, j, k)[0] - E0[0]) / h1 +
                            (AT(&s->E, i, j + 1, k)[1] - E0[1]) / h2 +
                            (AT(&s->E, i, j, k + 1)[2] - E0[2]) / h3 - *AT(&s->rho2, i, j, k);
                d0 = (fabs(t0) > d0) ? fabs(t0) : d0;
                d1 = (fabs(t1) > d1) ? fabs(t1) : d1;
            }
    res[0] = d0;
    res[1] = d1;
}

/* One iteration of the main loop, source/core/main.c:353-441. */
int opic_step(opic_t *s, double *diag, double *gauss)
{
    if (!s->inited) return -1;
    double WE, WM;
    em_HHalfStep(s);
    em_energy(s, &WE, &WM);
    if (gauss) rho_compute(s, &s->rho1);             /* gauss_before */
    /* plasma_move, source/core/plasma.c:268-305 */
    mesh_zero(&s->J);
    s->Wx = s->Wy = s->Wz = 0;
    long unprocessed = s->nCore;
    plasma_timestep(s, 1, s->nShell);
    plasma_walls(s);
    plasma_timestep(s, 0, s->nAll - unprocessed);
    if (diag) {
        diag[0] = s->time; diag[1] = WE; diag[2] = WM;
        diag[3] = s->Wx; diag[4] = s->Wy; diag[5] = s->Wz;
    }
    em_HStep(s);
    em_EStep_start(s);
    jbc_fold(s);
    em_EStep_finish(s);
    if (gauss) {
        rho_compute(s, &s->rho2);
        gauss_residuals(s, gauss);
    }
    s->time = s->time + s->tau;
    return 0;
}
