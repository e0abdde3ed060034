/*
 * mandor_setup.c -- the setup step of a run: reads an *.input deck and writes
 * record 0 (binData) and the control files of the sources, in the formats
 * setup.out writes (source/setup/main.c:155-256).
 *
 *   mandor_setup deck.input [create|estimate]
 *
 * The deck is a sequence of '[tag]' lines, each followed by its parameters:
 * '@' lines (one value, then a comment) and '>' option lines
 * (source/setup/tag.c:63-82, source/misc_cfgReader.c).  Lines that start
 * with neither are comments.  The first three tags are [units], [mesh] and
 * [boundaryConditions], in this order (source/setup/main.c:53-66); the
 * others run in file order:
 *
 *   [units]               @ lambda [micron], or -n_cr [cm^-3]   (tag_units.c:36)
 *   [mesh]                @ imax @ jmax @ kmax @ Lx @ Ly @ Lz [micron]
 *                         @ tau (negative: fraction of tau_max) (tag_mesh.c:40-106)
 *   [boundaryConditions]  @ xMin @ xMax @ yMin @ yMax @ zMin @ zMax
 *                         (0 periodic, 1 mirror, 2 Mur)         (tag_boundary.c:29)
 *   [plasma]              @ nx @ ny @ nz @ layers, then filters '@ boxes',
 *                         '@ planes', '@ spheres', '@ cylinders' with their
 *                         '>' blocks                            (tag_plasma.c:46-161)
 *   [DF:uniform]          @ rho [n_c] @ q/M @ ux @ uy @ uz @ dux @ duy @ duz
 *                                                               (tag_DF_uniform.c:36-70)
 *   [DF:Maxwell]          @ omega_pe^2 @ q/M @ ux @ uy @ uz @ tune energy
 *                         @ uniform weight                      (tag_maxwell.c:35-167)
 *   [DF:two-stream]       @ gamma @ q/M @ Nx @ Ny @ Nz @ mx @ my @ mz
 *                                                               (tag_twoStream.c:23-103)
 *   [DF:foil]             @ x0 @ y0 @ z0 [micron] @ lx @ ly @ lz @ width
 *                         [micron] (> shift) @ linear|exponential @ n1 @ n2
 *                         @ q @ M @ nx @ ny @ nz                (tag_foil.c:53-233)
 *   [gradient]            > x1 y1 z1 > x2 y2 z2 > d1 d2 type    (tag_gradient.c:22-51)
 *   [gaussSpot]           -> .gaussSpot.cfg                     (tag_gaussSpot.c:78-150)
 *   [TFSF]                -> .EM_sources/TFSF/TFSF.cfg, filter.cfg (tag_TFSF.c:78-182)
 *   [TFSF: open faces]    -> .EM_sources/TFSF/openFaces.cfg     (tag_TFSFFaces.c:44-80)
 *   [focused laser]       the reference's ten parameters (source/scpic/em_mirror.c:
 *                         292-328) -> .EM_sources/laser/mirror_NN; the map is a
 *                         paraxial Gaussian beam of waist 2 F lambda / pi, not
 *                         the Stratton-Chu integrals of em_mirror.c:497-660
 *
 * and three tags of this build:
 *   [RANMAR]              @ ij @ kl: seeds of RANMAR (source/misc_ranmar.c:78-146)
 *   [DF:thermal]          @ rho [n_c] @ q/M @ ux @ uy @ uz @ vth: Maxwellian
 *                         momenta (Box-Muller on RANMAR) about <u>
 *   [laser]               @ a0 @ waist [micron] @ xf @ yf @ zf [micron]
 *                         @ duration [fs, intensity FWHM] @ T0 [fs] @ phi0:
 *                         a y-polarised paraxial Gaussian beam
 * and one option: '> random' right after the four values of [plasma] places
 * its markers at RANMAR-drawn points of each cell instead of the lattice.
 * Every random draw comes from one RANMAR stream, seeded 1802 9373 unless a
 * [RANMAR] tag comes first.  [DF:uniform] draws its three numbers per
 * marker from it: the reference draws them with rand() after seeding it
 * from the wall clock (source/misc_partition.c:255), so its momenta cannot
 * be reproduced; positions, weights, charges and the marker order can.
 *
 * Plasma objects (source/setup/plasma.c:35-124): [plasma], [DF:foil] and
 * [DF:two-stream] each start a new object; [DF:uniform], [DF:Maxwell],
 * [DF:thermal] and [gradient] act on the newest one.  Record 0 holds the
 * objects newest first, with tracer ids 1..N in file order.
 */
#define _GNU_SOURCE
#include <dirent.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include "mandor_host.h"

/* CGS constants of source/misc_units.h:34-38 */
#define CGS_C     2.9979e+10
#define CGS_E     4.8032e-10
#define CGS_M     9.1094e-28
#define CGS_W_CM2 1e7

/* marker_t with the tracer id (source/type_marker.h), 64 bytes */
typedef struct {
    double x, y, z, vx, vy, vz;
    float  rho, qm;
    int    id, pad;
} rec64_t;

static void die(const char *what, const char *msg)
{
    fprintf(stderr, "mandor_setup: %s: %s\n", what, msg);
    exit(EXIT_FAILURE);
}

#define HOST(x, what)                          \
    do {                                       \
        if ((x) != 0) die(what, mh_error());   \
    } while (0)

/* ---- the deck ----------------------------------------------------------- */

typedef struct {
    char  kind;          /* '@' or '>' */
    int   line;
    char *text;          /* the line after its mark */
} item_t;

typedef struct {
    char    name[96];
    int     line, n, cap;
    item_t *it;
} sect_t;

typedef struct {
    const char *path;
    int         n, cap;
    sect_t     *s;        /* s[0]: lines before the first tag */
} deck_t;

static void *grow(void *p, int *cap, int need, size_t elem)
{
    if (need <= *cap) return p;
    *cap = *cap ? 2 * *cap : 16;
    if (*cap < need) *cap = need;
    p = realloc(p, elem * (size_t)*cap);
    if (!p) die("deck", "out of memory");
    return p;
}

static void deck_load(const char *path, deck_t *dk)
{
    FILE *fp = fopen(path, "rt");
    if (!fp) die(path, "cannot open");
    memset(dk, 0, sizeof(*dk));
    dk->path = path;
    dk->s = (sect_t *)grow(NULL, &dk->cap, 1, sizeof(sect_t));
    memset(dk->s, 0, sizeof(sect_t));
    dk->n = 1;
    char  *buf = NULL;
    size_t cap = 0;
    int    line = 0;
    while (getline(&buf, &cap, fp) >= 0) {
        ++line;
        char *s = buf;
        while (*s == ' ' || *s == '\t') ++s;
        s[strcspn(s, "\r\n")] = 0;
        if (*s == '[') {
            char *e = strchr(s, ']');
            if (!e || e - s - 1 >= (long)sizeof(dk->s->name)) {
                fprintf(stderr, "mandor_setup: %s:%d: bad tag line '%s'\n", path, line, s);
                exit(EXIT_FAILURE);
            }
            dk->s = (sect_t *)grow(dk->s, &dk->cap, dk->n + 1, sizeof(sect_t));
            sect_t *t = dk->s + dk->n++;
            memset(t, 0, sizeof(*t));
            memcpy(t->name, s + 1, (size_t)(e - s - 1));
            t->line = line;
        } else if (*s == '@' || *s == '>') {
            sect_t *t = dk->s + dk->n - 1;
            t->it = (item_t *)grow(t->it, &t->cap, t->n + 1, sizeof(item_t));
            item_t *q = t->it + t->n++;
            q->kind = *s;
            q->line = line;
            q->text = strdup(s + 1);
            if (!q->text) die("deck", "out of memory");
        }
    }
    free(buf);
    fclose(fp);
}

static void deck_free(deck_t *dk)
{
    for (int t = 0; t < dk->n; ++t) {
        for (int q = 0; q < dk->s[t].n; ++q) free(dk->s[t].it[q].text);
        free(dk->s[t].it);
    }
    free(dk->s);
}

/* A cursor over the parameters of one tag, read in order as the reference's
 * cfg_readDouble / cfg_readInt / cfg_readWord / cfg_readOpt* do. */
typedef struct {
    const deck_t *dk;
    const sect_t *s;
    int           k;
} cur_t;

static void fail_line(const cur_t *c, int line, const char *what, const char *msg)
{
    fprintf(stderr, "mandor_setup: %s:%d: [%s] %s: %s\n", c->dk->path, line, c->s->name, what, msg);
    exit(EXIT_FAILURE);
}

static void fail(const cur_t *c, const char *what, const char *msg)
{
    const int line = c->k < c->s->n ? c->s->it[c->k].line
                                    : (c->k ? c->s->it[c->k - 1].line : c->s->line);
    fail_line(c, line, what, msg);
}

static int is_param(const cur_t *c) { return c->k < c->s->n && c->s->it[c->k].kind == '@'; }
static int is_option(const cur_t *c) { return c->k < c->s->n && c->s->it[c->k].kind == '>'; }

static const item_t *take(cur_t *c, char kind, const char *what)
{
    if (c->k >= c->s->n || c->s->it[c->k].kind != kind)
        fail(c, what, kind == '@' ? "missing '@' parameter" : "missing '>' option");
    return c->s->it + c->k++;
}

static double num(cur_t *c, char kind, const char *what)
{
    const item_t *q = take(c, kind, what);
    char *e;
    const double v = strtod(q->text, &e);
    if (e == q->text) fail_line(c, q->line, what, "not a number");
    return v;
}

static int inum(cur_t *c, char kind, const char *what)
{
    const item_t *q = take(c, kind, what);
    char *e;
    const long v = strtol(q->text, &e, 10);
    if (e == q->text || v < INT_MIN || v > INT_MAX) fail_line(c, q->line, what, "not an integer");
    return (int)v;
}

/* k numbers on one line */
static void nums(cur_t *c, char kind, int k, double *v, const char *what)
{
    const item_t *q = take(c, kind, what);
    const char *s = q->text;
    for (int i = 0; i < k; ++i) {
        char *e;
        v[i] = strtod(s, &e);
        if (e == s) fail_line(c, q->line, what, "too few numbers");
        s = e;
    }
}

static const char *word(cur_t *c, char kind, const char *what)
{
    static char ring[8][200];
    static int  slot;
    const item_t *q = take(c, kind, what);
    const char *s = q->text;
    while (*s == ' ' || *s == '\t') ++s;
    size_t len = 0;
    while ((unsigned char)s[len] > ' ') ++len;
    if (!len || len >= sizeof(ring[0])) fail_line(c, q->line, what, "missing word");
    slot = (slot + 1) & 7;
    memcpy(ring[slot], s, len);
    ring[slot][len] = 0;
    return ring[slot];
}

/* ---- state of the run being set up -------------------------------------- */

static mh_units_t  U;
static mh_domain_t D;
static int         NC[3];          /* cells per axis */
static int         estimate_only;

/* parameters of the newest [plasma] tag (source/setup/tag_plasma.c:37) */
static int pl_nx, pl_ny, pl_nz, pl_layers, pl_ppc;

/* plasma objects, oldest first */
typedef struct {
    rec64_t *m;
    long     n, cap;
} obj_t;
static obj_t *objs;
static int    nobj, objcap;

static obj_t *new_object(void)
{
    objs = (obj_t *)grow(objs, &objcap, nobj + 1, sizeof(obj_t));
    obj_t *o = objs + nobj++;
    memset(o, 0, sizeof(*o));
    return o;
}

static rec64_t *add_marker(obj_t *o)
{
    if (o->n == o->cap) {
        o->cap = o->cap ? 2 * o->cap : 4096;
        o->m = (rec64_t *)realloc(o->m, sizeof(rec64_t) * (size_t)o->cap);
        if (!o->m) die("markers", "out of memory");
    }
    rec64_t *m = o->m + o->n++;
    memset(m, 0, sizeof(*m));
    return m;
}

static obj_t *newest(const cur_t *c, const char *what)
{
    if (!nobj) fail(c, what, "no plasma object to act on (a [plasma] tag must come first)");
    return objs + nobj - 1;
}

/* the RANMAR stream of every random draw */
static mh_ranmar_t RNG;
static int         rng_ready;

static void rng_draw(double *v, int n)
{
    if (!rng_ready) {
        HOST(mh_ranmar_init(&RNG, 1802, 9373), "RANMAR");
        rng_ready = 1;
    }
    mh_ranmar(&RNG, v, n);
}

static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* ---- [units] [mesh] [boundaryConditions] -------------------------------- */

static void tag_units(cur_t *c)
{
    const double p = num(c, '@', "lambda [micron] or -n_cr [cm^-3]");
    if (p > 0) {
        mh_units_from_lambda(p * 1e-4, &U);
    } else {
        /* omega_0 = omega_pe(n_cr), lambda = 2 pi c / omega_0 (misc_units.c:156-167) */
        const double ncr = -p;
        if (!(ncr > 0)) fail(c, "n_cr", "must be positive");
        const double omega = sqrt(4 * M_PI * ncr * CGS_E * CGS_E / CGS_M);
        mh_units_from_lambda(2 * M_PI * CGS_C / omega, &U);
    }
}

static void tag_mesh(cur_t *c)
{
    double L[3], tau;
    for (int a = 0; a < 3; ++a) NC[a] = inum(c, '@', "cells");
    for (int a = 0; a < 3; ++a) L[a] = num(c, '@', "length [micron]") * U.micron;
    tau = num(c, '@', "tau");
    memset(&D, 0, sizeof(D));
    for (int a = 0; a < 3; ++a) {
        NC[a] -= (NC[a] == 1);
        if (NC[a] < 2) fail(c, "cells", "this build is 3-D: every axis needs at least 2 cells");
        if (!(L[a] > 0)) fail(c, "length", "must be positive");
        D.L[a] = L[a];
        D.h[a] = L[a] / (double)NC[a];
        D.dmn_max[a] = NC[a];
    }
    const double s = 1.0 / (D.h[0] * D.h[0]) + 1.0 / (D.h[1] * D.h[1]) + 1.0 / (D.h[2] * D.h[2]);
    if (tau < 0) tau /= -sqrt(s);
    if (!(tau > 0)) fail(c, "tau", "must not be zero");
    if (!(tau * sqrt(s) < 0.99)) fail(c, "tau", "too big (Courant number >= 0.99)");
    if (!(D.h[0] < 0.2 && D.h[1] < 0.2 && D.h[2] < 0.2))
        fail(c, "mesh", "unresolved wavelength (a step >= 0.2 lambda)");
    D.tau = tau;
}

static void tag_boundary(cur_t *c)
{
    int v[6];
    for (int q = 0; q < 6; ++q) {
        v[q] = inum(c, '@', "boundary condition");
        if (v[q] < 0 || v[q] > 2) fail(c, "boundary condition", "0 periodic, 1 mirror or 2 Mur");
    }
    for (int a = 0; a < 3; ++a) {
        if ((v[2 * a] == 0) != (v[2 * a + 1] == 0))
            fail(c, "boundary condition", "a periodic axis is periodic on both faces");
        D.bc_min[a] = v[2 * a];
        D.bc_max[a] = v[2 * a + 1];
    }
}

/* ---- [plasma] and its shape filters (tag_plasma_*.c) ------------------- */

typedef struct { double c[3], d[3][3]; } box_t;        /* d: half-axes over their squared length */
typedef struct { int n; double (*pl)[4]; } poly_t;      /* planes A x + B y + C z + D <= 0 inside */
typedef struct { double c[3], r; } ball_t;
typedef struct { double p[3], a[3], r; } rod_t;

typedef struct {
    int     nbox, npoly, nball, nrod;
    box_t  *box;
    poly_t *poly;
    ball_t *ball;
    rod_t  *rod;
} shapes_t;

/* '> centre' '> half-axis 1' '> half-axis 2' '> half-axis 3' per box
 * (tag_plasma_box.c:24-47) */
static void read_boxes(cur_t *c, shapes_t *f)
{
    while (is_option(c)) {
        f->box = (box_t *)realloc(f->box, sizeof(box_t) * (size_t)(f->nbox + 1));
        if (!f->box) die("boxes", "out of memory");
        box_t *b = f->box + f->nbox++;
        nums(c, '>', 3, b->c, "box centre");
        const double cmp = 0.001 * (D.h[0] * D.h[0] + D.h[1] * D.h[1] + D.h[2] * D.h[2]);
        for (int q = 0; q < 3; ++q) {
            nums(c, '>', 3, b->d[q], "box half-axis");
            const double dd = dot3(b->d[q], b->d[q]);
            if (!(dd > cmp)) fail(c, "box half-axis", "too short");
            const double s = 1.0 / dd;
            for (int a = 0; a < 3; ++a) b->d[q][a] = s * b->d[q][a];
        }
    }
}

/* one polyhedron of '> point' '> normal' '> shift' planes (tag_plasma_convex.c:28-61) */
static void read_planes(cur_t *c, shapes_t *f)
{
    f->poly = (poly_t *)realloc(f->poly, sizeof(poly_t) * (size_t)(f->npoly + 1));
    if (!f->poly) die("planes", "out of memory");
    poly_t *P = f->poly + f->npoly++;
    P->n = 0;
    P->pl = NULL;
    while (is_option(c)) {
        double r[3], nv[3], shift;
        nums(c, '>', 3, r, "plane point");
        nums(c, '>', 3, nv, "plane normal");
        nums(c, '>', 1, &shift, "plane shift");
        const double n = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
        if (!(n > 0.1)) fail(c, "plane normal", "too short");
        P->pl = (double(*)[4])realloc(P->pl, sizeof(double[4]) * (size_t)(P->n + 1));
        if (!P->pl) die("planes", "out of memory");
        double *q = P->pl[P->n++];
        q[0] = nv[0] / n;
        q[1] = nv[1] / n;
        q[2] = nv[2] / n;
        q[3] = -(nv[0] * r[0] + nv[1] * r[1] + nv[2] * r[2]) / n - shift;
    }
}

/* '> x y z r' per sphere: r > 0 keeps the inside, r < 0 the outside
 * (tag_plasma_spheres.c:23-58) */
static void read_spheres(cur_t *c, shapes_t *f)
{
    while (is_option(c)) {
        f->ball = (ball_t *)realloc(f->ball, sizeof(ball_t) * (size_t)(f->nball + 1));
        if (!f->ball) die("spheres", "out of memory");
        double v[4];
        nums(c, '>', 4, v, "sphere");
        ball_t *b = f->ball + f->nball++;
        memcpy(b->c, v, sizeof(b->c));
        b->r = v[3];
    }
}

/* '> point' '> axis' '> r' per cylinder, r as for spheres
 * (tag_plasma_cylinders.c:24-71) */
static void read_cylinders(cur_t *c, shapes_t *f)
{
    while (is_option(c)) {
        f->rod = (rod_t *)realloc(f->rod, sizeof(rod_t) * (size_t)(f->nrod + 1));
        if (!f->rod) die("cylinders", "out of memory");
        rod_t *r = f->rod + f->nrod++;
        nums(c, '>', 3, r->p, "cylinder axis point");
        nums(c, '>', 3, r->a, "cylinder axis direction");
        nums(c, '>', 1, &r->r, "cylinder radius");
        const double aa = dot3(r->a, r->a);
        if (!(aa > 0.01)) fail(c, "cylinder axis direction", "too short");
        const double s = 1.0 / sqrt(aa);
        for (int a = 0; a < 3; ++a) r->a[a] = s * r->a[a];
    }
}

static void free_shapes(shapes_t *f)
{
    for (int q = 0; q < f->npoly; ++q) free(f->poly[q].pl);
    free(f->poly);
    free(f->box);
    free(f->ball);
    free(f->rod);
    memset(f, 0, sizeof(*f));
}

/* Each filter keeps the markers inside its shape; a removed marker is
 * replaced by the last untested one, which is tested next, so the survivors
 * come in the reference's order (tag_plasma_box.c:53-70 and the others). */
static int out_poly(const rec64_t *m, const poly_t *P)
{
    for (int q = 0; q < P->n; ++q) {
        const double *s = P->pl[q];
        if (s[0] * m->x + s[1] * m->y + s[2] * m->z + s[3] > 0) return 1;
    }
    return 0;
}

static int out_ball(const rec64_t *m, const ball_t *b)
{
    const double r2 = b->r * b->r;
    const double d2 = (m->x - b->c[0]) * (m->x - b->c[0]) + (m->y - b->c[1]) * (m->y - b->c[1]) +
                      (m->z - b->c[2]) * (m->z - b->c[2]);
    return (r2 - d2) * b->r < 0;
}

static int out_rod(const rec64_t *m, const rod_t *r)
{
    const double r2 = r->r * r->r;
    const double d[3] = {m->x - r->p[0], m->y - r->p[1], m->z - r->p[2]};
    const double x[3] = {d[1] * r->a[2] - d[2] * r->a[1], d[2] * r->a[0] - d[0] * r->a[2],
                         d[0] * r->a[1] - d[1] * r->a[0]};
    return (r2 - dot3(x, x)) * r->r < 0;
}

static int out_box(const rec64_t *m, const box_t *b)
{
    const double d[3] = {m->x - b->c[0], m->y - b->c[1], m->z - b->c[2]};
    return fabs(dot3(b->d[0], d)) > 1 || fabs(dot3(b->d[1], d)) > 1 || fabs(dot3(b->d[2], d)) > 1;
}

#define FILTER(p, N, test, shape)              \
    for (int q_ = 0; q_ < (N);) {              \
        if (test((p) + q_, shape))             \
            (p)[q_] = (p)[--(N)];              \
        else                                   \
            ++q_;                              \
    }

static void tag_plasma(cur_t *c)
{
    pl_nx = inum(c, '@', "nx");
    pl_ny = inum(c, '@', "ny");
    pl_nz = inum(c, '@', "nz");
    pl_layers = inum(c, '@', "velocity layers");
    int random_pos = 0;
    while (is_option(c)) {
        const char *w = word(c, '>', "option");
        if (strcmp(w, "random")) fail(c, w, "unknown [plasma] option (this build knows '> random')");
        random_pos = 1;
    }
    shapes_t f;
    memset(&f, 0, sizeof(f));
    while (is_param(c)) {
        const char *w = word(c, '@', "shape");
        if (!strcmp(w, "planes")) read_planes(c, &f);
        else if (!strcmp(w, "spheres")) read_spheres(c, &f);
        else if (!strcmp(w, "cylinders")) read_cylinders(c, &f);
        else if (!strcmp(w, "boxes")) read_boxes(c, &f);
        else fail(c, w, "bad shape ('planes', 'spheres', 'cylinders' or 'boxes')");
    }
    if (pl_nx < 1 || pl_ny < 1 || pl_nz < 1 || pl_layers < 1)
        fail(c, "sampling", "nx, ny, nz and layers must be positive");
    pl_ppc = pl_nx * pl_ny * pl_nz * pl_layers;

    /* the pattern of one cell: lattice points (i + 1/2)/n, each repeated per
     * velocity layer, vx holding the marker's place in the pattern */
    rec64_t *cell = (rec64_t *)malloc(sizeof(rec64_t) * 2 * (size_t)pl_ppc);
    if (!cell) die("[plasma]", "out of memory");
    rec64_t *kept = cell + pl_ppc;
    const double dx = D.h[0] / (double)pl_nx, dy = D.h[1] / (double)pl_ny,
                 dz = D.h[2] / (double)pl_nz;
    for (int i = 0, p = 0; i < pl_nx; ++i)
        for (int j = 0; j < pl_ny; ++j)
            for (int k = 0; k < pl_nz; ++k)
                for (int l = 0; l < pl_layers; ++l, ++p) {
                    rec64_t *m = cell + p;
                    memset(m, 0, sizeof(*m));
                    m->x = (i + 0.5) * dx;
                    m->y = (j + 0.5) * dy;
                    m->z = (k + 0.5) * dz;
                    m->vx = p;
                    m->vy = m->vz = NAN;
                    m->rho = m->qm = NAN;
                }
    obj_t *o = new_object();
    for (int I = 0; I < NC[0]; ++I)
        for (int J = 0; J < NC[1]; ++J)
            for (int K = 0; K < NC[2]; ++K) {
                for (int p = 0; p < pl_ppc; ++p) {
                    kept[p] = cell[p];
                    if (random_pos) {
                        double r[3];
                        rng_draw(r, 3);
                        kept[p].x = r[0] * D.h[0];
                        kept[p].y = r[1] * D.h[1];
                        kept[p].z = r[2] * D.h[2];
                    }
                    kept[p].x += I * D.h[0];
                    kept[p].y += J * D.h[1];
                    kept[p].z += K * D.h[2];
                }
                int n = pl_ppc;
                for (int q = 0; q < f.npoly; ++q) FILTER(kept, n, out_poly, f.poly + q);
                for (int q = 0; q < f.nball; ++q) FILTER(kept, n, out_ball, f.ball + q);
                for (int q = 0; q < f.nrod; ++q) FILTER(kept, n, out_rod, f.rod + q);
                for (int q = 0; q < f.nbox; ++q) FILTER(kept, n, out_box, f.box + q);
                for (int p = 0; p < n; ++p) *add_marker(o) = kept[p];
            }
    free(cell);
    free_shapes(&f);
}

/* ---- velocity distributions -------------------------------------------- */

static void tag_df_uniform(cur_t *c)
{
    double rho = num(c, '@', "rho [n_c]"), qm = num(c, '@', "q/M"), u[3], du[3];
    for (int a = 0; a < 3; ++a) u[a] = num(c, '@', "<u>");
    for (int a = 0; a < 3; ++a) du[a] = num(c, '@', "du");
    if (!(rho * qm > 0)) fail(c, "rho", "the sign of rho differs from the sign of q/M");
    if (!pl_ppc) fail(c, "rho", "needs a [plasma] tag before it");
    rho *= mh_rho_crit(&U) / pl_ppc;
    obj_t *o = newest(c, "[DF:uniform]");
    for (long q = 0; q < o->n; ++q) {
        rec64_t *m = o->m + q;
        double r[3];
        rng_draw(r, 3);
        m->vx = u[0] + (r[0] - 0.5) * du[0];
        m->vy = u[1] + (r[1] - 0.5) * du[1];
        m->vz = u[2] + (r[2] - 0.5) * du[2];
        m->rho = (float)rho;
        m->qm = (float)qm;
    }
}

static double gauss(void)
{
    double u[2];
    do rng_draw(u, 2);
    while (u[0] <= 0.0);
    return sqrt(-2.0 * log(u[0])) * cos(2.0 * M_PI * u[1]);
}

static void tag_df_thermal(cur_t *c)
{
    double rho = num(c, '@', "rho [n_c]"), qm = num(c, '@', "q/M"), u[3];
    for (int a = 0; a < 3; ++a) u[a] = num(c, '@', "<u>");
    const double vth = num(c, '@', "vth");
    if (!(rho * qm > 0)) fail(c, "rho", "the sign of rho differs from the sign of q/M");
    if (!pl_ppc) fail(c, "rho", "needs a [plasma] tag before it");
    rho *= mh_rho_crit(&U) / pl_ppc;
    obj_t *o = newest(c, "[DF:thermal]");
    for (long q = 0; q < o->n; ++q) {
        rec64_t *m = o->m + q;
        m->vx = u[0] + vth * gauss();
        m->vy = u[1] + vth * gauss();
        m->vz = u[2] + vth * gauss();
        m->rho = (float)rho;
        m->qm = (float)qm;
    }
}

/* Quiet start of Denavit and Walsh (Comments Plasma Phys. Control. Fusion 6,
 * 1981): sample i of a Hammersley set, the radix-2, -3 and -5 digit
 * reversals of i + 1 (source/setup/setup_denavit.c:24-53). */
static double radical_inverse(int num, int radix)
{
    const double inv = 1.0 / (double)radix;
    double res = 0, w = inv;
    for (; num > 0; num /= radix) {
        res += (num % radix) * w;
        w *= inv;
    }
    return res;
}

/* Inverse of the cumulative distribution of exp(-x^2) on [x0, x1], tabulated
 * at N + 1 points: the integral over each step from the two quadratic
 * splines through the neighbouring points, blended linearly, and inverted
 * with the quadratic through three table points on each side, blended the
 * same way (source/setup/setup_distrMapper.c:82-180). */
typedef struct {
    int     N;
    double *F, *X;
} cdf_t;

static double maxwell_f(double x) { return exp(-x * x); }

/* y(x) = a x^2 + b x + c through (-h, y0), (0, y1), (h, y2) */
static void quad_fit(double h, double y0, double y1, double y2, double *a, double *b, double *c)
{
    *a = (y2 - 2 * y1 + y0) / (2 * h * h);
    *b = (y2 - y0) / (2 * h);
    *c = y1;
}

/* integrals over [0, h] of y(x) and of y(x) x / h */
static double int_all(double h, double a, double b, double c)
{
    return a * h * h * h / 3.0 + 0.5 * b * h * h + c * h;
}

static double int_right(double h, double a, double b, double c)
{
    return a * h * h * h / 4.0 + b * h * h / 3.0 + 0.5 * c * h;
}

static double int_left(double h, double a, double b, double c)
{
    return int_all(h, a, b, c) - int_right(h, a, b, c);
}

static void cdf_build(cdf_t *m, double x0, double x1, int N)
{
    m->N = N;
    m->F = (double *)malloc(sizeof(double) * (size_t)(N + 1));
    m->X = (double *)malloc(sizeof(double) * (size_t)(N + 1));
    if (!m->F || !m->X) die("[DF:Maxwell]", "out of memory");
    const double dx = (x1 - x0) / (double)N;
    double x = x0 + dx;
    double y0 = maxwell_f(x - dx), y1 = maxwell_f(x), y2 = maxwell_f(x + dx), sum, a, b, cc;
    m->X[0] = x0;
    m->F[0] = 0;
    quad_fit(dx, y0, y1, y2, &a, &b, &cc);
    m->X[1] = x;
    sum = m->F[1] = int_all(dx, a, -b, cc);
    x += dx;
    for (int i = 2; i < N; ++i) {
        const double y3 = maxwell_f(x + dx);
        double A, B, C;
        quad_fit(dx, y1, y2, y3, &A, &B, &C);
        sum += int_left(dx, a, b, cc) + int_left(dx, A, -B, C);
        m->F[i] = sum;
        m->X[i] = x;
        x += dx;
        y1 = y2;
        y2 = y3;
        a = A;
        b = B;
        cc = C;
    }
    sum += int_all(dx, a, b, cc);
    m->F[N] = sum;
    m->X[N] = x;
    sum = 1.0 / m->F[N];
    for (int i = 1; i <= N; ++i) m->F[i] *= sum;
}

static double cdf_invert(double y, const cdf_t *m)
{
    if (!(y >= -1e-7 && y <= 1 + 1e-7)) die("[DF:Maxwell]", "quantile out of [0, 1]");
    const double *F = m->F, *X = m->X;
    int iL = 0, iR = m->N;
    while (iR - iL > 1) {
        const int i = (iL + iR) >> 1;
        if (y < F[i]) iR = i;
        else iL = i;
    }
    if (iL == 0 || iL == m->N - 1) {
        const double t = (y - F[iL]) / (F[iL + 1] - F[iL]);
        return X[iL] + t * (X[iL + 1] - X[iL]);
    }
    double g[2];
    for (int s = 0; s < 2; ++s) {
        const int i = iL + s;
        const double f0 = F[i - 1], f1 = F[i], f2 = F[i + 1];
        const double x0 = X[i - 1], x1 = X[i], x2 = X[i + 1];
        const double a = ((x2 - x1) / (f2 - f1) - (x0 - x1) / (f0 - f1)) / (f2 - f0);
        const double b = (x2 - x1) / (f2 - f1) - a * (f2 - f1);
        const double dy = y - F[i];
        g[s] = a * dy * dy + b * dy + X[i];
    }
    const double dy = y - F[iL + 1], w = dy / (F[iL + 1] - F[iL]);
    return -w * g[0] + (1.0 + w) * g[1];
}

static void cdf_free(cdf_t *m)
{
    free(m->F);
    free(m->X);
}

static void tag_df_maxwell(cur_t *c)
{
    const double w2 = num(c, '@', "omega_pe^2"), qm = num(c, '@', "q/M");
    double u[3];
    for (int a = 0; a < 3; ++a) u[a] = num(c, '@', "u_th");
    const int tune = inum(c, '@', "tune energy");
    const int uniform_weight = inum(c, '@', "uniform weight") == 1;
    if (!pl_ppc) fail(c, "omega_pe^2", "needs a [plasma] tag before it");
    const int ppc = pl_ppc;
    /* omega_pe^2 = rho q/M in units of omega_0^2 / (4 pi^2) */
    const double rho = w2 * 4 * M_PI * M_PI / qm / (double)ppc;
    if (!(qm * rho > 0)) fail(c, "omega_pe^2", "inconsistent q/M and density");
    if (pl_layers % 2) fail(c, "layers", "needs an even number of velocity layers");

    cdf_t all, pos;
    cdf_build(&all, -6.0, 6.0, 5000);
    cdf_build(&pos, 0.0, 6.0, 5000);
    /* one cell's pattern of velocities and weights, in mirrored pairs */
    rec64_t *pat = (rec64_t *)calloc((size_t)ppc, sizeof(rec64_t));
    if (!pat) die("[DF:Maxwell]", "out of memory");
    for (int p = 0; p < ppc; p += 2) {
        double vx = radical_inverse(p / 2 + 1, 2), vy = radical_inverse(p / 2 + 1, 3),
               vz = radical_inverse(p / 2 + 1, 5);
        if (uniform_weight) {
            vx = cdf_invert(vx, &pos);
            vy = cdf_invert(vy, &all);
            vz = cdf_invert(vz, &all);
            pat[p].rho = 1.0f;
        } else {
            vx *= 1.6;
            vy = 1.6 * 2 * (vy - 0.5);
            vz = 1.6 * 2 * (vz - 0.5);
            pat[p].rho = (float)(maxwell_f(vx) * maxwell_f(vy) * maxwell_f(vz));
        }
        pat[p].vx = vx * u[0];
        pat[p].vy = vy * u[1];
        pat[p].vz = vz * u[2];
        pat[p + 1].vx = -vx * u[0];
        pat[p + 1].vy = -vy * u[1];
        pat[p + 1].vz = -vz * u[2];
        pat[p + 1].rho = pat[p].rho;
    }
    cdf_free(&all);
    cdf_free(&pos);

    double norm = 0;
    for (int p = 0; p < ppc; ++p) norm += pat[p].rho;
    norm = ppc * rho / norm;
    for (int p = 0; p < ppc; ++p) pat[p].rho *= norm;
    if (tune) {
        double W[3] = {0, 0, 0};
        for (int p = 0; p < ppc; ++p) {
            W[0] += pat[p].vx * pat[p].vx * pat[p].rho;
            W[1] += pat[p].vy * pat[p].vy * pat[p].rho;
            W[2] += pat[p].vz * pat[p].vz * pat[p].rho;
        }
        if (W[0] == 0 || W[1] == 0 || W[2] == 0) fail(c, "u_th", "a zero temperature cannot be tuned");
        for (int a = 0; a < 3; ++a) W[a] = u[a] * sqrt(rho * ppc * 0.5 / W[a]);
        for (int p = 0; p < ppc; ++p) {
            pat[p].vx *= W[0];
            pat[p].vy *= W[1];
            pat[p].vz *= W[2];
        }
    }
    obj_t *o = newest(c, "[DF:Maxwell]");
    for (long q = 0; q < o->n; ++q) {
        rec64_t *m = o->m + q;
        const int l = (int)(m->vx + 0.1 + 128.0) - 128;   /* the place stored by [plasma] */
        if (l < 0 || l >= ppc) fail(c, "[DF:Maxwell]", "the newest object is not a [plasma] lattice");
        m->vx = pat[l].vx;
        m->vy = pat[l].vy;
        m->vz = pat[l].vz;
        m->rho = pat[l].rho;
        m->qm = (float)qm;
    }
    free(pat);
}

/* two cold counter-streaming beams tuned to the growth rate gamma of the
 * mode (mx, my, mz) (tag_twoStream.c:23-103) */
static void tag_two_stream(cur_t *c)
{
    const double g = num(c, '@', "growth rate"), qm = num(c, '@', "q/M");
    int Nb[3], mode[3];
    for (int a = 0; a < 3; ++a) Nb[a] = inum(c, '@', "markers per beam and axis");
    for (int a = 0; a < 3; ++a) mode[a] = inum(c, '@', "mode");
    const int Nbeam = Nb[0] * Nb[1] * Nb[2];
    double rho = g * g / qm;
    double k[3];
    for (int a = 0; a < 3; ++a) k[a] = 2 * M_PI * mode[a] / D.L[a];
    const double kabs = sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
    if (!(kabs > 0)) fail(c, "mode", "the mode vector is zero");
    const double V = g * sqrt(3.0) / (2 * kabs);
    double v[3];
    for (int a = 0; a < 3; ++a) v[a] = V * k[a] / kabs;
    if (!(Nb[0] > 0 && Nb[1] > 0 && Nb[2] > 0 && g > 0 && fabs(qm) <= 1))
        fail(c, "parameters", "need N > 0, gamma > 0 and |q/M| <= 1");
    rho *= D.L[0] * D.L[1] * D.L[2] / (D.h[0] * D.h[1] * D.h[2] * Nbeam);
    double d[3];
    for (int a = 0; a < 3; ++a) d[a] = D.L[a] / (double)Nb[a];
    obj_t *o = new_object();
    for (int i = 0; i < Nb[0]; ++i)
        for (int j = 0; j < Nb[1]; ++j)
            for (int kk = 0; kk < Nb[2]; ++kk)
                for (int s = 0; s < 2; ++s) {
                    rec64_t *m = add_marker(o);
                    m->x = (i + 0.5) * d[0];
                    m->y = (j + 0.5) * d[1];
                    m->z = (kk + 0.5) * d[2];
                    m->vx = s ? -v[0] : v[0];
                    m->vy = s ? -v[1] : v[1];
                    m->vz = s ? -v[2] : v[2];
                    m->rho = (float)rho;
                    m->qm = (float)qm;
                }
}

/* a layer of width w normal to l through (x0, y0, z0) + shift l, density
 * n1 -> n2 across it, at nx*ny*nz lattice markers per cell, 5 cells off
 * non-periodic walls (tag_foil.c:53-233) */
static void tag_df_foil(cur_t *c)
{
    double r0[3], l[3];
    for (int a = 0; a < 3; ++a) r0[a] = num(c, '@', "centre [micron]") * U.micron;
    for (int a = 0; a < 3; ++a) l[a] = num(c, '@', "normal");
    const double len = sqrt(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
    if (!(len > 1.0e-2)) fail(c, "normal", "too short");
    for (int a = 0; a < 3; ++a) l[a] /= len;
    const double width = num(c, '@', "width [micron]") * U.micron;
    double shift = 0;
    if (is_option(c)) shift = num(c, '>', "shift [micron]") * U.micron;
    const char *prof = word(c, '@', "profile");
    const int expo = !strcmp(prof, "exponential");
    if (!expo && strcmp(prof, "linear")) fail(c, prof, "profile is 'linear' or 'exponential'");
    double n1 = num(c, '@', "n1"), n2 = num(c, '@', "n2");
    const double q = num(c, '@', "q [e]"), M = num(c, '@', "M [m_e]");
    n1 *= (n1 < 0) ? (-U.ncr) : 1;       /* negative: in units of n_c */
    n2 *= (n2 < 0) ? (-U.ncr) : 1;
    int ns[3];
    for (int a = 0; a < 3; ++a) ns[a] = inum(c, '@', "markers per cell and axis");
    if (!(width >= (D.h[0] + D.h[1] + D.h[2]) / 3)) fail(c, "width", "thinner than a cell");
    if (ns[0] < 1 || ns[1] < 1 || ns[2] < 1) fail(c, "markers per cell", "must be positive");
    n1 *= q * CGS_E / U.rho;
    n2 *= q * CGS_E / U.rho;
    const double d[3] = {D.h[0] / (double)ns[0], D.h[1] / (double)ns[1], D.h[2] / (double)ns[2]};
    const double margin = 1e-10 * (D.h[0] + D.h[1] + D.h[2]) / (3 * width);
    int lo[3], hi[3];
    for (int a = 0; a < 3; ++a) {
        const int wall = 5 * (D.bc_min[a] != 0);
        lo[a] = wall;
        hi[a] = NC[a] - wall - 1;
    }
    obj_t *o = new_object();
    for (int I = lo[0]; I <= hi[0]; ++I)
        for (int J = lo[1]; J <= hi[1]; ++J)
            for (int K = lo[2]; K <= hi[2]; ++K)
                for (int i = 0; i < ns[0]; ++i)
                    for (int j = 0; j < ns[1]; ++j)
                        for (int k = 0; k < ns[2]; ++k) {
                            const double x = I * D.h[0] + (i + 0.5) * d[0];
                            const double y = J * D.h[1] + (j + 0.5) * d[1];
                            const double z = K * D.h[2] + (k + 0.5) * d[2];
                            double alpha = shift - ((x - r0[0]) * l[0] + (y - r0[1]) * l[1] +
                                                    (z - r0[2]) * l[2]);
                            alpha = alpha / width + 0.5;
                            if (alpha < margin || alpha > 1.0 - margin) continue;
                            rec64_t *m = add_marker(o);
                            m->x = x;
                            m->y = y;
                            m->z = z;
                            m->qm = (float)(q / M);
                            m->rho = (float)alpha;     /* the place across the layer, for now */
                        }
    const double weight = 1.0 / ((double)ns[0] * ns[1] * ns[2]);
    for (long k = 0; k < o->n; ++k) {
        rec64_t *m = o->m + k;
        m->rho = expo ? (float)(weight * n1 * pow(n2 / n1, m->rho))
                      : (float)(weight * (n1 + (n2 - n1) * m->rho));
    }
}

/* weights scaled by d1 -> d2 (linear, type 0, or exponential) between two
 * points given in code units (tag_gradient.c:22-51) */
static void tag_gradient(cur_t *c)
{
    double p1[3], p2[3];
    nums(c, '>', 3, p1, "point 1");
    nums(c, '>', 3, p2, "point 2");
    const item_t *q = take(c, '>', "d1 d2 type");
    double d1, d2;
    int type;
    if (sscanf(q->text, "%lf %lf %d", &d1, &d2, &type) != 3) fail_line(c, q->line, "d1 d2 type", "cannot read");
    const double l[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
    const double width = l[0] * l[0] + l[1] * l[1] + l[2] * l[2];
    obj_t *o = newest(c, "[gradient]");
    for (long k = 0; k < o->n; ++k) {
        rec64_t *m = o->m + k;
        const double s = ((m->x - p1[0]) * l[0] + (m->y - p1[1]) * l[1] + (m->z - p1[2]) * l[2]) / width;
        if (s > 0 && s < 1) {
            if (type == 0) m->rho *= d1 + (d2 - d1) * s;
            else m->rho *= d1 * pow(d2 / d1, s);
        }
    }
}

/* ---- sources ------------------------------------------------------------ */

static int spot_placed, tfsf_placed, open_faces_seen, nlasers;
static int faces_open[6];

static void tag_gauss_spot(cur_t *c, const char *deckname)
{
    const double I = num(c, '@', "intensity [W/cm^2] or -a0");
    const double Omega = num(c, '@', "frequency [omega_0]");
    const double X = num(c, '@', "X [micron]") * U.micron, Y = num(c, '@', "Y [micron]") * U.micron,
                 Z = num(c, '@', "Z [micron]") * U.micron;
    double R = num(c, '@', "width"), Tau = num(c, '@', "duration");
    const double front = num(c, '@', "front offset [fs]") * U.femtosecond;
    double plateau = 0;
    if (is_option(c)) plateau = num(c, '>', "plateau [fs]") * U.femtosecond;
    int EY = inum(c, '@', "Ey"), EZ = inum(c, '@', "Ez");
    if (is_option(c)) (void)num(c, '>', "TFSF position [micron]");
    if (spot_placed) fail(c, "[gaussSpot]", "only one gaussian hard source");
    spot_placed = 1;
    R = (R > 0) ? R * U.micron / sqrt(2.0 * log(2.0)) : -R;
    Tau = (Tau > 0) ? Tau * U.femtosecond / sqrt(2.0 * log(2.0)) : -Tau;
    if ((X - 5 * D.h[0]) * (D.L[0] - 5 * D.h[0] - X) < 0 || (Y - 5 * D.h[1]) * (D.L[1] - 5 * D.h[1] - Y) < 0 ||
        (Z - 5 * D.h[2]) * (D.L[2] - 5 * D.h[2] - Z) < 0)
        fail(c, "position", "the spot must lie 5 cells or more inside the domain");
    if (!(plateau > -1.0e-10 * D.tau)) fail(c, "plateau", "negative duration");
    EY = (EY > 0) - (EY < 0);
    EZ = (EZ > 0) - (EZ < 0);
    if (!EY && !EZ) fail(c, "polarisation", "both components are zero");
    const double A = (I > 0) ? sqrt(4 * M_PI * I * CGS_W_CM2 / ((abs(EY) + abs(EZ)) * 0.5 * CGS_C)) / U.E
                             : -I * U.A / U.E;
    if (estimate_only) return;
    FILE *fp = fopen(".gaussSpot.cfg", "wt");
    if (!fp) die(".gaussSpot.cfg", "cannot create");
    fprintf(fp, "@ %e       Amplitude [E0].\n", A);
    fprintf(fp, "@ %e       Cyclic frequency [omega_0].\n", Omega);
    fprintf(fp, "@ %e       Focal plane X [r0].\n", X);
    fprintf(fp, "@ %e       Focus Y [r0].\n", Y);
    fprintf(fp, "@ %e       Focus Z [r0].\n", Z);
    fprintf(fp, "@ %e       Spot size [r0].\n", R);
    fprintf(fp, "@ %e       Duration [t0].\n", Tau);
    fprintf(fp, "@ %e       Front offset [t0].\n", front);
    fprintf(fp, "@ %e       Plateau duration [t0].\n", plateau);
    fprintf(fp, "@ %d       Ey polarisation.\n", EY);
    fprintf(fp, "@ %d       Ez polarisation.\n", EZ);
    fprintf(fp, "\n\n   Written by mandor_setup from '%s'.\n", deckname);
    fclose(fp);
}

static void remove_dir_files(const char *dir)
{
    DIR *d = opendir(dir);
    if (!d) return;
    struct dirent *e;
    char p[1024];
    while ((e = readdir(d))) {
        if (!strcmp(e->d_name, ".") || !strcmp(e->d_name, "..")) continue;
        snprintf(p, sizeof(p), "%s/%s", dir, e->d_name);
        remove(p);
    }
    closedir(d);
}

/* consistencyCheck.dat of an earlier recording: the region's sizes, tau
 * and the steps must be the same to add to it (tag_TFSF.c:51-72) */
static int tfsf_same_region(const int reg[6])
{
    FILE *fp = fopen(".EM_sources/TFSF/consistencyCheck.dat", "rt");
    if (!fp) return 1;
    fclose(fp);
    deck_t dk;
    deck_load(".EM_sources/TFSF/consistencyCheck.dat", &dk);
    cur_t c = {&dk, dk.s, 0};
    int di[3];
    double v[4];
    for (int a = 0; a < 3; ++a) di[a] = inum(&c, '@', "region size");
    for (int a = 0; a < 4; ++a) v[a] = num(&c, '@', "tau and steps");
    deck_free(&dk);
    int same = fabs(v[0] - D.tau) < 1e-13 * D.tau;
    for (int a = 0; a < 3; ++a)
        same = same && reg[2 * a + 1] - reg[2 * a] == di[a] && fabs(v[1 + a] - D.h[a]) < 1e-13 * D.h[a];
    return same;
}

static void tag_tfsf(cur_t *c)
{
    int reg[6];   /* imin imax jmin jmax kmin kmax */
    for (int q = 0; q < 6; ++q) reg[q] = inum(c, '@', "bounding box");
    const char *mode_s = word(c, '@', "mode");
    int mode = -1;
    if (!strcmp(mode_s, "record")) mode = 1;
    else if (!strcmp(mode_s, "playForward")) mode = 0;
    else if (!strcmp(mode_s, "playBackward")) mode = 2;
    else fail(c, mode_s, "mode is 'record', 'playForward' or 'playBackward'");
    if (tfsf_placed) fail(c, "[TFSF]", "only one [TFSF] tag");
    tfsf_placed = 1;
    for (int a = 0; a < 3; ++a)
        if (reg[2 * a] <= 5 || reg[2 * a + 1] >= NC[a] - 5)
            fail(c, "bounding box", "the interface must lie more than 5 cells inside the domain");
    int add = 0, fmin = INT_MIN, fmax = INT_MAX, shift = 0;
    if (is_option(c)) {
        add = !strcmp(word(c, '>', "blend mode"), "add");
        if (add && !tfsf_same_region(reg))
            fail(c, "add", "the region differs from the one already recorded");
        fmin = inum(c, '>', "I0");
        fmax = inum(c, '>', "I1");
        if (is_option(c)) shift = inum(c, '>', "shift frames");
    }
    if (estimate_only) return;
    if (mode == 1 && !add) remove_dir_files(".EM_sources/TFSF");
    FILE *fp = fopen(".EM_sources/TFSF/TFSF.cfg", "wt");
    if (!fp) die("TFSF.cfg", "cannot create");
    static const char *lim[6] = {"reg.min[0]", "reg.max[0]", "reg.min[1]", "reg.max[1]", "reg.min[2]", "reg.max[2]"};
    for (int q = 0; q < 6; ++q) fprintf(fp, "@ %5d       Interface bounding box: %s.\n", reg[q], lim[q]);
    fprintf(fp, "@ %5d       Interface regime (%s).\n", mode, mode_s);
    fprintf(fp, "\n\n   Written by mandor_setup.\n");
    fclose(fp);
    fp = fopen(".EM_sources/TFSF/filter.cfg", "wt");
    if (!fp) die("filter.cfg", "cannot create");
    fprintf(fp, "@ % 5d       I0.\n", fmin);
    fprintf(fp, "@ % 5d       I1.\n", fmax);
    fprintf(fp, "@ % 5d       Shift frames.\n", shift);
    fclose(fp);
}

static void tag_tfsf_faces(cur_t *c)
{
    static const char *names[6] = {"xMin", "xMax", "yMin", "yMax", "zMin", "zMax"};
    if (open_faces_seen) fprintf(stderr, "mandor_setup: warning: [TFSF: open faces] is used twice\n");
    open_faces_seen = 1;
    while (is_option(c)) {
        const char *w = word(c, '>', "face");
        int f = 0;
        while (f < 6 && strcmp(names[f], w)) ++f;
        if (f == 6) fail(c, w, "faces are xMin, xMax, yMin, yMax, zMin, zMax");
        faces_open[f] = 1;
    }
    if (estimate_only) return;
    FILE *fp = fopen(".EM_sources/TFSF/openFaces.cfg", "wt");
    if (!fp) die("openFaces.cfg", "cannot create");
    for (int f = 0; f < 6; ++f)
        fprintf(fp, "@ %d        %s (0 engaged, 1 opened)\n", faces_open[f], names[f]);
    fprintf(fp, "\n\n  Written by mandor_setup.\n");
    fclose(fp);
}

/* y-polarised paraxial Gaussian beam focused at f; records {y_cos, y_sin,
 * z_cos, z_sin} of the tangential incident field on the source planes for
 * j, k in [5, n-5] (E map at x = (I-1)h drives H at plane I; H map at
 * x = (I-1/2)h drives E at plane I-1). */
static void beam_map(double w0, const double f[3], double *map)
{
    const int I = 5;
    const double k0 = 2 * M_PI, zr = M_PI * w0 * w0;
    const long nj = NC[1] + 1 - 10, nk = NC[2] + 1 - 10, nodes = nj * nk;
    for (int pass = 0; pass < 2; ++pass) {
        const double x = (pass == 0 ? I - 1.0 : I - 0.5) * D.h[0];
        const double xi = x - f[0];
        const double w = w0 * sqrt(1 + (xi / zr) * (xi / zr));
        const double invR = xi / (xi * xi + zr * zr);
        const double gouy = atan2(xi, zr);
        for (long j = 0; j < nj; ++j)
            for (long k = 0; k < nk; ++k) {
                const double y = (j + 5 - 0.5) * D.h[1], z = (k + 5) * D.h[2];
                const double r2 = (y - f[1]) * (y - f[1]) + (z - f[2]) * (z - f[2]);
                const double amp = (w0 / w) * exp(-r2 / (w * w));
                const double th = -k0 * r2 * invR / 2 + gouy - k0 * (x - I * D.h[0]);
                double *m = map + 4 * (pass * nodes + j * nk + k);
                const double cs = amp * cos(th), sn = -amp * sin(th);
                if (pass == 0) {
                    m[0] = cs; m[1] = sn; m[2] = 0; m[3] = 0;
                } else {
                    m[0] = cs; m[1] = sn; m[2] = -cs; m[3] = -sn;
                }
            }
    }
}

/* p: A (intensity or -a0), f-number, xf, yf, zf [micron], duration [fs],
 * T0 [fs], phi0; w0: waist [lambda] */
static void write_laser(cur_t *c, const double p[8], double w0)
{
    if (NC[1] < 11 || NC[2] < 11) fail(c, "laser", "needs at least 11 cells along y and z");
    if (estimate_only) return;
    if (!nlasers) remove_dir_files(".EM_sources/laser");
    const long nodes = (long)(NC[1] + 1 - 10) * (NC[2] + 1 - 10);
    double *map = (double *)malloc(sizeof(double) * 8 * (size_t)nodes);
    if (!map) die("laser", "out of memory");
    const double f[3] = {p[2] * U.micron, p[3] * U.micron, p[4] * U.micron};
    beam_map(w0, f, map);
    const double focus_um[3] = {p[2], p[3], p[4]};
    char path[64];
    snprintf(path, sizeof(path), ".EM_sources/laser/mirror_%02d", nlasers++);
    HOST(mh_laser_save(path, &D, &U, p[0], p[1], focus_um, p[5], p[6], p[7], 1, map), path);
    free(map);
}

/* [focused laser] with the reference's parameters (em_mirror.c:292-328) */
static void tag_focused_laser(cur_t *c)
{
    double p[8];
    p[0] = num(c, '@', "intensity [W/cm^2] or -a0");
    p[1] = num(c, '@', "f-number");
    for (int q = 2; q < 5; ++q) p[q] = num(c, '@', "focus [micron]");
    p[5] = num(c, '@', "duration [fs]");
    p[6] = num(c, '@', "focus arrival time [fs]");
    p[7] = num(c, '@', "phase");
    (void)num(c, '@', "incident to mirror radius ratio");
    if (inum(c, '@', "direction") <= 0) fail(c, "direction", "this build injects along +x only");
    if (!(p[1] > 0.25)) fail(c, "f-number", "must be greater than 0.25");
    /* far-field half-angle 1/(2F) of a Gaussian beam: w0 = 2 F lambda / pi */
    write_laser(c, p, 2 * p[1] / M_PI);
}

static void tag_laser(cur_t *c)
{
    double v[8];
    for (int q = 0; q < 8; ++q) v[q] = num(c, '@', "laser parameter");
    const double p[8] = {-v[0], 0.9, v[2], v[3], v[4], v[5], v[6], v[7]};
    write_laser(c, p, v[1] * U.micron);
}

/* ---- record 0 ----------------------------------------------------------- */

/* plasma_rho (source/setup/plasma.c:129-144): CIC charge density of the
 * markers in record order on the nodes [-2, n+2] (k fastest), each marker's
 * eight shares added x fastest, then z (plasmaRho_add with the staged
 * mf_PIC_distribute of source/misc_PIC.h:164-201), then the wall folds in
 * the order xMin xMax yMin yMax zMin zMax (plasmaRho_applyBoundary,
 * source/core/plasma_rho.c:18-96). */
static void cic_rho(const rec64_t *M, long nm, double *rho)
{
    const long W[3] = {(long)(NC[1] + 5) * (NC[2] + 5), NC[2] + 5, 1};
#define RHO(i, j, k) rho[((long)(i) + 2) * W[0] + ((long)(j) + 2) * W[1] + (k) + 2]
    for (long q = 0; q < nm; ++q) {
        const double r[3] = {M[q].x, M[q].y, M[q].z};
        int c[3];
        double s[3];
        for (int a = 0; a < 3; ++a) {
            s[a] = r[a] / D.h[a];
            c[a] = (int)(s[a] + 128.0) - 128;
            s[a] -= c[a];
        }
        const double v = M[q].rho;
        for (int dk = 0; dk < 2; ++dk) {
            const double ez = dk ? s[2] * v : (1 - s[2]) * v;
            for (int dj = 0; dj < 2; ++dj) {
                const double ey = dj ? s[1] * ez : (1 - s[1]) * ez;
                RHO(c[0], c[1] + dj, c[2] + dk) += (1 - s[0]) * ey;
                RHO(c[0] + 1, c[1] + dj, c[2] + dk) += s[0] * ey;
            }
        }
    }
    for (int b = 0; b < 6; ++b) {
        const int ea = b >> 1, top = b & 1, ep = ea == 0 ? 1 : 0, eq = ea == 2 ? 1 : 2;
        const int bc = top ? D.bc_max[ea] : D.bc_min[ea], N = NC[ea];
        if (bc == 0 && top) continue;            /* periodic: folded once, at the min wall */
        for (int p = -2; p <= NC[ep] + 2; ++p)
            for (int q = -2; q <= NC[eq] + 2; ++q) {
                long o[3];
                o[ep] = p;
                o[eq] = q;
#define AT(a) (o[ea] = (a), &RHO(o[0], o[1], o[2]))
                if (bc == 0) {
                    for (int a = -2; a <= 2; ++a) {
                        double *lo = AT(a), *hi = AT(a + N);
                        *hi += *lo;
                        *lo = *hi;
                    }
                } else {
                    const int pl = top ? N : 0, so = top ? 1 : -1;
                    *AT(pl - so) += *AT(pl + so);
                    *AT(pl - 2 * so) += *AT(pl + 2 * so);
                    *AT(pl) = 0;
                    *AT(pl + so) = -*AT(pl - so);
                    *AT(pl + 2 * so) = -*AT(pl - 2 * so);
                }
#undef AT
            }
    }
#undef RHO
}

/* ---- main --------------------------------------------------------------- */

static void clear_old_outputs(void)
{
    remove(".gaussSpot.cfg");
    remove(".EM_sources/TFSF/TFSF.cfg");
    remove(".EM_sources/TFSF/openFaces.cfg");
    DIR *d = opendir("binData");
    if (!d) return;
    struct dirent *e;
    char p[1024];
    while ((e = readdir(d))) {
        if (e->d_name[0] == '.' || !strchr(e->d_name, '.')) continue;
        snprintf(p, sizeof(p), "binData/%s", e->d_name);
        remove(p);
    }
    closedir(d);
}

int main(int argc, char **argv)
{
    if (argc < 2 || (argc > 2 && strcmp(argv[2], "create") && strcmp(argv[2], "estimate"))) {
        fprintf(stderr, "usage: %s deck.input [create|estimate]\n", argv[0]);
        return EXIT_FAILURE;
    }
    estimate_only = argc > 2 && !strcmp(argv[2], "estimate");
    deck_t dk;
    deck_load(argv[1], &dk);
    static const char *first[3] = {"units", "mesh", "boundaryConditions"};
    for (int t = 0; t < 3; ++t)
        if (dk.n <= t + 1 || strcmp(dk.s[t + 1].name, first[t]))
            die(argv[1], "the first tags must be [units], [mesh] and [boundaryConditions]");
    if (!estimate_only) {
        mkdir("binData", 0755);
        mkdir("tmp", 0755);
        mkdir("output", 0755);
        mkdir(".EM_sources", 0755);
        mkdir(".EM_sources/laser", 0755);
        mkdir(".EM_sources/TFSF", 0755);
        clear_old_outputs();
    }
    for (int t = 1; t < dk.n; ++t) {
        cur_t c = {&dk, dk.s + t, 0};
        const char *name = dk.s[t].name;
        if (!strcmp(name, "units")) tag_units(&c);
        else if (!strcmp(name, "mesh")) tag_mesh(&c);
        else if (!strcmp(name, "boundaryConditions")) tag_boundary(&c);
        else if (!strcmp(name, "plasma")) tag_plasma(&c);
        else if (!strcmp(name, "DF:uniform")) tag_df_uniform(&c);
        else if (!strcmp(name, "DF:Maxwell")) tag_df_maxwell(&c);
        else if (!strcmp(name, "DF:two-stream")) tag_two_stream(&c);
        else if (!strcmp(name, "DF:foil")) tag_df_foil(&c);
        else if (!strcmp(name, "gradient")) tag_gradient(&c);
        else if (!strcmp(name, "gaussSpot")) tag_gauss_spot(&c, argv[1]);
        else if (!strcmp(name, "TFSF")) tag_tfsf(&c);
        else if (!strcmp(name, "TFSF: open faces")) tag_tfsf_faces(&c);
        else if (!strcmp(name, "focused laser")) tag_focused_laser(&c);
        else if (!strcmp(name, "RANMAR")) {
            const int ij = inum(&c, '@', "ij"), kl = inum(&c, '@', "kl");
            HOST(mh_ranmar_init(&RNG, ij, kl), "[RANMAR]");
            rng_ready = 1;
        } else if (!strcmp(name, "DF:thermal")) tag_df_thermal(&c);
        else if (!strcmp(name, "laser")) tag_laser(&c);
        else {
            fprintf(stderr, "mandor_setup: %s:%d: unknown or unsupported tag [%s]\n", argv[1], dk.s[t].line,
                    name);
            return EXIT_FAILURE;
        }
        if (c.k < c.s->n)
            fprintf(stderr, "mandor_setup: %s:%d: [%s]: %d parameter line(s) ignored\n", argv[1],
                    c.s->it[c.k].line, name, c.s->n - c.k);
    }

    /* record 0 holds the objects newest first (source/setup/plasma.c:93-124) */
    long nm = 0;
    for (int q = 0; q < nobj; ++q) nm += objs[q].n;
    if (nm >= INT_MAX) die("markers", "too many markers for 32-bit tracer ids");
    printf("mandor_setup: %d x %d x %d cells, h = (%g, %g, %g), tau = %g, %ld markers (%.3e bytes)\n", NC[0],
           NC[1], NC[2], D.h[0], D.h[1], D.h[2], D.tau, nm, (double)nm * sizeof(rec64_t));
    if (estimate_only) {
        deck_free(&dk);
        return 0;
    }
    rec64_t *M = (rec64_t *)malloc(sizeof(rec64_t) * (size_t)(nm ? nm : 1));
    if (!M) die("markers", "out of memory");
    long at = 0;
    for (int q = nobj - 1; q >= 0; --q) {
        if (objs[q].n) memcpy(M + at, objs[q].m, sizeof(rec64_t) * (size_t)objs[q].n);
        at += objs[q].n;
        free(objs[q].m);
    }
    free(objs);
    for (long q = 0; q < nm; ++q) {
        M[q].id = (int)(q + 1);
        M[q].pad = 0;
    }

    HOST(mh_domain_save("binData/domain.bin", &D), "domain.bin");
    HOST(mh_units_save("binData/units.bin", &U), "units.bin");
    const long nodes = (long)(NC[0] + 1) * (NC[1] + 1) * (NC[2] + 1);
    const long gnodes = (long)(NC[0] + 5) * (NC[1] + 5) * (NC[2] + 5);
    double *zero = (double *)calloc(3 * (size_t)nodes, sizeof(double));
    double *rho = (double *)calloc((size_t)gnodes, sizeof(double));
    double *rho0 = (double *)malloc(sizeof(double) * (size_t)nodes);
    if (!zero || !rho || !rho0) die("fields", "out of memory");
    cic_rho(M, nm, rho);
    for (int i = 0, q = 0; i <= NC[0]; ++i)
        for (int j = 0; j <= NC[1]; ++j)
            for (int k = 0; k <= NC[2]; ++k, ++q)
                rho0[q] = rho[(((long)i + 2) * (NC[1] + 5) + j + 2) * (NC[2] + 5) + k + 2];
    const int lo[3] = {0, 0, 0};
    /* Tecplot record 0 (tecIO_saveFields / tecIO_saveCurrents): file map 0;
     * the setup meshes span [-2, n+2] (source/setup/main.c:76-80) */
    HOST(mh_filemap_save("binData", 0, NC, 1) < 0, "file map");
    HOST(mh_mesh_save_ex("binData/tec_E_000000(000).bin", NC, zero, 3, lo, NC, 2, "setup_E"), "tec_E");
    HOST(mh_mesh_save_ex("binData/tec_H_000000(000).bin", NC, zero, 3, lo, NC, 2, "setup_H"), "tec_H");
    HOST(mh_tec_write("binData", 0, 0.0, NC, 0), "tec record descriptor");
    HOST(mh_mesh_save_ex("binData/tec_J_000000(000).bin", NC, zero, 3, lo, NC, 2, "setup_J"), "tec_J");
    HOST(mh_mesh_save_ex("binData/tec_rho_000000(000).bin", NC, rho0, 1, lo, NC, 2, "setup_rho"),
         "tec_rho");
    HOST(mh_tec_count("binData", 1), "tecplot.N");
    /* checkpoint 0 (sysIO_save, plasma_save): file map 1 */
    HOST(mh_filemap_save("binData", 1, NC, 1) < 0, "file map");
    HOST(mh_mesh_save_ex("binData/sys_E_000000(000).bin", NC, zero, 3, lo, NC, 2, "setup_E"), "E record");
    HOST(mh_mesh_save_ex("binData/sys_H_000000(000).bin", NC, zero, 3, lo, NC, 2, "setup_H"), "H record");
    HOST(mh_sys_write("binData", 0, 0.0, NC, 1), "record descriptor");
    HOST(mh_markers_save("binData/plasma_000000_000.bin", M, nm, (int)sizeof(rec64_t)), "plasma record");
    free(rho);
    free(rho0);
    FILE *fp = fopen("binData/tracer.inf", "wt");      /* source/setup/main.c:245-250 */
    if (!fp) die("binData/tracer.inf", "cannot create");
    fprintf(fp, "%le \t\xcf\x84 (time step)\n%d  \tnumber of trajectories", D.tau, 0);
    fclose(fp);
    free(zero);
    free(M);
    deck_free(&dk);
    return 0;
}
