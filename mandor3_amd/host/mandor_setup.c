/*
 * mandor_setup.c -- writes record 0 of a run from a *.input deck, in the
 * binData formats setup.out writes (source/setup/main.c:155-256).
 *
 *   mandor_setup deck.input
 *
 * Tags, in the reference's parameter order:
 *   [units]              @ lambda [micron]                (tag_units.c:36)
 *   [mesh]               @ imax @ jmax @ kmax @ Lx @ Ly @ Lz [micron]
 *                        @ tau (negative: fraction of tau_max) (tag_mesh.c:40-106)
 *   [boundaryConditions] @ xMin @ xMax @ yMin @ yMax @ zMin @ zMax (tag_boundary.c:29)
 *   [plasma]             @ nx @ ny @ nz @ layers           (tag_plasma.c:46-161)
 *                        optional '> random' (RANMAR in-cell positions instead
 *                        of the (i + 0.5)/n lattice) and '> xmin' '> xmax'
 *                        [micron] (a slab along x instead of the whole box)
 *   [DF:uniform]         @ rho [n_c] @ q/M @ ux @ uy @ uz @ dux @ duy @ duz
 *                        u = <u> + (r - 0.5) du            (tag_DF_uniform.c:36-70)
 * and two tags of this build:
 *   [RANMAR]             @ ij @ kl   seeds of the generator (default 1802 9373);
 *                        the reference draws with unseeded rand()
 *   [DF:thermal]         @ rho [n_c] @ q/M @ ux @ uy @ uz @ vth:  Maxwellian
 *                        momenta (Box-Muller on RANMAR) about <u>
 *   [laser]              @ a0 @ waist [micron] @ xf @ yf @ zf [micron]
 *                        @ duration [fs, intensity FWHM] @ T0 [fs] @ phi0:
 *                        a y-polarised paraxial Gaussian beam written as the
 *                        soft-source map .EM_sources/laser/mirror_00
 *                        (the reference computes its map with Stratton-Chu
 *                        integrals, source/scpic/em_mirror.c:497-660)
 * Each [plasma] tag starts a species; the [DF:*] tag after it sets its
 * momenta and charge.  Marker ids run 1..N (tracer records of 64 bytes).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include "mandor_host.h"

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

static const mh_tag_t *find(const mh_cfg_t *c, const char *name, int from)
{
    for (int t = from; t < c->ntags; ++t)
        if (!strcmp(c->tags[t].name, name)) return c->tags + t;
    return NULL;
}

static int has_opt(const mh_tag_t *t, const char *word)
{
    for (int q = 0; q < t->nopt; ++q)
        if (!strcmp(t->opt[q], word)) return 1;
    return 0;
}

static int opt_double(const mh_tag_t *t, const char *key, double *v)
{
    for (int q = 0; q + 1 < t->nopt; ++q)
        if (!strcmp(t->opt[q], key)) {
            *v = atof(t->opt[q + 1]);
            return 1;
        }
    return 0;
}

static double gauss(mh_ranmar_t *r)
{
    double u[2];
    do mh_ranmar(r, u, 2);
    while (u[0] <= 0.0);
    return sqrt(-2.0 * log(u[0])) * cos(2.0 * M_PI * u[1]);
}

/* y-polarised paraxial Gaussian beam focused at f; records {y_cos, y_sin,
 * z_cos, z_sin} of the tangential incident field on the source planes for
 * j, k in [5, n-5] (E map at x = (I-1)h drives H at plane I; H map at
 * x = (I-1/2)h drives E at plane I-1). */
static void beam_map(const int n[3], const double h[3], double w0, const double f[3],
                     double *map)
{
    const int I = 5;
    const double k0 = 2 * M_PI, zr = M_PI * w0 * w0;
    const long nj = n[1] + 1 - 10, nk = n[2] + 1 - 10, nodes = nj * nk;
    for (int pass = 0; pass < 2; ++pass) {
        const double x = (pass == 0 ? I - 1.0 : I - 0.5) * h[0];
        const double xi = x - f[0];
        const double w = w0 * sqrt(1 + (xi / zr) * (xi / zr));
        const double invR = xi / (xi * xi + zr * zr);
        const double gouy = atan2(xi, zr);
        for (long j = 0; j < nj; ++j)
            for (long k = 0; k < nk; ++k) {
                const double y = (j + 5 - 0.5) * h[1], z = (k + 5) * h[2];
                const double r2 = (y - f[1]) * (y - f[1]) + (z - f[2]) * (z - f[2]);
                const double amp = (w0 / w) * exp(-r2 / (w * w));
                const double th = -k0 * r2 * invR / 2 + gouy - k0 * (x - I * h[0]);
                double *m = map + 4 * (pass * nodes + j * nk + k);
                const double c = amp * cos(th), s = -amp * sin(th);
                if (pass == 0) {
                    m[0] = c; m[1] = s; m[2] = 0; m[3] = 0;
                } else {
                    m[0] = c; m[1] = s; m[2] = -c; m[3] = -s;
                }
            }
    }
}

int main(int argc, char **argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: %s deck.input [create]\n", argv[0]);
        return EXIT_FAILURE;
    }
    mh_cfg_t c;
    HOST(mh_cfg_load(argv[1], &c), argv[1]);
    const mh_tag_t *tu = find(&c, "units", 0), *tm = find(&c, "mesh", 0),
                   *tb = find(&c, "boundaryConditions", 0), *tr = find(&c, "RANMAR", 0);
    if (!tu || !tm || !tb) die(argv[1], "[units], [mesh] and [boundaryConditions] are required");

    double lambda_um;
    HOST(mh_tag_double(tu, 0, &lambda_um), "[units]");
    if (lambda_um <= 0) die("[units]", "lambda must be positive");
    mh_units_t u;
    mh_units_from_lambda(lambda_um * 1e-4, &u);

    int n[3];
    double Lum[3], tau;
    for (int a = 0; a < 3; ++a) {
        HOST(mh_tag_int(tm, a, n + a), "[mesh]");
        HOST(mh_tag_double(tm, 3 + a, Lum + a), "[mesh]");
    }
    HOST(mh_tag_double(tm, 6, &tau), "[mesh]");
    mh_domain_t d;
    memset(&d, 0, sizeof(d));
    double inv2 = 0;
    for (int a = 0; a < 3; ++a) {
        if (n[a] < 1) die("[mesh]", "every axis needs at least one cell");
        d.L[a] = Lum[a] * u.micron;
        d.h[a] = d.L[a] / n[a];
        d.dmn_max[a] = n[a];
        inv2 += 1.0 / (d.h[a] * d.h[a]);
    }
    if (tau < 0) tau /= -sqrt(inv2);                      /* tag_mesh.c:59-63 */
    if (tau * sqrt(inv2) >= 0.99) die("[mesh]", "tau is too big (Courant >= 0.99)");
    d.tau = tau;
    int bc[6];
    for (int q = 0; q < 6; ++q) HOST(mh_tag_int(tb, q, bc + q), "[boundaryConditions]");
    for (int a = 0; a < 3; ++a) {                         /* xMin xMax yMin yMax zMin zMax */
        d.bc_min[a] = bc[2 * a];
        d.bc_max[a] = bc[2 * a + 1];
    }
    int ij = 1802, kl = 9373;
    if (tr) {
        HOST(mh_tag_int(tr, 0, &ij), "[RANMAR]");
        HOST(mh_tag_int(tr, 1, &kl), "[RANMAR]");
    }
    mh_ranmar_t rng;
    HOST(mh_ranmar_init(&rng, ij, kl), "[RANMAR]");

    mkdir("binData", 0755);
    mkdir("tmp", 0755);
    mkdir("output", 0755);
    mkdir(".EM_sources", 0755);
    mkdir(".EM_sources/laser", 0755);
    mkdir(".EM_sources/TFSF", 0755);

    /* species: every [plasma] tag and the [DF:*] tag that follows it */
    const double rho_crit = mh_rho_crit(&u);
    rec64_t *M = NULL;
    long nm = 0, cap = 0;
    for (int t = 0; t < c.ntags; ++t) {
        if (strcmp(c.tags[t].name, "plasma")) continue;
        const mh_tag_t *tp = c.tags + t;
        const mh_tag_t *tdf = t + 1 < c.ntags ? c.tags + t + 1 : NULL;
        if (!tdf || strncmp(tdf->name, "DF:", 3)) die("[plasma]", "must be followed by a [DF:*] tag");
        int sub[4];
        for (int q = 0; q < 4; ++q) HOST(mh_tag_int(tp, q, sub + q), "[plasma]");
        if (sub[0] < 1 || sub[1] < 1 || sub[2] < 1 || sub[3] < 1) die("[plasma]", "bad sampling");
        const int ppc = sub[0] * sub[1] * sub[2] * sub[3];
        const int random_pos = has_opt(tp, "random");
        double xlo = 0, xhi = d.L[0], v;
        if (opt_double(tp, "xmin", &v)) xlo = v * u.micron;
        if (opt_double(tp, "xmax", &v)) xhi = v * u.micron;
        double rho_nc, qm, um[3], spread[3] = {0, 0, 0}, vth = 0;
        HOST(mh_tag_double(tdf, 0, &rho_nc), tdf->name);
        HOST(mh_tag_double(tdf, 1, &qm), tdf->name);
        for (int a = 0; a < 3; ++a) HOST(mh_tag_double(tdf, 2 + a, um + a), tdf->name);
        const int thermal = !strcmp(tdf->name, "DF:thermal");
        if (thermal) {
            HOST(mh_tag_double(tdf, 5, &vth), tdf->name);
        } else if (!strcmp(tdf->name, "DF:uniform")) {
            for (int a = 0; a < 3; ++a) HOST(mh_tag_double(tdf, 5 + a, spread + a), tdf->name);
        } else {
            die(tdf->name, "supported: [DF:uniform], [DF:thermal]");
        }
        if (!(rho_nc * qm > 0)) die(tdf->name, "sign of rho differs from the sign of q/M");
        const float w = (float)(rho_nc * rho_crit / ppc);
        for (int I = 0; I < n[0]; ++I) {
            const double xc = (I + 0.5) * d.h[0];
            if (xc < xlo || xc >= xhi) continue;
            for (int J = 0; J < n[1]; ++J)
                for (int K = 0; K < n[2]; ++K)
                    for (int i = 0; i < sub[0]; ++i)
                        for (int j = 0; j < sub[1]; ++j)
                            for (int k = 0; k < sub[2]; ++k)
                                for (int l = 0; l < sub[3]; ++l) {
                                    if (nm == cap) {
                                        cap = cap ? 2 * cap : 1 << 16;
                                        M = (rec64_t *)realloc(M, sizeof(rec64_t) * cap);
                                        if (!M) die("markers", "out of memory");
                                    }
                                    rec64_t *m = M + nm++;
                                    double r[3];
                                    if (random_pos) {
                                        mh_ranmar(&rng, r, 3);
                                    } else {
                                        r[0] = (i + 0.5) / sub[0];
                                        r[1] = (j + 0.5) / sub[1];
                                        r[2] = (k + 0.5) / sub[2];
                                    }
                                    m->x = (I + r[0]) * d.h[0];
                                    m->y = (J + r[1]) * d.h[1];
                                    m->z = (K + r[2]) * d.h[2];
                                    if (thermal) {
                                        m->vx = um[0] + vth * gauss(&rng);
                                        m->vy = um[1] + vth * gauss(&rng);
                                        m->vz = um[2] + vth * gauss(&rng);
                                    } else {
                                        double q[3];
                                        mh_ranmar(&rng, q, 3);
                                        m->vx = um[0] + (q[0] - 0.5) * spread[0];
                                        m->vy = um[1] + (q[1] - 0.5) * spread[1];
                                        m->vz = um[2] + (q[2] - 0.5) * spread[2];
                                    }
                                    m->rho = w;
                                    m->qm = (float)qm;
                                    m->id = (int)nm;
                                    m->pad = 0;
                                }
        }
    }

    const mh_tag_t *tl = find(&c, "laser", 0);
    if (tl) {
        double p[8];
        for (int q = 0; q < 8; ++q) HOST(mh_tag_double(tl, q, p + q), "[laser]");
        if (n[1] < 11 || n[2] < 11) die("[laser]", "needs at least 11 cells along y and z");
        const long nodes = (long)(n[1] + 1 - 10) * (n[2] + 1 - 10);
        double *map = (double *)malloc(sizeof(double) * 8 * nodes);
        if (!map) die("[laser]", "out of memory");
        const double f[3] = {p[2] * u.micron, p[3] * u.micron, p[4] * u.micron};
        beam_map(n, d.h, p[1] * u.micron, f, map);
        const double focus_um[3] = {p[2], p[3], p[4]};
        HOST(mh_laser_save(".EM_sources/laser/mirror_00", &d, &u, -p[0], 0.9, focus_um, p[5], p[6],
                           p[7], 1, map),
             "laser source");
        free(map);
    }

    HOST(mh_domain_save("binData/domain.bin", &d), "domain.bin");
    HOST(mh_units_save("binData/units.bin", &u), "units.bin");
    const long nodes = (long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
    double *zero = (double *)calloc(3 * nodes, sizeof(double));
    if (!zero) die("fields", "out of memory");
    const int lo[3] = {0, 0, 0};
    HOST(mh_mesh_save("binData/sys_E_000000(000).bin", n, zero, lo, n, "E"), "E record");
    HOST(mh_mesh_save("binData/sys_H_000000(000).bin", n, zero, lo, n, "H"), "H record");
    HOST(mh_markers_save("binData/plasma_000000_000.bin", M, nm, (int)sizeof(rec64_t)), "plasma record");
    HOST(mh_sys_write("binData", 0, 0.0, n, 1), "record descriptor");
    printf("mandor_setup: %d x %d x %d cells, h = (%g, %g, %g), tau = %g, %ld markers\n", n[0], n[1],
           n[2], d.h[0], d.h[1], d.h[2], d.tau, nm);
    free(zero);
    free(M);
    mh_cfg_free(&c);
    return 0;
}
