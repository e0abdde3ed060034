/*
 * mandor_core.c -- the core.out main loop on one MI355X.
 *
 *   mandor_core [--fp32] [--device N]     (run in the directory of a setup)
 *
 * Reads run_mandor.cfg, binData/domain.bin, binData/units.bin, the start
 * record (record 0, or the last one with the continue flag), the laser
 * sources in .EM_sources/laser, the TFSF recordings in .EM_sources/TFSF
 * (play modes) and the hard spot of .gaussSpot.cfg, then
 * runs the time step through the C ABI of include/mpic.h, writing
 * checkpoints, the energy history and honouring tmp/stop.flag as
 * source/core/main.c:195-490 does for one rank.
 * Tecplot and spectral dumps are not produced (a warning says so).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "mandor_host.h"
#include "mpic.h"

static double now(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

static void die(const char *what, const char *msg)
{
    fprintf(stderr, "mandor_core: %s: %s\n", what, msg);
    exit(EXIT_FAILURE);
}

#define HOST(x, what)                               \
    do {                                            \
        if ((x) != 0) die(what, mh_error());        \
    } while (0)
#define GPU(x, what)                                \
    do {                                            \
        if ((x) != MPIC_OK) die(what, mpic_last_error()); \
    } while (0)

typedef struct {
    mpic_run_t *run;
    int         n[3];
    long        nodes;
    double     *E, *H;
    void       *rec;
    long        cap;
    int         recsize;
} state_t;

/* plasma_save + sysIO_save (source/core/main.c:444-462). */
static void checkpoint(state_t *s, int rec, double time)
{
    char p[256];
    long n = s->cap;
    GPU(mpic_markers_save(s->run, s->rec, &n, s->recsize), "mpic_markers_save");
    snprintf(p, sizeof(p), "binData/plasma_%06d_%03d.bin", rec, 0);
    HOST(mh_markers_save(p, s->rec, n, s->recsize), "plasma record");
    GPU(mpic_get_fields(s->run, s->E, s->H), "mpic_get_fields");
    const int lo[3] = {0, 0, 0};
    snprintf(p, sizeof(p), "binData/sys_E_%06d(%03d).bin", rec, 0);
    HOST(mh_mesh_save(p, s->n, s->E, lo, s->n, "E"), "E record");
    snprintf(p, sizeof(p), "binData/sys_H_%06d(%03d).bin", rec, 0);
    HOST(mh_mesh_save(p, s->n, s->H, lo, s->n, "H"), "H record");
    HOST(mh_sys_write("binData", rec, time, s->n, 0), "checkpoint descriptor");
}

int main(int argc, char **argv)
{
    int precision = MPIC_FP64, device = 0;
    for (int a = 1; a < argc; ++a) {
        if (!strcmp(argv[a], "--fp32")) {
            precision = MPIC_FP32_SPLIT;
        } else if (!strcmp(argv[a], "--device") && a + 1 < argc) {
            device = atoi(argv[++a]);
        } else {
            fprintf(stderr, "usage: %s [--fp32] [--device N]\n", argv[0]);
            return EXIT_FAILURE;
        }
    }
    mh_runcfg_t rc;
    HOST(mh_runcfg_load("run_mandor.cfg", &rc), "run_mandor.cfg");
    HOST(mh_stopflag_default("tmp/stop.flag"), "tmp/stop.flag");
    if (rc.chpoint_tecplot > 0 || rc.chpoint_spectr > 0)
        fprintf(stderr, "mandor_core: tecplot / spectral dumps are not produced by this driver\n");

    mh_domain_t dom;
    mh_units_t units;
    HOST(mh_domain_load("binData/domain.bin", &dom), "domain.bin");
    HOST(mh_units_load("binData/units.bin", &units), "units.bin");

    /* start record (source/core/main.c:238-257) */
    const int nrec = mh_sys_records("binData");
    if (nrec < 1) die("binData/sys.N", "initial checkpoint is not prepared");
    int point = nrec - 1, next;
    int cont = rc.continue_run && point != 0;
    if (cont) {
        next = nrec;
    } else {
        point = 0;
        next = 1;
    }
    double time;
    int cpus, fmap;
    HOST(mh_sys_info("binData", point, &time, &cpus, &fmap), "record descriptor");

    state_t s;
    memset(&s, 0, sizeof(s));
    for (int a = 0; a < 3; ++a) s.n[a] = dom.dmn_max[a] - dom.dmn_min[a];
    s.nodes = (long)(s.n[0] + 1) * (s.n[1] + 1) * (s.n[2] + 1);
    s.E = (double *)calloc(3 * s.nodes, sizeof(double));
    s.H = (double *)calloc(3 * s.nodes, sizeof(double));
    if (!s.E || !s.H) die("fields", "out of memory");

    /* every rank's piece of the record (file map of the record) */
    void *rec = NULL;
    long nm = 0;
    int recsize = 64;
    for (int cpu = 0; cpu < cpus; ++cpu) {
        char p[256];
        snprintf(p, sizeof(p), "binData/sys_E_%06d(%03d).bin", point, cpu);
        HOST(mh_mesh_load(p, s.n, s.E), "E record");
        snprintf(p, sizeof(p), "binData/sys_H_%06d(%03d).bin", point, cpu);
        HOST(mh_mesh_load(p, s.n, s.H), "H record");
        snprintf(p, sizeof(p), "binData/plasma_%06d_%03d.bin", point, cpu);
        void *r1;
        long n1;
        int rs1;
        HOST(mh_markers_load(p, &r1, &n1, &rs1), "plasma record");
        if (cpu && rs1 != recsize && n1) die(p, "mixed marker record sizes");
        if (n1) recsize = rs1;
        void *grown = realloc(rec, (size_t)(nm + n1) * recsize + 1);
        if (!grown) die("markers", "out of memory");
        rec = grown;
        memcpy((char *)rec + (size_t)nm * recsize, r1, (size_t)n1 * recsize);
        nm += n1;
        free(r1);
    }

    mpic_config_t cfg;
    memset(&cfg, 0, sizeof(cfg));
    for (int a = 0; a < 3; ++a) {
        cfg.n[a] = s.n[a];
        cfg.h[a] = dom.h[a];
        cfg.L[a] = dom.L[a];
        cfg.bc[a] = dom.bc_min[a];
        cfg.bc[3 + a] = dom.bc_max[a];
    }
    cfg.tau = dom.tau;
    cfg.time = time;
    cfg.precision = precision;
    cfg.max_markers = nm > 0 ? nm : 1;
    cfg.device = device;
    GPU(mpic_create(&cfg, &s.run), "mpic_create");
    GPU(mpic_set_fields(s.run, s.E, s.H), "mpic_set_fields");
    GPU(mpic_markers_load(s.run, rec, nm, recsize), "mpic_markers_load");
    s.rec = rec;
    s.cap = nm;
    s.recsize = recsize;
    for (int k = 0; k < 100; ++k) {
        char p[256];
        snprintf(p, sizeof(p), ".EM_sources/laser/mirror_%02d", k);
        FILE *fp = fopen(p, "rb");
        if (!fp) break;
        fclose(fp);
        mh_laser_t l;
        HOST(mh_laser_load(p, &dom, &units, &l), p);
        GPU(mpic_add_laser(s.run, l.A, l.T0, l.delay, l.dt, l.phi0, l.forward, l.plane, l.map),
            "mpic_add_laser");
        mh_laser_free(&l);
    }
    FILE *spot = fopen(".gaussSpot.cfg", "rt");       /* source/core/em_gaussSpot.c:42-55 */
    if (spot) {
        fclose(spot);
        mh_cfg_t sc;
        HOST(mh_cfg_load(".gaussSpot.cfg", &sc), ".gaussSpot.cfg");
        double p[11];
        for (int q = 0; q < 11; ++q) HOST(mh_tag_double(&sc.tags[0], q, p + q), ".gaussSpot.cfg");
        mh_cfg_free(&sc);
        GPU(mpic_add_gauss_spot(s.run, p), "mpic_add_gauss_spot");
    }
    mh_tfsf_t tfsf;                                    /* source/core/em_TFSF.c:474-545 */
    HOST(mh_tfsf_load(".EM_sources/TFSF", &dom, &tfsf), ".EM_sources/TFSF");
    if (tfsf.mode == 0 || tfsf.mode == 2)
        GPU(mpic_add_tfsf(s.run, tfsf.reg, tfsf.mode, tfsf.nframes,
                          (const double *const *)tfsf.dataE, (const double *const *)tfsf.dataH),
            "mpic_add_tfsf");
    else if (tfsf.mode == 1)
        fprintf(stderr, "mandor_core: TFSF record mode is not produced by this driver; "
                        "the interface stays off\n");
    mh_tfsf_free(&tfsf);
    GPU(mpic_init(s.run), "mpic_init");
    printf("mandor_core: %s form, %d x %d x %d cells, %ld markers, start record %d, t = %.6f\n",
           precision == MPIC_FP64 ? "FP64 plain" : "FP32 split", s.n[0], s.n[1], s.n[2], nm, point,
           time);

    /* energy history: fresh unless continuing (source/core/main.c:259) */
    int fresh = !cont;
    double *pts = (double *)malloc(sizeof(double) * 6 * (rc.total_steps + 1));
    if (!pts) die("energy history", "out of memory");
    long npts = 0;

    const double t_start = now();
    double last_check = t_start, dT = 0, save_time = 0, saved_at = -1;
    int first_check = 1, stop = 0;
    for (int t = 1; t <= rc.total_steps && !stop; ++t) {
        int stop_now = 0;
        if (rc.chpoint_stop > 0 && t % rc.chpoint_stop == 0) {
            int force = 0, limit = 0;
            HOST(mh_stopflag_read("tmp/stop.flag", &force, &limit), "tmp/stop.flag");
            const double el = now() - t_start;
            if (force || el + 1.2 * dT + 1.5 * save_time >= limit) stop_now = 1;
            const double tn = now();
            if (!first_check) dT = 0.8 * (tn - last_check) + 0.2 * dT;
            first_check = 0;
            last_check = tn;
        }
        GPU(mpic_step(s.run, pts + 6 * npts), "mpic_step");
        if (!isfinite(pts[6 * npts + 1] + pts[6 * npts + 2] + pts[6 * npts + 3] +
                      pts[6 * npts + 4] + pts[6 * npts + 5]))
            die("energy diagnostic", "nan detected");
        ++npts;
        time = mpic_time(s.run);
        if (t % rc.chpoint_full == 0) {
            const double w0 = now();
            checkpoint(&s, next++, time);
            HOST(mh_diag_append("binData/diag.dat", pts, npts, dom.tau, fresh), "diag.dat");
            fresh = 0;
            npts = 0;
            saved_at = time;
            save_time = now() - w0;
        }
        stop = stop_now;
    }
    GPU(mpic_sync(s.run), "mpic_sync");
    printf("main: main loop have taken %f sec.\n", now() - t_start);
    if (saved_at != time) checkpoint(&s, next++, time);
    HOST(mh_diag_append("binData/diag.dat", pts, npts, dom.tau, fresh), "diag.dat");
    remove("tmp/stop.flag");
    mpic_destroy(s.run);
    free(pts);
    free(rec);
    free(s.E);
    free(s.H);
    return 0;
}
