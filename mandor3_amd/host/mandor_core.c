/*
 * mandor_core.c -- the core.out main loop on one MI355X.
 *
 *   mandor_core [--fp32] [--device N]     (run in the directory of a setup)
 *   mandor_core --tfsf-defrag             (TFSFdefrag: joins recorded TFSF fragments)
 *
 * Reads run_mandor.cfg, binData/domain.bin, binData/units.bin, the start
 * record (record 0, or the last one with the continue flag), the laser
 * sources in .EM_sources/laser, the TFSF recordings in .EM_sources/TFSF
 * (play modes) and the hard spot of .gaussSpot.cfg, then
 * runs the time step through the C ABI of include/mpic.h, writing
 * checkpoints, Tecplot records (every chPoint_tecplot steps: E and H after
 * the push, J and the charge density at the end of the step), probe samples
 * (run_probes.cfg, every step after the first H half step) and the energy
 * history, and honouring tmp/stop.flag, as source/core/main.c:195-490 does
 * for one rank.  A TFSF interface in record mode stores its faces on the
 * device every step and writes the fragment files at the end of the run.
 * Spectral dumps are not produced (a warning says so).
 */
#define _GNU_SOURCE
#include <dirent.h>
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
    int         fmap;         /* next file map id (source/IO_fileMap.c:40-68) */
    int         tec;          /* next Tecplot record */
    double     *node, *rho, *rho0;
    int         nprobes;
    int        *probe;        /* nodes, 3 ints each */
    FILE      **pf;           /* binData/probe_%d_%d_%d.bin */
    double     *sample;
} state_t;

/* plasma_save + sysIO_save (source/core/main.c:444-462). */
static void checkpoint(state_t *s, int rec, double time)
{
    const int fmap = mh_filemap_save("binData", s->fmap++, s->n, 0);
    if (fmap < 0) die("file map", mh_error());
    char p[256];
    long n = s->cap;
    GPU(mpic_markers_save(s->run, s->rec, &n, s->recsize), "mpic_markers_save");
    snprintf(p, sizeof(p), "binData/plasma_%06d_%03d.bin", rec, 0);
    HOST(mh_markers_save(p, s->rec, n, s->recsize), "plasma record");
    GPU(mpic_get_fields(s->run, s->E, s->H), "mpic_get_fields");
    const int lo[3] = {0, 0, 0};
    snprintf(p, sizeof(p), "binData/sys_E_%06d(%03d).bin", rec, 0);
    HOST(mh_mesh_save(p, s->n, s->E, lo, s->n, "vec_E"), "E record");
    snprintf(p, sizeof(p), "binData/sys_H_%06d(%03d).bin", rec, 0);
    HOST(mh_mesh_save(p, s->n, s->H, lo, s->n, "vec_H"), "H record");
    HOST(mh_sys_write("binData", rec, time, s->n, fmap), "checkpoint descriptor");
}

/* tecIO_saveFields (source/IO_tecplot.c:218-245): node values of E and H
 * as they are after the push, file map, descriptor. */
static void tec_fields(state_t *s, double time)
{
    char p[256];
    const int lo[3] = {0, 0, 0};
    const int fmap = mh_filemap_save("binData", s->fmap++, s->n, 0);
    if (fmap < 0) die("file map", mh_error());
    GPU(mpic_node_fields(s->run, 0, s->node), "mpic_node_fields");
    snprintf(p, sizeof(p), "binData/tec_E_%06d(%03d).bin", s->tec, 0);
    HOST(mh_mesh_save_ex(p, s->n, s->node, 3, lo, s->n, 1, "vec_E"), "tec_E");
    GPU(mpic_node_fields(s->run, 1, s->node), "mpic_node_fields");
    snprintf(p, sizeof(p), "binData/tec_H_%06d(%03d).bin", s->tec, 0);
    HOST(mh_mesh_save_ex(p, s->n, s->node, 3, lo, s->n, 1, "vec_H"), "tec_H");
    HOST(mh_tec_write("binData", s->tec, time, s->n, fmap), "tec descriptor");
}

/* tecIO_saveCurrents (:251-264): node values of J after the fold and the
 * charge density taken after the push, then tecplot.N. */
static void tec_currents(state_t *s)
{
    char p[256];
    const int lo[3] = {0, 0, 0};
    GPU(mpic_node_fields(s->run, 2, s->node), "mpic_node_fields");
    snprintf(p, sizeof(p), "binData/tec_J_%06d(%03d).bin", s->tec, 0);
    HOST(mh_mesh_save_ex(p, s->n, s->node, 3, lo, s->n, 2, "vec_J"), "tec_J");
    for (int i = 0, q = 0; i <= s->n[0]; ++i)
        for (int j = 0; j <= s->n[1]; ++j)
            for (int k = 0; k <= s->n[2]; ++k, ++q)
                s->rho0[q] = s->rho[(((long)i + 2) * (s->n[1] + 5) + j + 2) * (s->n[2] + 5) + k + 2];
    snprintf(p, sizeof(p), "binData/tec_rho_%06d(%03d).bin", s->tec, 0);
    HOST(mh_mesh_save_ex(p, s->n, s->rho0, 1, lo, s->n, 2, "scal_rho"), "tec_rho");
    ++s->tec;
    HOST(mh_tec_count("binData", s->tec), "tecplot.N");
}

/* probe_allocate / probe_touch (source/core/diag_probe.c:47-100,171-357):
 * old probe files removed unless continuing, globals and locations written,
 * one file per probe node opened (header on creation, append otherwise). */
static void probes_open(state_t *s, const mh_domain_t *d, const mh_units_t *u, int cont)
{
    DIR *dir = cont ? NULL : opendir("binData");
    if (dir) {
        struct dirent *e;
        char p[512];
        while ((e = readdir(dir))) {
            const size_t len = strlen(e->d_name);
            if (!strncmp(e->d_name, "probe_", 6) && len > 4 && !strcmp(e->d_name + len - 4, ".bin")) {
                snprintf(p, sizeof(p), "binData/%s", e->d_name);
                remove(p);
            }
        }
        closedir(dir);
    }
    HOST(mh_probes_parse("run_probes.cfg", d, u, &s->probe, &s->nprobes), "run_probes.cfg");
    for (int p = 0; p < s->nprobes; ++p)
        for (int a = 0; a < 3; ++a)
            if (s->probe[3 * p + a] < 0 || s->probe[3 * p + a] >= s->n[a])
                die("run_probes.cfg", "a probe lies outside the domain");
    const double mic = u->micron;
    FILE *fp = fopen("output/probe_locations.dat", "wt");
    if (fp) {
        fprintf(fp, "variables = \"x [<greek>m</greek>m]\", \"y [<greek>m</greek>m]\", "
                    "\"z [<greek>m</greek>m]\"\nzone t=\"domain\", f= point\n");
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k)
                    fprintf(fp, "%e %e %e\n", (d->dmn_min[0] * i + d->dmn_max[0] * (1 - i)) * d->h[0] / mic,
                            (d->dmn_min[1] * j + d->dmn_max[1] * (1 - j)) * d->h[1] / mic,
                            (d->dmn_min[2] * k + d->dmn_max[2] * (1 - k)) * d->h[2] / mic);
        fprintf(fp, "zone t=\"probes\", f= point\n");
        for (int p = 0; p < s->nprobes; ++p)
            fprintf(fp, "%e %e %e\n", s->probe[3 * p] * d->h[0] / mic, s->probe[3 * p + 1] * d->h[1] / mic,
                    s->probe[3 * p + 2] * d->h[2] / mic);
        fclose(fp);
    }
    fp = fopen("binData/probe_globals.bin", "wb");
    if (!fp) die("binData/probe_globals.bin", "cannot create");
    const double glob[10] = {d->tau, d->h[0], d->h[1], d->h[2],
                             d->dmn_min[0] * d->h[0], d->dmn_min[1] * d->h[1], d->dmn_min[2] * d->h[2],
                             d->dmn_max[0] * d->h[0], d->dmn_max[1] * d->h[1], d->dmn_max[2] * d->h[2]};
    if (fwrite(glob, sizeof(double), 10, fp) != 10) die("binData/probe_globals.bin", "write error");
    fclose(fp);
    if (!s->nprobes) return;
    s->pf = (FILE **)calloc((size_t)s->nprobes, sizeof(FILE *));
    s->sample = (double *)malloc(sizeof(double) * 6 * (size_t)s->nprobes);
    if (!s->pf || !s->sample) die("probes", "out of memory");
    for (int p = 0; p < s->nprobes; ++p) {
        char name[256];
        const int *q = s->probe + 3 * p;
        snprintf(name, sizeof(name), "binData/probe_%d_%d_%d.bin", q[0], q[1], q[2]);
        FILE *t = fopen(name, "rb");
        if (t) {
            fclose(t);
            s->pf[p] = fopen(name, "ab");
        } else {
            s->pf[p] = fopen(name, "wb");
            if (s->pf[p]) {
                struct {
                    char    magic[8];
                    int32_t i, j, k;
                } h;
                memset(&h, 0, sizeof(h));
                memcpy(h.magic, "probes", 6);
                h.i = q[0];
                h.j = q[1];
                h.k = q[2];
                if (fwrite(&h, sizeof(h), 1, s->pf[p]) != 1) die(name, "write error");
            }
        }
        if (!s->pf[p]) die(name, "cannot open");
    }
    GPU(mpic_set_probes(s->run, s->probe, s->nprobes), "mpic_set_probes");
}

/* One step through the stage entry points, with the probe samples after
 * the first H half step and a Tecplot record when 'tec' is set
 * (source/core/main.c:365-435). */
static void step_stages(state_t *s, double *pt, int tec)
{
    const double t0 = mpic_time(s->run);
    double WE, WM;
    GPU(mpic_em_h_half(s->run), "mpic_em_h_half");
    GPU(mpic_em_energy(s->run, &WE, &WM), "mpic_em_energy");
    if (s->nprobes) {
        GPU(mpic_probe_sample(s->run, s->sample), "mpic_probe_sample");
        for (int p = 0; p < s->nprobes; ++p) {
            double rec[7] = {t0};
            memcpy(rec + 1, s->sample + 6 * p, sizeof(double) * 6);
            if (fwrite(rec, sizeof(double), 7, s->pf[p]) != 7) die("probe file", "write error");
        }
    }
    GPU(mpic_push_deposit(s->run), "mpic_push_deposit");
    GPU(mpic_temperature(s->run, pt + 3, pt + 4, pt + 5), "mpic_temperature");
    if (tec) {
        GPU(mpic_rho(s->run, s->rho), "mpic_rho");
        tec_fields(s, t0);
    }
    GPU(mpic_em_h_step(s->run), "mpic_em_h_step");
    GPU(mpic_em_e_start(s->run), "mpic_em_e_start");
    GPU(mpic_j_finish(s->run), "mpic_j_finish");
    GPU(mpic_em_e_finish(s->run), "mpic_em_e_finish");
    if (tec) tec_currents(s);
    pt[0] = t0;
    pt[1] = WE;
    pt[2] = WM;
}

int main(int argc, char **argv)
{
    int precision = MPIC_FP64, device = 0;
    double vsp_weight = 1.0;
    for (int a = 1; a < argc; ++a) {
        if (!strcmp(argv[a], "--fp32")) {
            precision = MPIC_FP32_SPLIT;
        } else if (!strncmp(argv[a], "--VSP:weight(", 13)) {
            /* core.out's option (source/import.pl) and its clamp
             * (source/core/main.c:316-320) */
            vsp_weight = atof(argv[a] + 13);
            vsp_weight = (vsp_weight > 0.5 && vsp_weight < 1.00001) ? vsp_weight : 1.00;
        } else if (!strcmp(argv[a], "--device") && a + 1 < argc) {
            device = atoi(argv[++a]);
        } else if (!strcmp(argv[a], "--tfsf-defrag")) {
            HOST(mh_tfsf_defrag(".EM_sources/TFSF"), "TFSF defragmentation");
            return 0;
        } else {
            fprintf(stderr, "usage: %s [--fp32] [--device N] [--VSP:weight(w)] | --tfsf-defrag\n",
                    argv[0]);
            return EXIT_FAILURE;
        }
    }
    mh_runcfg_t rc;
    HOST(mh_runcfg_load("run_mandor.cfg", &rc), "run_mandor.cfg");
    HOST(mh_stopflag_default("tmp/stop.flag"), "tmp/stop.flag");
    if (rc.chpoint_spectr > 0)
        fprintf(stderr, "mandor_core: spectral dumps are not produced by this driver\n");

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
    s.node = (double *)calloc(3 * s.nodes, sizeof(double));
    s.rho0 = (double *)calloc(s.nodes, sizeof(double));
    s.rho = (double *)calloc((size_t)(s.n[0] + 5) * (s.n[1] + 5) * (s.n[2] + 5), sizeof(double));
    if (!s.E || !s.H || !s.node || !s.rho0 || !s.rho) die("fields", "out of memory");
    /* file map and Tecplot record counters (source/core/main.c:245-257) */
    s.fmap = mh_filemap_next("binData", cont);
    if (s.fmap < 0) die(cont ? "binData/fileMap.N" : "binData/fileMap_start.txt", "missing or unreadable");
    s.tec = cont ? mh_tec_records("binData") : 1;
    if (s.tec < 1) s.tec = 1;

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
    GPU(mpic_set_vsp(s.run, vsp_weight), "mpic_set_vsp");
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
    const int recording = tfsf.mode == 1;
    int rec_reg[6];
    memcpy(rec_reg, tfsf.reg, sizeof(rec_reg));
    if (recording) {                                   /* TFSF_turnRec, source/core/em_TFSF.c:273-330 */
        GPU(mpic_tfsf_record(s.run, rec_reg, rc.total_steps > 0 ? rc.total_steps : 1), "mpic_tfsf_record");
        HOST(mh_tfsf_record_begin(".EM_sources/TFSF", rec_reg, &dom), "TFSF record");
    }
    mh_tfsf_free(&tfsf);
    GPU(mpic_init(s.run), "mpic_init");
    probes_open(&s, &dom, &units, cont);
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
        const int tec = rc.chpoint_tecplot > 0 && t % rc.chpoint_tecplot == 0;
        if (tec || s.nprobes)
            step_stages(&s, pts + 6 * npts, tec);
        else
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
    if (recording) {                                   /* TFSF_shutDown, :866-890 */
        int nf = 0;
        GPU(mpic_tfsf_recorded(s.run, &nf, NULL, NULL), "mpic_tfsf_recorded");
        long vol[24], total = 0;
        GPU(mpic_tfsf_recorded(s.run, &nf, vol, NULL), "mpic_tfsf_recorded");
        for (int q = 0; q < 24; ++q) total += vol[q] * nf;
        double *frames = (double *)malloc(sizeof(double) * (size_t)(total > 0 ? total : 1));
        if (!frames) die("TFSF record", "out of memory");
        GPU(mpic_tfsf_recorded(s.run, &nf, vol, frames), "mpic_tfsf_recorded");
        HOST(mh_tfsf_record_save(".EM_sources/TFSF", rec_reg, nf, frames), "TFSF record");
        printf("TFSF_shutDown: %d frames written.\n", nf);
        free(frames);
    }
    for (int p = 0; p < s.nprobes; ++p) fclose(s.pf[p]);
    mpic_destroy(s.run);
    free(s.pf);
    free(s.sample);
    free(s.probe);
    free(s.node);
    free(s.rho);
    free(s.rho0);
    free(pts);
    free(rec);
    free(s.E);
    free(s.H);
    return 0;
}
