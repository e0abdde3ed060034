/*
 * mandor_host.c -- configuration and binData record formats, energy
 * history, stop flag, laser source files and RANMAR for the C drivers.
 * Each function cites the reference file:line whose format it reads or
 * writes (cvarin/mandor3).
 */
#define _GNU_SOURCE
#include "mandor_host.h"

#include <errno.h>
#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

static char g_err[512];

static int fail(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return -1;
}

const char *mh_error(void) { return g_err; }

/* ---- '@' parameter files (source/misc_cfgReader.c:20-111) -------------- */

static mh_tag_t *new_tag(mh_cfg_t *c, const char *name)
{
    mh_tag_t *t = (mh_tag_t *)realloc(c->tags, sizeof(mh_tag_t) * (c->ntags + 1));
    if (!t) return NULL;
    c->tags = t;
    t = c->tags + c->ntags++;
    memset(t, 0, sizeof(*t));
    snprintf(t->name, sizeof(t->name), "%s", name);
    return t;
}

static int push_value(char ***list, int *n, const char *word)
{
    char **l = (char **)realloc(*list, sizeof(char *) * (*n + 1));
    if (!l) return -1;
    *list = l;
    l[*n] = strdup(word);
    if (!l[*n]) return -1;
    ++*n;
    return 0;
}

int mh_cfg_load(const char *path, mh_cfg_t *cfg)
{
    memset(cfg, 0, sizeof(*cfg));
    FILE *fp = fopen(path, "rt");
    if (!fp) return fail("cannot open '%s': %s", path, strerror(errno));
    mh_tag_t *cur = new_tag(cfg, "");
    char line[4096];
    while (cur && fgets(line, sizeof(line), fp)) {
        if (line[0] == '[') {
            char *e = strchr(line, ']');
            if (!e) continue;
            *e = 0;
            cur = new_tag(cfg, line + 1);
        } else if (line[0] == '@' || line[0] == '>') {
            char word[512];
            if (sscanf(line + 1, "%511s", word) != 1) {
                fclose(fp);
                return fail("'%s': empty parameter line", path);
            }
            if (line[0] == '@') {
                if (push_value(&cur->req, &cur->nreq, word)) break;
            } else if (push_value(&cur->opt, &cur->nopt, word)) {
                break;
            }
        }
    }
    fclose(fp);
    if (!cur) return fail("out of memory");
    return 0;
}

void mh_cfg_free(mh_cfg_t *cfg)
{
    for (int t = 0; t < cfg->ntags; ++t) {
        for (int q = 0; q < cfg->tags[t].nreq; ++q) free(cfg->tags[t].req[q]);
        for (int q = 0; q < cfg->tags[t].nopt; ++q) free(cfg->tags[t].opt[q]);
        free(cfg->tags[t].req);
        free(cfg->tags[t].opt);
    }
    free(cfg->tags);
    memset(cfg, 0, sizeof(*cfg));
}

int mh_tag_int(const mh_tag_t *t, int idx, int *v)
{
    if (idx < 0 || idx >= t->nreq) return fail("[%s]: parameter %d missing", t->name, idx + 1);
    char *end;
    long x = strtol(t->req[idx], &end, 10);
    if (*end) return fail("[%s]: parameter %d '%s' is not an integer", t->name, idx + 1, t->req[idx]);
    *v = (int)x;
    return 0;
}

int mh_tag_double(const mh_tag_t *t, int idx, double *v)
{
    if (idx < 0 || idx >= t->nreq) return fail("[%s]: parameter %d missing", t->name, idx + 1);
    char *end;
    double x = strtod(t->req[idx], &end);
    if (*end) return fail("[%s]: parameter %d '%s' is not a number", t->name, idx + 1, t->req[idx]);
    *v = x;
    return 0;
}

/* run_mandor.cfg, source/core/main.c:205-228. */
int mh_runcfg_load(const char *path, mh_runcfg_t *rc)
{
    mh_cfg_t c;
    if (mh_cfg_load(path, &c)) return -1;
    int v[6], rcode = 0;
    for (int q = 0; q < 6 && !rcode; ++q) rcode = mh_tag_int(&c.tags[0], q, v + q);
    mh_cfg_free(&c);
    if (rcode) return -1;
    rc->total_steps = v[0];
    rc->chpoint_full = v[1];
    rc->chpoint_tecplot = v[2];
    rc->continue_run = (v[3] == 1);
    rc->chpoint_stop = v[4];
    rc->chpoint_spectr = v[5];
    if (v[0] == 0 || v[1] == 0 || v[2] == 0 || v[4] == 0 || v[5] == 0)
        return fail("'%s' contains zero step(s)", path);
    return 0;
}

/* ---- domain.bin, source/misc_parameters.c:98-131 ----------------------- */

int mh_domain_load(const char *path, mh_domain_t *d)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail("cannot open '%s'", path);
    double v[7];
    int i[12];
    const int ok = fread(v, sizeof(double), 7, fp) == 7 && fread(i, sizeof(int), 12, fp) == 12;
    fclose(fp);
    if (!ok) return fail("'%s': short file", path);
    for (int a = 0; a < 3; ++a) {
        d->h[a] = v[a];
        d->L[a] = v[3 + a];
        d->dmn_min[a] = i[a];
        d->dmn_max[a] = i[3 + a];
        d->bc_min[a] = i[6 + a];
        d->bc_max[a] = i[9 + a];
    }
    d->tau = v[6];
    return 0;
}

int mh_domain_save(const char *path, const mh_domain_t *d)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail("cannot create '%s'", path);
    const double v[7] = {d->h[0], d->h[1], d->h[2], d->L[0], d->L[1], d->L[2], d->tau};
    int i[12];
    for (int a = 0; a < 3; ++a) {
        i[a] = d->dmn_min[a];
        i[3 + a] = d->dmn_max[a];
        i[6 + a] = d->bc_min[a];
        i[9 + a] = d->bc_max[a];
    }
    const int ok = fwrite(v, sizeof(double), 7, fp) == 7 && fwrite(i, sizeof(int), 12, fp) == 12;
    return (fclose(fp) == 0 && ok) ? 0 : fail("write error on '%s'", path);
}

/* ---- units.bin, source/misc_units.c:39-118 ----------------------------- */

#define CGS_C 2.9979e+10
#define CGS_E 4.8032e-10
#define CGS_M 9.1094e-28

void mh_units_from_lambda(double lambda, mh_units_t *u)
{
    const double pi = M_PI;
    u->lambda = lambda;
    u->v = CGS_C;
    u->t = lambda / u->v;
    u->omega = 2 * pi / u->t;
    u->A = CGS_M * CGS_C * u->omega / CGS_E;
    u->E = u->A / (2 * pi);
    u->n = CGS_M * u->omega * u->omega / (16 * pi * pi * pi * CGS_E * CGS_E);
    u->rho = CGS_E * u->n;
    u->micron = 1.0e-4 / lambda;
    u->femtosecond = 1.0e-15 / u->t;
    u->wpe = 2 * CGS_E * sqrt(pi * u->n / CGS_M) * u->t;
    u->Tpe = 2 * pi / u->wpe;
    u->ncr = u->omega * u->omega * CGS_M / (4 * pi * CGS_E * CGS_E);
    u->S = -1;
    u->debye = -1;
}

double mh_rho_crit(const mh_units_t *u) { return CGS_E * u->ncr / u->rho; }

int mh_units_load(const char *path, mh_units_t *u)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail("cannot open '%s'", path);
    double v[15];
    const int ok = fread(v, sizeof(double), 15, fp) == 15;
    fclose(fp);
    if (!ok) return fail("'%s': short file", path);
    u->lambda = v[0]; u->v = v[1]; u->t = v[2]; u->omega = v[3]; u->E = v[4];
    u->S = v[5]; u->A = v[6]; u->n = v[7]; u->rho = v[8]; u->ncr = v[9];
    u->debye = v[10]; u->micron = v[11]; u->femtosecond = v[12]; u->wpe = v[13]; u->Tpe = v[14];
    return 0;
}

int mh_units_save(const char *path, const mh_units_t *u)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail("cannot create '%s'", path);
    const double v[15] = {u->lambda, u->v, u->t, u->omega, u->E, u->S, u->A, u->n,
                          u->rho, u->ncr, u->debye, u->micron, u->femtosecond, u->wpe, u->Tpe};
    const int ok = fwrite(v, sizeof(double), 15, fp) == 15;
    return (fclose(fp) == 0 && ok) ? 0 : fail("write error on '%s'", path);
}

/* ---- mesh records, source/type_mesh.c:285-357 -------------------------- */

/* The reference's mesh_t as laid out on disk on x86-64 (88 bytes). */
typedef struct {
    int32_t imin, imax, jmin, jmax, kmin, kmax;
    int32_t type, width_z, width_yz, size;
    uint64_t storage, origin;
    char     name[30];
    char     pad[2];
} disk_mesh_t;

int mh_mesh_load(const char *path, const int n[3], double *block)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail("cannot open '%s'", path);
    int32_t r[6];
    disk_mesh_t m;
    if (fread(r, sizeof(int32_t), 6, fp) != 6 || fread(&m, sizeof(m), 1, fp) != 1) {
        fclose(fp);
        return fail("'%s': short header", path);
    }
    if (m.type != 1) {
        fclose(fp);
        return fail("'%s': not a vector mesh (type %d)", path, m.type);
    }
    const long span = r[5] - r[2] + 1;      /* kmax - kmin + 1 */
    double *line = (double *)malloc(sizeof(double) * 3 * (span > 0 ? span : 1));
    if (!line) {
        fclose(fp);
        return fail("out of memory");
    }
    int rc = 0;
    for (int i = r[0]; i <= r[3] && !rc; ++i)
        for (int j = r[1]; j <= r[4] && !rc; ++j) {
            if (fread(line, sizeof(double) * 3, span, fp) != (size_t)span) {
                rc = fail("'%s': short data", path);
                break;
            }
            if (i < 0 || i > n[0] || j < 0 || j > n[1]) continue;
            for (int k = r[2]; k <= r[5]; ++k) {
                if (k < 0 || k > n[2]) continue;
                double *dst = block + 3 * (((long)i * (n[1] + 1) + j) * (n[2] + 1) + k);
                memcpy(dst, line + 3 * (k - r[2]), sizeof(double) * 3);
            }
        }
    free(line);
    fclose(fp);
    return rc;
}

int mh_mesh_save_ex(const char *path, const int n[3], const double *block, int ncomp,
                    const int lo[3], const int hi[3], int ghost, const char *name)
{
    if (ncomp != 1 && ncomp != 3) return fail("'%s': %d components per node", path, ncomp);
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail("cannot create '%s'", path);
    const int32_t r[6] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
    /* the mesh_t of the saved mesh: storage [-ghost, n + ghost] per axis */
    disk_mesh_t m;
    memset(&m, 0, sizeof(m));
    m.imin = -ghost; m.imax = n[0] + ghost;
    m.jmin = -ghost; m.jmax = n[1] + ghost;
    m.kmin = -ghost; m.kmax = n[2] + ghost;
    m.type = ncomp == 3;
    m.width_z = n[2] + 1 + 2 * ghost;
    m.width_yz = (n[1] + 1 + 2 * ghost) * m.width_z;
    m.size = (int32_t)((long)(n[0] + 1 + 2 * ghost) * m.width_yz * 8 * ncomp);
    snprintf(m.name, sizeof(m.name), "%s", name);
    int ok = fwrite(r, sizeof(int32_t), 6, fp) == 6 && fwrite(&m, sizeof(m), 1, fp) == 1;
    const long span = hi[2] - lo[2] + 1;
    for (int i = lo[0]; i <= hi[0] && ok; ++i)
        for (int j = lo[1]; j <= hi[1] && ok; ++j) {
            const double *src = block + ncomp * (((long)i * (n[1] + 1) + j) * (n[2] + 1) + lo[2]);
            ok = fwrite(src, sizeof(double) * ncomp, span, fp) == (size_t)span;
        }
    return (fclose(fp) == 0 && ok) ? 0 : fail("write error on '%s'", path);
}

int mh_mesh_save(const char *path, const int n[3], const double *block, const int lo[3],
                 const int hi[3], const char *name)
{
    /* the core.out E/H meshes span [min - 1, max + 1] (source/core/main.c:109-115) */
    return mh_mesh_save_ex(path, n, block, 3, lo, hi, 1, name);
}

/* ---- marker records, source/core/plasma_IO.c:273-287 ------------------- */

int mh_markers_load(const char *path, void **rec, long *n, int *recsize)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail("cannot open '%s'", path);
    int64_t N = 0;
    struct stat st;
    if (fread(&N, sizeof(N), 1, fp) != 1 || fstat(fileno(fp), &st) || N < 0) {
        fclose(fp);
        return fail("'%s': bad header", path);
    }
    const long bytes = (long)st.st_size - 8;
    int rs = 64;
    if (N > 0) {
        if (bytes % N || (bytes / N != 56 && bytes / N != 64)) {
            fclose(fp);
            return fail("'%s': %ld bytes for %ld markers", path, bytes, (long)N);
        }
        rs = (int)(bytes / N);
    }
    void *buf = malloc(N > 0 ? (size_t)bytes : 1);
    if (!buf) {
        fclose(fp);
        return fail("out of memory (%ld bytes)", bytes);
    }
    if (N > 0 && fread(buf, rs, N, fp) != (size_t)N) {
        free(buf);
        fclose(fp);
        return fail("'%s': short data", path);
    }
    fclose(fp);
    *rec = buf;
    *n = (long)N;
    *recsize = rs;
    return 0;
}

int mh_markers_save(const char *path, const void *rec, long n, int recsize)
{
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail("cannot create '%s'", path);
    const int64_t N = n;
    int ok = fwrite(&N, sizeof(N), 1, fp) == 1;
    if (ok && n > 0) ok = fwrite(rec, recsize, n, fp) == (size_t)n;
    return (fclose(fp) == 0 && ok) ? 0 : fail("write error on '%s'", path);
}

/* ---- checkpoint descriptors, source/IO_sys.c:20-124, IO_fileMap.c ----- */

static int read_first_ints(const char *path, int *v, int nv)
{
    mh_cfg_t c;
    if (mh_cfg_load(path, &c)) return -1;
    int rc = 0;
    for (int q = 0; q < nv && !rc; ++q) rc = mh_tag_int(&c.tags[0], q, v + q);
    mh_cfg_free(&c);
    return rc;
}

int mh_sys_records(const char *bindir)
{
    char p[1024];
    snprintf(p, sizeof(p), "%s/sys.N", bindir);
    FILE *fp = fopen(p, "rt");
    if (!fp) return -1;
    fclose(fp);
    int v;
    return read_first_ints(p, &v, 1) ? -1 : v;
}

int mh_sys_info(const char *bindir, int rec, double *time, int *cpus, int *fmap)
{
    char p[1024];
    snprintf(p, sizeof(p), "%s/sys_%06d.txt", bindir, rec);
    mh_cfg_t c;
    if (mh_cfg_load(p, &c)) return -1;
    int rc = mh_tag_double(&c.tags[0], 0, time);
    if (!rc) rc = mh_tag_int(&c.tags[0], 1, cpus);
    if (!rc) rc = mh_tag_int(&c.tags[0], 2, fmap);
    mh_cfg_free(&c);
    return rc;
}

static int write_text(const char *path, const char *text)
{
    FILE *fp = fopen(path, "wt");
    if (!fp) return fail("cannot create '%s'", path);
    const int ok = fputs(text, fp) >= 0;
    return (fclose(fp) == 0 && ok) ? 0 : fail("write error on '%s'", path);
}

int mh_filemap_next(const char *bindir, int cont)
{
    char p[1024];
    snprintf(p, sizeof(p), "%s/%s", bindir, cont ? "fileMap.N" : "fileMap_start.txt");
    FILE *fp = fopen(p, "rt");
    if (!fp) return -1;
    int v = -1;
    if (fscanf(fp, " @ %d", &v) != 1) v = -1;
    fclose(fp);
    return v;
}

int mh_filemap_save(const char *bindir, int id, const int n[3], int setup_mode)
{
    char p[1024], t[512];
    snprintf(p, sizeof(p), "%s/fileMap_%06d.txt", bindir, id);
    snprintf(t, sizeof(t), "1 cpus total.\n0: (0 0 0) - (%d %d %d)\n", n[0], n[1], n[2]);
    if (write_text(p, t)) return -1;
    if (setup_mode) {
        snprintf(p, sizeof(p), "%s/fileMap_start.txt", bindir);
        snprintf(t, sizeof(t),
                 "@ %d filemaps are generated by setup - keep them on the start of simulation.\n",
                 id + 1);
        if (write_text(p, t)) return -1;
    }
    snprintf(p, sizeof(p), "%s/fileMap.N", bindir);
    snprintf(t, sizeof(t), "@ %d filemaps total is written.\n", id + 1);
    if (write_text(p, t)) return -1;
    return id;
}

/* partition_save of one rank (source/misc_partition.c:422-435,
 * source/misc_partitionPack.c:35-76): the packed size, then the domain box,
 * no split axis, 1 cpu, 1 tree node; the root record is depth 0, cpus
 * [0, 0], no split, next record at 18.  The reference leaves the children
 * offset of a leaf unset; it is written as 0 here. */
static int prtn_save(const char *path, const int n[3])
{
    const int32_t v[19] = {18, 0, 0, 0, n[0], n[1], n[2], 0, 0, 0, 1, 1, 0, 0, 0, 0, 0, 0, 18};
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail("cannot create '%s'", path);
    const int ok = fwrite(v, sizeof(int32_t), 19, fp) == 19;
    return (fclose(fp) == 0 && ok) ? 0 : fail("write error on '%s'", path);
}

int mh_sys_write(const char *bindir, int rec, double time, const int n[3], int fmap)
{
    char p[1024], t[512];
    snprintf(p, sizeof(p), "%s/sys_prtn_%06d.bin", bindir, rec);
    if (prtn_save(p, n)) return -1;
    snprintf(p, sizeof(p), "%s/sys_%06d.txt", bindir, rec);
    snprintf(t, sizeof(t),
             "@ %.14e    time.\n@ %d       cpus total.\n@ %d       file map id.\n"
             "That is checkpoint (record of the full system state) #%d.\n",
             time, 1, fmap, rec);
    if (write_text(p, t)) return -1;
    snprintf(p, sizeof(p), "%s/sys.N", bindir);
    snprintf(t, sizeof(t), "@ %d\t  system check-point records written.\n", rec + 1);
    return write_text(p, t);
}

int mh_tec_write(const char *bindir, int rec, double time, const int n[3], int fmap)
{
    char p[1024];
    snprintf(p, sizeof(p), "%s/tec_%06d.txt", bindir, rec);
    FILE *fp = fopen(p, "wb");
    if (!fp) return fail("cannot create '%s'", p);
    const int32_t cf[2] = {1, fmap};
    int ok = fwrite(&time, sizeof(double), 1, fp) == 1 && fwrite(cf, sizeof(int32_t), 2, fp) == 2;
    if (fclose(fp) != 0 || !ok) return fail("write error on '%s'", p);
    snprintf(p, sizeof(p), "%s/tec_prtn_%06d.bin", bindir, rec);
    return prtn_save(p, n);
}

int mh_tec_count(const char *bindir, int count)
{
    char p[1024], t[128];
    snprintf(p, sizeof(p), "%s/tecplot.N", bindir);
    snprintf(t, sizeof(t), "@ %d  tecplots' records written.\n", count);
    return write_text(p, t);
}

int mh_tec_records(const char *bindir)
{
    char p[1024];
    snprintf(p, sizeof(p), "%s/tecplot.N", bindir);
    FILE *fp = fopen(p, "rt");
    if (!fp) return -1;
    int v = -1;
    if (fscanf(fp, " @ %d", &v) != 1) v = -1;
    fclose(fp);
    return v;
}

/* run_probes.cfg (source/core/diag_probe.c:155-330) */
static int probe_add(int **ijk, int *n, int *cap, const mh_domain_t *d, const mh_units_t *u,
                     double x, double y, double z)
{
    if (*n == *cap) {
        *cap = *cap ? 2 * *cap : 64;
        int *g = (int *)realloc(*ijk, sizeof(int) * 3 * (size_t)*cap);
        if (!g) return fail("out of memory");
        *ijk = g;
    }
    const double r[3] = {x, y, z};
    for (int a = 0; a < 3; ++a)
        (*ijk)[3 * *n + a] = (int)(r[a] * u->micron / d->h[a] + 128.0) - 128;
    ++*n;
    return 0;
}

int mh_probes_parse(const char *path, const mh_domain_t *d, const mh_units_t *u, int **ijk, int *n)
{
    *ijk = NULL;
    *n = 0;
    FILE *fp = fopen(path, "rt");
    if (!fp) return 0;                       /* no file: no probes */
    /* the '@' words and '>' values in file order */
    char line[4096];
    char kind[4096];
    double *v = NULL;
    int nv = 0, vcap = 0, cap = 0, rc = 0;
    char words[64][16];
    int nw = 0, wat[64];
    while (fgets(line, sizeof(line), fp) && !rc) {
        const char *s = line;
        while (*s == ' ' || *s == '\t') ++s;
        if (*s == '@') {
            if (sscanf(s + 1, "%4095s", kind) != 1) continue;
            if (nw == 64) {
                rc = fail("'%s': too many probe groups", path);
                break;
            }
            snprintf(words[nw], sizeof(words[nw]), "%s", kind);
            wat[nw++] = nv;
            if (!strcmp(kind, "end")) break;
        } else if (*s == '>') {
            if (nv == vcap) {
                vcap = vcap ? 2 * vcap : 64;
                double *g = (double *)realloc(v, sizeof(double) * vcap);
                if (!g) {
                    rc = fail("out of memory");
                    break;
                }
                v = g;
            }
            if (sscanf(s + 1, "%lf", v + nv) != 1) {
                rc = fail("'%s': bad probe value '%s'", path, s);
                break;
            }
            ++nv;
        }
    }
    fclose(fp);
    for (int w = 0; w < nw && !rc; ++w) {
        const double *a = v + wat[w];
        const int have = (w + 1 < nw ? wat[w + 1] : nv) - wat[w];
        const char *k = words[w];
        if (!strcmp(k, "end")) break;
        if (!strcmp(k, "point")) {
            if (have < 3) rc = fail("'%s': a point needs 3 values", path);
            else rc = probe_add(ijk, n, &cap, d, u, a[0], a[1], a[2]);
        } else if (!strcmp(k, "line")) {
            if (have < 7) {
                rc = fail("'%s': a line needs 7 values", path);
                break;
            }
            const int N = 2 + abs((int)a[6]);
            double x = a[0], y = a[1], z = a[2];
            const double dx = (a[3] - x) / (N - 1), dy = (a[4] - y) / (N - 1), dz = (a[5] - z) / (N - 1);
            for (int i = 0; i < N && !rc; ++i, x += dx, y += dy, z += dz)
                rc = probe_add(ijk, n, &cap, d, u, x, y, z);
        } else if (!strcmp(k, "rectangle")) {
            if (have < 11) {
                rc = fail("'%s': a rectangle needs 11 values", path);
                break;
            }
            const int Np = 2 + abs((int)a[9]), Nq = 2 + abs((int)a[10]);
            const double px = (a[3] - a[0]) / (Np - 1), py = (a[4] - a[1]) / (Np - 1),
                         pz = (a[5] - a[2]) / (Np - 1), qx = (a[6] - a[0]) / (Nq - 1),
                         qy = (a[7] - a[1]) / (Nq - 1), qz = (a[8] - a[2]) / (Nq - 1);
            for (int p = 0; p < Np && !rc; ++p)
                for (int q = 0; q < Nq && !rc; ++q)
                    rc = probe_add(ijk, n, &cap, d, u, a[0] + p * px + q * qx, a[1] + p * py + q * qy,
                                   a[2] + p * pz + q * qz);
        } else if (!strcmp(k, "volume")) {
            if (have < 15) {
                rc = fail("'%s': a volume needs 15 values", path);
                break;
            }
            const int Np = 2 + abs((int)a[12]), Nq = 2 + abs((int)a[13]), Nr = 2 + abs((int)a[14]);
            const double px = (a[3] - a[0]) / (Np - 1), py = (a[4] - a[1]) / (Np - 1),
                         pz = (a[5] - a[2]) / (Np - 1), qx = (a[6] - a[0]) / (Nq - 1),
                         qy = (a[7] - a[1]) / (Nq - 1), qz = (a[8] - a[2]) / (Nq - 1),
                         rx = (a[9] - a[0]) / (Nr - 1), ry = (a[10] - a[1]) / (Nr - 1),
                         rz = (a[11] - a[2]) / (Nr - 1);
            for (int p = 0; p < Np && !rc; ++p)
                for (int q = 0; q < Nq && !rc; ++q)
                    for (int r = 0; r < Nr && !rc; ++r)
                        rc = probe_add(ijk, n, &cap, d, u, a[0] + p * px + q * qx + r * rx,
                                       a[1] + p * py + q * qy + r * ry, a[2] + p * pz + q * qz + r * rz);
        } else {
            rc = fail("'%s': bad probe group '%s' (point, line, rectangle, volume or end)", path, k);
        }
    }
    free(v);
    if (rc) {
        free(*ijk);
        *ijk = NULL;
        *n = 0;
        return rc;
    }
    /* repeated nodes: the later one replaced by the last probe, which is
     * tested in its place (diag_probe.c:302-308) */
    for (int p1 = 0; p1 < *n; ++p1)
        for (int p2 = p1 + 1; p2 < *n; ++p2) {
            int *A = *ijk + 3 * p1, *B = *ijk + 3 * p2;
            if (A[0] == B[0] && A[1] == B[1] && A[2] == B[2]) {
                --*n;
                memcpy(B, *ijk + 3 * *n, sizeof(int) * 3);
                --p2;
            }
        }
    return 0;
}

/* ---- diag.dat, source/core/diag_WDensity.c:81-105 ---------------------- */

int mh_diag_append(const char *path, const double *pts, long npts, double tau, int fresh)
{
    FILE *fp = fresh ? NULL : fopen(path, "rt");
    const int newfile = !fp;
    if (fp) fclose(fp);
    fp = fopen(path, newfile ? "wt" : "at");
    if (!fp) return fail("cannot open '%s'", path);
    if (newfile)
        fprintf(fp, "variables = t, W_E, W_M, W_T_x, W_T_y, W_T_z, W_p_e_r_p, W_k_i_n, W_E_M, W_`S\n"
                    "zone f=point\n");
    for (long q = 0; q < npts; ++q) {
        const double *d = pts + 6 * q;
        if (d[0] < 0.5 * tau) continue;
        fprintf(fp, "%e %e %e %e %e %e %e %e %e %e\n", d[0], d[1], d[2], d[3], d[4], d[5],
                d[4] + d[5], d[3] + d[4] + d[5], d[1] + d[2], d[1] + d[2] + d[3] + d[4] + d[5]);
    }
    return fclose(fp) == 0 ? 0 : fail("write error on '%s'", path);
}

/* ---- tmp/stop.flag, source/core/main.c:121-160,231-236 ----------------- */

int mh_stopflag_read(const char *path, int *force, int *limit)
{
    FILE *fp = fopen(path, "rt");
    if (!fp) return fail("cannot open '%s'", path);
    const int k = fscanf(fp, "%d %d", force, limit);
    fclose(fp);
    return k == 2 ? 0 : fail("'%s': two integers expected", path);
}

int mh_stopflag_default(const char *path)
{
    FILE *fp = fopen(path, "rt");
    if (fp) {
        fclose(fp);
        return 0;
    }
    char t[64];
    snprintf(t, sizeof(t), "0\n%d\n", 23 * 60 * 60 + 30 * 60);
    return write_text(path, t);
}

/* ---- laser source files, source/scpic/em_mirror.c:291-485 -------------- */

#define SOURCE_PLANE 5          /* OFFSET / SOURCE_X_OFFSET of em_mirror.c */

int mh_laser_load(const char *path, const mh_domain_t *d, const mh_units_t *u, mh_laser_t *l)
{
    memset(l, 0, sizeof(*l));
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail("cannot open '%s'", path);
    char line[4096];
    double v[10];
    int nv = 0, dashes = 0;
    /* header: title, dashes, 10 '@' lines, dashes; the maps follow */
    while (dashes < 2 && fgets(line, sizeof(line), fp)) {
        if (!strncmp(line, "----", 4)) {
            ++dashes;
        } else if (line[0] == '@' && nv < 10) {
            if (sscanf(line + 1, "%lf", v + nv) != 1) break;
            ++nv;
        }
    }
    if (dashes < 2 || nv != 10) {
        fclose(fp);
        return fail("'%s': bad header (%d values)", path, nv);
    }
    double A = v[0];
    if (A > 0)   /* intensity [W/cm^2] -> e E/(m omega c) */
        A = u->lambda * CGS_E * sqrt(8 * M_PI * A * 1e7 / CGS_C) / (CGS_M * CGS_C * CGS_C);
    else
        A *= -2 * M_PI;
    const double xf = v[2] * 1e-4 / u->lambda;
    l->A = A;
    l->dt = v[5] * 1e-15 / u->t * sqrt(2.0 / log(2.0));
    l->T0 = v[6] * 1e-15 / u->t;
    l->phi0 = v[7];
    l->forward = v[9] > 0;
    l->plane = SOURCE_PLANE;
    l->delay = l->forward ? xf - SOURCE_PLANE * d->h[0]
                          : d->dmn_max[0] - SOURCE_PLANE * d->h[0] - xf;
    l->nodes = (long)(d->dmn_max[1] + 1 - 2 * SOURCE_PLANE) * (d->dmn_max[2] + 1 - 2 * SOURCE_PLANE);
    if (l->nodes <= 0) {
        fclose(fp);
        return fail("'%s': domain too small for a laser source", path);
    }
    l->map = (double *)malloc(sizeof(double) * 8 * l->nodes);
    if (!l->map) {
        fclose(fp);
        return fail("out of memory");
    }
    const size_t want = 8 * (size_t)l->nodes;
    const size_t got = fread(l->map, sizeof(double), want, fp);
    fclose(fp);
    if (got != want) {
        mh_laser_free(l);
        return fail("'%s': %zu of %zu map values", path, got, want);
    }
    return 0;
}

int mh_laser_save(const char *path, const mh_domain_t *d, const mh_units_t *u,
                  double a, double fnum, const double focus_um[3], double dt_fs, double T0_fs,
                  double phi0, int forward, const double *map)
{
    (void)u;
    FILE *fp = fopen(path, "wb");
    if (!fp) return fail("cannot create '%s'", path);
    fprintf(fp, "Mirror focused laser beam #%d\n", 0);
    fprintf(fp, "---------------------------------------\n");
    fprintf(fp, "@ %.15le   Poynting vector at the best focus: positive => I [W/cm^2] / negative => e*E/(m*omega*c).\n", a);
    fprintf(fp, "@ %.15le   Mirror f-number.\n", fnum);
    fprintf(fp, "@ %.15le   Focus X coordinate [micron].\n", focus_um[0]);
    fprintf(fp, "@ %.15le   Focus Y coordinate [micron].\n", focus_um[1]);
    fprintf(fp, "@ %.15le   Focus Z coordinate [micron].\n", focus_um[2]);
    fprintf(fp, "@ %.15le   Pulse duration, [fs] (intensity FWHM), Gaussian envelope.\n", dt_fs);
    fprintf(fp, "@ %.15le   Pulse to focus arrival time [fs].\n", T0_fs);
    fprintf(fp, "@ %.15le   Phase shift.\n", phi0);
    fprintf(fp, "@ %.15le   Ratio between incident pulse and mirror radii (-1: plane wave).\n", -1.0);
    fprintf(fp, "@ %d       Direction of propagation (positive or negative).\n", forward ? 1 : -1);
    fprintf(fp, "---------------------------------------\n");
    const long nodes = (long)(d->dmn_max[1] + 1 - 2 * SOURCE_PLANE) * (d->dmn_max[2] + 1 - 2 * SOURCE_PLANE);
    const int ok = fwrite(map, sizeof(double), 8 * nodes, fp) == (size_t)(8 * nodes);
    return (fclose(fp) == 0 && ok) ? 0 : fail("write error on '%s'", path);
}

void mh_laser_free(mh_laser_t *l)
{
    free(l->map);
    l->map = NULL;
}

/* ---- RANMAR, source/misc_ranmar.c:78-146 ------------------------------ */

int mh_ranmar_init(mh_ranmar_t *r, int ij, int kl)
{
    if (ij < 0 || ij > 31328 || kl < 0 || kl > 30081)
        return fail("RANMAR seeds out of range (0 <= ij <= 31328, 0 <= kl <= 30081)");
    int i = (ij / 177) % 177 + 2, j = ij % 177 + 2, k = (kl / 169) % 178 + 1, l = kl % 169;
    for (int ii = 0; ii < 97; ++ii) {
        double s = 0.0, t = 0.5;
        for (int jj = 0; jj < 24; ++jj) {
            const int m = (((i * j) % 179) * k) % 179;
            i = j;
            j = k;
            k = m;
            l = (53 * l + 1) % 169;
            if ((l * m) % 64 >= 32) s += t;
            t *= 0.5;
        }
        r->u[ii] = s;
    }
    r->c = 362436.0 / 16777216.0;
    r->cd = 7654321.0 / 16777216.0;
    r->cm = 16777213.0 / 16777216.0;
    r->i97 = 96;
    r->j97 = 32;
    return 0;
}

void mh_ranmar(mh_ranmar_t *r, double *out, int len)
{
    for (int q = 0; q < len; ++q) {
        double uni = r->u[r->i97] - r->u[r->j97];
        if (uni < 0.0) uni += 1.0;
        r->u[r->i97] = uni;
        if (--r->i97 < 0) r->i97 = 96;
        if (--r->j97 < 0) r->j97 = 96;
        r->c -= r->cd;
        if (r->c < 0.0) r->c += r->cm;
        uni -= r->c;
        if (uni < 0.0) uni += 1.0;
        out[q] = uni;
    }
}

/* len deviates of N(0, sigma^2) from the stream: the Box-Muller transform
 * (G. E. P. Box and M. E. Muller, 1958) of consecutive uniform pairs
 * (u1, u2), r = sqrt(-2 ln(1 - u1)) (1 - u1 lies in (0, 1]), angle 2 pi u2,
 * giving r cos and r sin in that order; an odd len drops the last sine. */
void mh_ranmar_normal(mh_ranmar_t *r, double *out, long len, double sigma)
{
    double u[2];
    for (long q = 0; q < len; q += 2) {
        mh_ranmar(r, u, 2);
        const double rad = sigma * sqrt(-2.0 * log1p(-u[0])), a = 2.0 * M_PI * u[1];
        out[q] = rad * cos(a);
        if (q + 1 < len) out[q + 1] = rad * sin(a);
    }
}

/* ---- TFSF recordings, source/core/em_TFSF.c ---------------------------- */

/* Face file names by slot axis*4 + top*2 + comp (tfsfE / tfsfH,
 * source/core/em_TFSF.c:82-123): the interface's file-name contract. */
static const char *const TFSF_E_NAME[12] = {
    "Ey_xMin", "Ez_xMin", "Ey_xMax", "Ez_xMax", "Ez_yMin", "Ex_yMin",
    "Ez_yMax", "Ex_yMax", "Ex_zMin", "Ey_zMin", "Ex_zMax", "Ey_zMax"};
static const char *const TFSF_H_NAME[12] = {
    "Hz_xMin", "Hy_xMin", "Hz_xMax", "Hy_xMax", "Hx_yMin", "Hz_yMin",
    "Hx_yMax", "Hz_yMax", "Hy_zMin", "Hx_zMin", "Hy_zMax", "Hx_zMax"};

static int read_frames(const char *path, long count, double **out)
{
    FILE *fp = fopen(path, "rb");
    if (!fp) return fail("cannot open '%s'", path);
    double *d = (double *)malloc(sizeof(double) * (size_t)count);
    int ok = d && fseek(fp, 11 * sizeof(int32_t), SEEK_SET) == 0 &&
             fread(d, sizeof(double), (size_t)count, fp) == (size_t)count;
    fclose(fp);
    if (!ok) {
        free(d);
        return fail("'%s': fewer frames than lastFrameTime.bin promises", path);
    }
    *out = d;
    return 0;
}

/* Node box of face slot q of the interface reg (E faces 0..11, H faces
 * 12..23), as TFSF_init lays them out (source/core/em_TFSF.c:500-537). */
static void tfsf_face_box(const int reg[6], int q, int lo[3], int hi[3], int *ep, int *eq)
{
    const int f = q % 12, ea = f >> 2, top = (f >> 1) & 1, comp = f & 1;
    const int side[2][3] = {{reg[0], reg[2], reg[4]}, {reg[1], reg[3], reg[5]}};
    for (int a = 0; a < 3; ++a) {
        lo[a] = side[0][a];
        hi[a] = side[1][a];
    }
    lo[ea] = hi[ea] = side[top][ea];
    lo[(ea + 1 + comp) % 3] += 1;
    if (q >= 12) {
        lo[ea] += top;
        hi[ea] += top;
    }
    *ep = (ea + 1) % 3;
    *eq = (ea + 2) % 3;
}

static const char *tfsf_face_name(int q) { return q < 12 ? TFSF_E_NAME[q] : TFSF_H_NAME[q - 12]; }

int mh_tfsf_record_begin(const char *dir, const int reg[6], const mh_domain_t *d)
{
    char p[1024], t[2048];
    snprintf(p, sizeof(p), "%s/consistencyCheck.dat", dir);
    snprintf(t, sizeof(t),
             "@ %d\tInterface size in X direction (nodes)\n@ %d\tInterface size in Y direction (nodes)\n"
             "@ %d\tInterface size in Z direction (nodes)\n@ %.12e\tTime step\n"
             "@ %.12e\tSpatial step along X\n@ %.12e\tSpatial step along Y\n"
             "@ %.12e\tSpatial step along Z\n",
             reg[1] - reg[0], reg[3] - reg[2], reg[5] - reg[4], d->tau, d->h[0], d->h[1], d->h[2]);
    if (write_text(p, t)) return -1;
    /* TFSF_defragInit: appended, E faces then H faces */
    snprintf(p, sizeof(p), "%s/TFSF_defrag.map", dir);
    FILE *fp = fopen(p, "at");
    if (!fp) return fail("cannot open '%s'", p);
    for (int q = 0; q < 24; ++q) {
        int lo[3], hi[3], ep, eq;
        tfsf_face_box(reg, q, lo, hi, &ep, &eq);
        fprintf(fp, "[%d %d] [%d %d] [%d %d]\nep = %d, eq = %d\n", lo[0], hi[0], lo[1], hi[1], lo[2], hi[2],
                ep, eq);
        fprintf(fp, "%s/%s_99999.bin\n%s/%s_000.bin\n---\n", dir, tfsf_face_name(q), dir, tfsf_face_name(q));
    }
    return fclose(fp) == 0 ? 0 : fail("write error on '%s'", p);
}

int mh_tfsf_record_save(const char *dir, const int reg[6], int nframes, const double *data)
{
    char p[1024];
    long at = 0;
    for (int q = 0; q < 24; ++q) {
        int lo[3], hi[3], ep, eq;
        tfsf_face_box(reg, q, lo, hi, &ep, &eq);
        const long vol = (long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
        /* reg_t of the fragment: box, cpu 0, mode 1 (record), no wrap */
        const int32_t h[11] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], 0, 1, 0, 0, 0};
        snprintf(p, sizeof(p), "%s/%s_000.bin", dir, tfsf_face_name(q));
        FILE *fp = fopen(p, "wb");
        if (!fp) return fail("cannot create '%s'", p);
        int ok = fwrite(h, sizeof(int32_t), 11, fp) == 11 &&
                 fwrite(data + at, sizeof(double), (size_t)(vol * nframes), fp) == (size_t)(vol * nframes);
        if (fclose(fp) != 0 || !ok) return fail("write error on '%s'", p);
        at += vol * nframes;
    }
    snprintf(p, sizeof(p), "%s/lastFrameTime.bin", dir);
    FILE *fp = fopen(p, "wb");
    if (!fp) return fail("cannot create '%s'", p);
    const int32_t nf = nframes;
    const int ok = fwrite(&nf, sizeof(nf), 1, fp) == 1;
    return (fclose(fp) == 0 && ok) ? 0 : fail("write error on '%s'", p);
}

/* One job of TFSF_defrag.map: the face box, its unrolling, the joined file
 * and its fragments; each frame of the joined file is its old content (or
 * zero) plus every fragment's nodes inside the filter (TFSFdefrag,
 * source/TFSFdefrag/main.c:50-198). */
typedef struct {
    int    lo[3], hi[3];
    int    ep, eq;
    char   joined[1024];
    int    nfrag;
    char **frag;
} defrag_job_t;

static int defrag_job(const defrag_job_t *j, const int flo[3], const int fhi[3], int skip, int *frames_out,
                      int *frames_before)
{
    const long nodes = (long)(j->hi[0] - j->lo[0] + 1) * (j->hi[1] - j->lo[1] + 1) * (j->hi[2] - j->lo[2] + 1);
    const int32_t hdr[11] = {j->lo[0], j->lo[1], j->lo[2], j->hi[0], j->hi[1], j->hi[2], 0, 0, 0, 0, 0};
    FILE *out = fopen(j->joined, "r+b");
    if (!out) out = fopen(j->joined, "w+b");
    if (!out) return fail("cannot open '%s'", j->joined);
    double *full = (double *)calloc((size_t)nodes, sizeof(double));
    FILE **in = (FILE **)calloc((size_t)j->nfrag, sizeof(FILE *));
    int32_t (*pr)[11] = (int32_t(*)[11])calloc((size_t)j->nfrag, sizeof(*pr));
    int rc = (!full || !in || !pr) ? fail("out of memory") : 0;
    long maxpiece = 1;
    for (int f = 0; f < j->nfrag && !rc; ++f) {
        in[f] = fopen(j->frag[f], "rb");
        if (!in[f] || fread(pr[f], sizeof(int32_t), 11, in[f]) != 11) {
            rc = fail("cannot read fragment '%s'", j->frag[f]);
            break;
        }
        const long v = (long)(pr[f][3] - pr[f][0] + 1) * (pr[f][4] - pr[f][1] + 1) * (pr[f][5] - pr[f][2] + 1);
        if (v > maxpiece) maxpiece = v;
    }
    double *piece = rc ? NULL : (double *)malloc(sizeof(double) * (size_t)maxpiece);
    if (!rc && !piece) rc = fail("out of memory");
    int frames = 0;
    *frames_before = 0;
    if (!rc) {
        /* the header, then 'skip' + 1 zero frames when the file is shorter
         * than that; the frames it held count as written */
        fseek(out, 0L, SEEK_SET);
        fwrite(hdr, sizeof(int32_t), 11, out);
        fseek(out, 0L, SEEK_END);
        const long tail = ftell(out);
        if (tail < (long)sizeof(hdr) + (long)(skip * nodes * sizeof(double))) {
            fseek(out, 0L, SEEK_SET);
            fwrite(hdr, sizeof(int32_t), 11, out);
            for (int k = 0; k <= skip; ++k) fwrite(full, sizeof(double), (size_t)nodes, out);
        }
        *frames_before = (int)((tail - (long)sizeof(hdr)) / (nodes * (long)sizeof(double)));
        const int er = 3 - j->ep - j->eq;
        for (;; ++frames) {
            const long pos = (long)sizeof(hdr) + (long)(skip + frames) * nodes * (long)sizeof(double);
            fseek(out, pos, SEEK_SET);
            const size_t got = fread(full, sizeof(double), (size_t)nodes, out);
            if (got == 0)
                memset(full, 0, sizeof(double) * (size_t)nodes);
            else if (got != (size_t)nodes) {
                rc = fail("'%s': incomplete frame", j->joined);
                break;
            }
            int any = 0;
            for (int f = 0; f < j->nfrag; ++f) {
                const int32_t *r = pr[f];
                const long v = (long)(r[3] - r[0] + 1) * (r[4] - r[1] + 1) * (r[5] - r[2] + 1);
                const size_t g = fread(piece, sizeof(double), (size_t)v, in[f]);
                if (g == 0) continue;
                if (g != (size_t)v) {
                    rc = fail("'%s': incomplete fragment", j->frag[f]);
                    break;
                }
                any = 1;
                if (r[er] < flo[er] || r[er] > fhi[er]) continue;
                const int ep = j->ep, eq = j->eq;
                const int out_span = j->hi[eq] - j->lo[eq] + 1, in_span = r[3 + eq] - r[eq] + 1;
                const int p0 = r[ep] < flo[ep] ? flo[ep] : r[ep], p1 = r[3 + ep] > fhi[ep] ? fhi[ep] : r[3 + ep];
                const int q0 = r[eq] < flo[eq] ? flo[eq] : r[eq], q1 = r[3 + eq] > fhi[eq] ? fhi[eq] : r[3 + eq];
                for (int pp = p0; pp <= p1; ++pp)
                    for (int qq = q0; qq <= q1; ++qq)
                        full[(qq - j->lo[eq]) + (long)out_span * (pp - j->lo[ep])] +=
                            piece[(qq - r[eq]) + (long)in_span * (pp - r[ep])];
            }
            if (rc || !any) break;
            fseek(out, pos, SEEK_SET);
            if (fwrite(full, sizeof(double), (size_t)nodes, out) != (size_t)nodes) {
                rc = fail("write error on '%s'", j->joined);
                break;
            }
        }
    }
    for (int f = 0; f < j->nfrag; ++f)
        if (in && in[f]) fclose(in[f]);
    if (fclose(out) != 0 && !rc) rc = fail("write error on '%s'", j->joined);
    if (!rc)
        for (int f = 0; f < j->nfrag; ++f) remove(j->frag[f]);
    free(full);
    free(in);
    free(pr);
    free(piece);
    *frames_out = frames;
    return rc;
}

int mh_tfsf_defrag(const char *dir)
{
    char p[1024], line[2048];
    /* filter.cfg: I0, I1 (along x) and the frame shift (tag_TFSF.c:170-176) */
    int flo[3] = {INT_MIN, INT_MIN, INT_MIN}, fhi[3] = {INT_MAX, INT_MAX, INT_MAX}, skip = 0;
    snprintf(p, sizeof(p), "%s/filter.cfg", dir);
    int fv[3];
    if (read_first_ints(p, fv, 3)) return fail("'%s': 3 integers expected", p);
    flo[0] = fv[0];
    fhi[0] = fv[1];
    skip = fv[2];
    if (skip < 0) return fail("unsupported frame shift (%d)", skip);
    snprintf(p, sizeof(p), "%s/lastFrameTime.bin", dir);
    int32_t written = 0;
    FILE *fp = fopen(p, "rb");
    if (!fp || fread(&written, sizeof(written), 1, fp) != 1) {
        if (fp) fclose(fp);
        return fail("'%s': no frame count", p);
    }
    fclose(fp);
    snprintf(p, sizeof(p), "%s/TFSF_defrag.map", dir);
    FILE *map = fopen(p, "rt");
    if (!map) return 0;                                      /* nothing to join */
    int rc = 0;
    for (;;) {
        defrag_job_t j;
        memset(&j, 0, sizeof(j));
        if (!fgets(line, sizeof(line), map)) break;
        if (sscanf(line, "[%d %d] [%d %d] [%d %d]", &j.lo[0], &j.hi[0], &j.lo[1], &j.hi[1], &j.lo[2],
                   &j.hi[2]) != 6)
            break;
        if (!fgets(line, sizeof(line), map) || sscanf(line, "ep = %d, eq = %d", &j.ep, &j.eq) != 2 ||
            !fgets(line, sizeof(line), map) || sscanf(line, "%1023s", j.joined) != 1) {
            rc = fail("'%s': malformed job", p);
            break;
        }
        while (fgets(line, sizeof(line), map) && line[0] != '-') {
            char name[1024];
            if (sscanf(line, "%1023s", name) != 1) continue;
            char **g = (char **)realloc(j.frag, sizeof(char *) * (size_t)(j.nfrag + 1));
            if (!g) {
                rc = fail("out of memory");
                break;
            }
            j.frag = g;
            j.frag[j.nfrag++] = strdup(name);
        }
        int frames = 0, before = 0;
        if (!rc && j.nfrag) {
            rc = defrag_job(&j, flo, fhi, skip, &frames, &before);
            if (!rc && before > written) written = before;
            if (!rc && frames + skip > written) written = frames + skip;
        }
        for (int f = 0; f < j.nfrag; ++f) free(j.frag[f]);
        free(j.frag);
        if (rc) break;
    }
    fclose(map);
    if (rc) return rc;
    remove(p);
    snprintf(p, sizeof(p), "%s/lastFrameTime.bin", dir);
    fp = fopen(p, "wb");
    if (!fp) return fail("cannot create '%s'", p);
    const int ok = fwrite(&written, sizeof(written), 1, fp) == 1;
    return (fclose(fp) == 0 && ok) ? 0 : fail("write error on '%s'", p);
}

void mh_tfsf_free(mh_tfsf_t *t)
{
    for (int f = 0; f < 12; ++f) {
        free(t->dataE[f]);
        free(t->dataH[f]);
        t->dataE[f] = t->dataH[f] = NULL;
    }
}

int mh_tfsf_load(const char *dir, const mh_domain_t *d, mh_tfsf_t *t)
{
    memset(t, 0, sizeof(*t));
    t->mode = -1;
    char p[1024];
    snprintf(p, sizeof(p), "%s/TFSF.cfg", dir);
    FILE *probe = fopen(p, "rt");
    if (!probe) return 0;                                    /* no interface */
    fclose(probe);
    int v[7];
    if (read_first_ints(p, v, 7)) return fail("'%s': 7 integers expected", p);
    memcpy(t->reg, v, sizeof(t->reg));
    t->mode = v[6];
    if (t->mode == 1) return 0;                              /* record: not played */
    if (t->mode != 0 && t->mode != 2) return fail("'%s': unknown mode %d", p, t->mode);
    snprintf(p, sizeof(p), "%s/TFSF_defrag.map", dir);     /* TFSF_init, :477-478 */
    probe = fopen(p, "rt");
    if (probe) {
        fclose(probe);
        return fail("TFSF input data is not compiled (use 'TFSFdefrag')");
    }
    /* consistencyCheck.dat: sizes, tau, h (TFSF_turnPlay, :432-453) */
    snprintf(p, sizeof(p), "%s/consistencyCheck.dat", dir);
    mh_cfg_t c;
    if (mh_cfg_load(p, &c)) return fail("there are no recorded data ('%s')", p);
    int sz[3] = {0, 0, 0};
    double dt = 0, dh[3] = {0, 0, 0};
    int rc = 0;
    for (int a = 0; a < 3 && !rc; ++a) rc = mh_tag_int(&c.tags[0], a, sz + a);
    if (!rc) rc = mh_tag_double(&c.tags[0], 3, &dt);
    for (int a = 0; a < 3 && !rc; ++a) rc = mh_tag_double(&c.tags[0], 4 + a, dh + a);
    mh_cfg_free(&c);
    if (rc) return fail("'%s': malformed", p);
    for (int a = 0; a < 3; ++a)
        if (sz[a] != t->reg[2 * a + 1] - t->reg[2 * a])
            return fail("inconsistent TFSF size along axis '%c'", 'X' + a);
    if (!(fabs(1 - dt / d->tau) < 1e-11)) return fail("inconsistent TFSF time step");
    for (int a = 0; a < 3; ++a)
        if (!(fabs(1 - dh[a] / d->h[a]) < 1e-11))
            return fail("inconsistent TFSF spatial step along %c", 'X' + a);
    snprintf(p, sizeof(p), "%s/lastFrameTime.bin", dir);
    FILE *fp = fopen(p, "rb");
    int32_t nf = 0;
    const int okf = fp && fread(&nf, sizeof(nf), 1, fp) == 1;
    if (fp) fclose(fp);
    if (!okf || nf < 1) return fail("'%s': no frame count", p);
    t->nframes = nf;
    /* opened faces (TFSF_openFaces, :339-371): xMin xMax yMin yMax zMin zMax */
    int opened[6] = {0, 0, 0, 0, 0, 0};
    snprintf(p, sizeof(p), "%s/openFaces.cfg", dir);
    probe = fopen(p, "rt");
    if (probe) {
        fclose(probe);
        if (read_first_ints(p, opened, 6)) return fail("'%s': 6 integers expected", p);
    }
    /* face boxes as TFSF_init (:500-537); E and H faces of a slot share the
     * tangential extents, so one node count serves both recordings */
    for (int ea = 0; ea < 3; ++ea) {
        const int ep = (ea + 1) % 3, eq = (ea + 2) % 3;
        for (int top = 0; top < 2; ++top)
            for (int comp = 0; comp < 2; ++comp) {
                const int f = ea * 4 + top * 2 + comp, cE = (ea + 1 + comp) % 3;
                int lo[3] = {t->reg[0], t->reg[2], t->reg[4]}, hi[3] = {t->reg[1], t->reg[3], t->reg[5]};
                lo[cE] += 1;
                t->vol[f] = (long)(hi[ep] - lo[ep] + 1) * (hi[eq] - lo[eq] + 1);
                if (opened[2 * ea + top]) continue;
                snprintf(p, sizeof(p), "%s/%s_99999.bin", dir, TFSF_H_NAME[f]);
                if (read_frames(p, t->vol[f] * nf, t->dataE + f)) {
                    mh_tfsf_free(t);
                    return -1;
                }
                snprintf(p, sizeof(p), "%s/%s_99999.bin", dir, TFSF_E_NAME[f]);
                if (read_frames(p, t->vol[f] * nf, t->dataH + f)) {
                    mh_tfsf_free(t);
                    return -1;
                }
            }
    }
    return 0;
}
