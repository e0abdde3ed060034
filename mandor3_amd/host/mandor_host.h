/*
 * mandor_host.h -- host side of the MI355X PIC step in C: the reference's
 * configuration and binData record formats, RANMAR, and the pieces of the
 * core.out / setup.out drivers built on them (cvarin/mandor3 file:line
 * citations below name the format each function reads or writes).
 *
 * Everything here is plain C99 and runs on the CPU; the drivers reach the
 * GPU only through include/mpic.h.
 */
#ifndef MANDOR_HOST_H
#define MANDOR_HOST_H

#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- configuration files ------------------------------------------------ */

/* A parsed '@'/'>' parameter file (source/misc_cfgReader.c:20-111): the
 * values of the '@' lines in order, grouped by the '[tag]' line above them
 * (tag "" before the first tag); '>' lines are kept as optional values of
 * the same tag. */
typedef struct {
    char   name[64];
    int    nreq, nopt;
    char **req;           /* '@' values (first word after '@') */
    char **opt;           /* '>' values */
} mh_tag_t;

typedef struct {
    int       ntags;
    mh_tag_t *tags;
} mh_cfg_t;

int  mh_cfg_load(const char *path, mh_cfg_t *cfg);
void mh_cfg_free(mh_cfg_t *cfg);
/* Reads parameter 'idx' of a tag as a number; -1 when missing/malformed. */
int  mh_tag_int(const mh_tag_t *t, int idx, int *v);
int  mh_tag_double(const mh_tag_t *t, int idx, double *v);

/* run_mandor.cfg (source/core/main.c:205-228): six '@' ints. */
typedef struct {
    int total_steps, chpoint_full, chpoint_tecplot, continue_run, chpoint_stop, chpoint_spectr;
} mh_runcfg_t;
int mh_runcfg_load(const char *path, mh_runcfg_t *rc);

/* ---- binData records --------------------------------------------------- */

/* binData/domain.bin (source/misc_parameters.c:98-110): 7 doubles
 * {h1,h2,h3,Lx,Ly,Lz,tau}, int dmn_min[3], dmn_max[3], int bc_min[3],
 * bc_max[3]. */
typedef struct {
    double h[3], L[3], tau;
    int    dmn_min[3], dmn_max[3];
    int    bc_min[3], bc_max[3];
} mh_domain_t;
int mh_domain_load(const char *path, mh_domain_t *d);
int mh_domain_save(const char *path, const mh_domain_t *d);

/* binData/units.bin (source/misc_units.c:39-61): 15 doubles. */
typedef struct {
    double lambda, v, t, omega, E, S, A, n, rho, ncr, debye, micron, femtosecond, wpe, Tpe;
} mh_units_t;
void mh_units_from_lambda(double lambda_cm, mh_units_t *u);   /* source/misc_units.c:99-118 */
int  mh_units_load(const char *path, mh_units_t *u);
/* Charge density of the critical plasma in code units, e n_cr / rho0
 * (source/setup/tag_DF_uniform.c:38); 4 pi^2 up to rounding. */
double mh_rho_crit(const mh_units_t *u);
int  mh_units_save(const char *path, const mh_units_t *u);

/* sys_E/H_%06d(%03d).bin (source/type_mesh.c:285-311): int32 region
 * {imin,jmin,kmin,imax,jmax,kmax}, the 88-byte mesh_t header, then the
 * region's nodes (i, j, k order, k fastest) of 3 doubles.  'block' is the
 * node block [0,n] of the whole domain (k fastest, xyz interleaved); load
 * copies a record's region into it, save writes the region [lo, hi]. */
int mh_mesh_load(const char *path, const int n[3], double *block);
int mh_mesh_save(const char *path, const int n[3], const double *block, const int lo[3],
                 const int hi[3], const char *name);
/* The same for a vector (ncomp 3) or scalar (ncomp 1) block [0,n] whose
 * mesh_t spans [-ghost, n+ghost] (core.out E/H: 1; J, rho and every setup
 * mesh: 2, source/core/main.c:109-115, source/setup/main.c:76-80). */
int mh_mesh_save_ex(const char *path, const int n[3], const double *block, int ncomp,
                    const int lo[3], const int hi[3], int ghost, const char *name);

/* plasma_%06d_%03d.bin (source/core/plasma_IO.c:273-287): int64 N, then N
 * marker_t records (56 B, or 64 B with the tracer id).  load returns the
 * records in a malloc'd buffer and their size in *recsize. */
int mh_markers_load(const char *path, void **rec, long *n, int *recsize);
int mh_markers_save(const char *path, const void *rec, long n, int recsize);

/* Record descriptors of one rank owning the whole domain [0, n].
 * File maps (source/IO_fileMap.c:40-160): checkpoints and Tecplot records
 * draw their ids from one counter; mh_filemap_next gives the first id of a
 * run (fileMap_start.txt, or fileMap.N when continuing; -1 if absent) and
 * mh_filemap_save writes fileMap_%06d.txt, fileMap.N and, for setup,
 * fileMap_start.txt.  mh_sys_write: sys_%06d.txt, sys.N and sys_prtn_%06d.bin
 * (source/IO_sys.c:99-124); mh_tec_write: tec_%06d.txt {double time, int
 * cpus, int file map} and tec_prtn_%06d.bin (source/IO_tecplot.c:225-245);
 * mh_tec_count / mh_tec_records write / read tecplot.N (:259-263). */
int mh_sys_records(const char *bindir);                 /* sys.N, -1 if absent */
int mh_sys_info(const char *bindir, int rec, double *time, int *cpus, int *fmap);
int mh_filemap_next(const char *bindir, int cont);
int mh_filemap_save(const char *bindir, int id, const int n[3], int setup_mode);
int mh_sys_write(const char *bindir, int rec, double time, const int n[3], int fmap);
int mh_tec_write(const char *bindir, int rec, double time, const int n[3], int fmap);
int mh_tec_count(const char *bindir, int count);
int mh_tec_records(const char *bindir);

/* run_probes.cfg (source/core/diag_probe.c:171-330): '@ point', '@ line',
 * '@ rectangle', '@ volume' groups of '>' values [micron] up to '@ end',
 * turned into nodes (floor(x/h) with the +128 trick) without repeats.
 * *ijk is malloc'd (3 ints per probe); no file means no probes. */
int mh_probes_parse(const char *path, const mh_domain_t *d, const mh_units_t *u, int **ijk, int *n);

/* Energy history (source/core/diag_WDensity.c:30-158): points {t, W_E, W_M,
 * WTx, WTy, WTz} merged into binData/diag.dat (Tecplot ASCII, 10 columns,
 * points with t < tau/2 skipped). */
int mh_diag_append(const char *path, const double *pts, long npts, double tau, int fresh);

/* tmp/stop.flag (source/core/main.c:121-236): {force, time limit [s]}. */
int mh_stopflag_read(const char *path, int *force, int *limit);
int mh_stopflag_default(const char *path);

/* .EM_sources/laser/mirror_%02d (source/scpic/em_mirror.c:291-330,393-485):
 * text header of '@' values, then the E map and the H map, each
 * (n1+1-10)*(n2+1-10) records {y_cos,y_sin,z_cos,z_sin}.  Converted to the
 * soft-source parameters of mpic_add_laser. */
typedef struct {
    double A, T0, delay, dt, phi0;
    int    forward, plane;
    long   nodes;
    double *map;          /* 2 * nodes * 4 doubles: E map then H map */
} mh_laser_t;
int  mh_laser_load(const char *path, const mh_domain_t *d, const mh_units_t *u, mh_laser_t *l);
int  mh_laser_save(const char *path, const mh_domain_t *d, const mh_units_t *u,
                   double a0_or_intensity, double fnum, const double focus_um[3], double dt_fs,
                   double T0_fs, double phi0, int forward, const double *map);
void mh_laser_free(mh_laser_t *l);

/* .EM_sources/TFSF in play mode (source/core/em_TFSF.c:151-180 TFSF.cfg,
 * :339-371 openFaces.cfg, :400-470 consistencyCheck.dat, lastFrameTime.bin
 * and the defragmented recordings <face>_99999.bin: a reg_t header of 11
 * int32, then frames x face nodes doubles).  mode -1: no TFSF.cfg (no
 * interface); 1 (record) is reported, not loaded.  dataE[f] / dataH[f] are
 * the frames played onto E / H face slot f (the recordings of the H / E
 * face of that slot), NULL for opened faces: the arguments of
 * mpic_add_tfsf. */
typedef struct {
    int     reg[6];           /* imin imax jmin jmax kmin kmax */
    int     mode;             /* 0 playForward, 1 record, 2 playBackward, -1 none */
    int     nframes;
    long    vol[12];          /* nodes per face */
    double *dataE[12], *dataH[12];
} mh_tfsf_t;
int  mh_tfsf_load(const char *dir, const mh_domain_t *d, mh_tfsf_t *t);
/* Record mode (TFSF_turnRec / TFSF_shutDown, source/core/em_TFSF.c:273-330,
 * 866-890): mh_tfsf_record_begin writes consistencyCheck.dat and appends
 * the 24 jobs of TFSF_defrag.map (E faces, then H faces); mh_tfsf_record_save
 * writes the one-rank fragments <face>_000.bin (reg_t header, then the
 * frames: E faces 0..11 then H faces 0..11 in 'data', each frames x face
 * nodes) and lastFrameTime.bin.  mh_tfsf_defrag does what TFSFdefrag
 * (source/TFSFdefrag/main.c) does: joins the fragments of every map job into
 * <face>_99999.bin (adding to what the file holds, with the I0/I1 filter and
 * frame shift of filter.cfg), removes fragments and map, and updates
 * lastFrameTime.bin. */
int  mh_tfsf_record_begin(const char *dir, const int reg[6], const mh_domain_t *d);
int  mh_tfsf_record_save(const char *dir, const int reg[6], int nframes, const double *data);
int  mh_tfsf_defrag(const char *dir);
void mh_tfsf_free(mh_tfsf_t *t);

/* ---- RANMAR (source/misc_ranmar.c:78-146) ------------------------------ */

/* Marsaglia-Zaman-Tsang lagged Fibonacci generator (97, 33) combined with
 * an arithmetic sequence, as published by F. James (1990): 24-bit doubles
 * in [0, 1).  Seeds 0 <= ij <= 31328, 0 <= kl <= 30081. */
typedef struct {
    double u[97], c, cd, cm;
    int    i97, j97;
} mh_ranmar_t;
int  mh_ranmar_init(mh_ranmar_t *r, int ij, int kl);
void mh_ranmar(mh_ranmar_t *r, double *out, int len);
/* Normal deviates N(0, sigma^2) by the Box-Muller transform of the stream's
 * consecutive uniform pairs (r cos, r sin of each pair, in that order). */
void mh_ranmar_normal(mh_ranmar_t *r, double *out, long len, double sigma);

/* Last error message of this library. */
const char *mh_error(void);

#ifdef __cplusplus
}
#endif
#endif
