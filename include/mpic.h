/*
 * mpic.h -- C ABI of the MI355X PIC time step (libmpic.so).
 *
 * Replaces, for one GPU holding the whole domain, the physics API that the
 * reference driver source/core/main.c:353-476 calls (cvarin/mandor3):
 *
 *   em_init / em_HHalfStep / em_HStep / em_EStep_start / em_EStep_finish /
 *   em_energy                          source/core/em.h:32-39
 *   plasma_init / plasma_move / plasma_temperature / jbc_finish /
 *   plasma_rho                         source/core/plasma.h:82-87
 *   plasma_load / plasma_save          source/core/plasma_IO.c:227,273
 *   gauss_before / gauss_after         source/core/plasma_gaussTest.h:21-22
 *
 * The reference keeps E, H, J and the markers in process globals
 * (source/core/main.c:90-91, source/core/plasma.c:40-43); here they live in
 * device memory behind an opaque mpic_run_t, so several runs can coexist on
 * one device (ensemble / serving).  Plain pointers and sizes only.
 *
 * Error convention: the reference logs and exits (ENSURE/DIE,
 * source/log.h:45-60); every call here returns MPIC_OK (0) or a negative
 * code, with a message in mpic_last_error().  The C host driver maps a
 * failure to "log + exit(1)" like the reference.
 *
 * Threads: each run has its own stream and state, the last error is per
 * thread, and the fused ensemble cache is guarded by a lock, so one thread
 * may create, load, initialise and close runs while another advances other
 * runs (the serving loop does).  Calls naming the same run must not overlap.
 *
 * Layouts at this boundary are the reference's binData layouts:
 *   fields  node block [0,n] per axis, interleaved {x,y,z} doubles, k
 *           fastest (the payload of sys_E/H_*.bin, source/type_mesh.c:285-311)
 *   markers marker_t records (source/type_marker.h:12-25): x y z vx vy vz
 *           (double), rho qDivM (float), optional int id; recsize 56 or 64.
 */
#ifndef MPIC_H
#define MPIC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPIC_OK            0
#define MPIC_ERR_ARG      -1
#define MPIC_ERR_DEVICE   -2
#define MPIC_ERR_ORDER    -3   /* stage called out of order (emStage) */
#define MPIC_ERR_CAPACITY -4
#define MPIC_ERR_STATE    -5

/* Boundary codes, reference enum source/misc_parameters.h:38-42. */
#define MPIC_BC_PERIODIC 0
#define MPIC_BC_MIRROR   1
#define MPIC_BC_OPEN     2     /* first-order Mur for fields, deletion for markers */

/* Arithmetic of a run. */
#define MPIC_FP64        0     /* plain form: FP64 everywhere, reference arithmetic */
#define MPIC_FP32_SPLIT  1     /* int32 cell + FP32 offset markers, FP32 push/deposit/FDTD,
                                  FP64 J flush and diagnostics */

typedef struct mpic_run mpic_run_t;

/* Domain description: the content of binData/domain.bin
 * (source/misc_parameters.c:98-110) plus the start time of the record. */
typedef struct {
    int    n[3];          /* cells per axis (dmn_max) */
    double h[3];          /* mesh steps */
    double L[3];          /* box lengths */
    double tau;           /* time step */
    int    bc[6];         /* {xmin, ymin, zmin, xmax, ymax, zmax} */
    double time;          /* Time of the loaded record */
    int    precision;     /* MPIC_FP64 or MPIC_FP32_SPLIT */
    long   max_markers;   /* device capacity in markers */
    int    device;        /* HIP device ordinal */
} mpic_config_t;

const char *mpic_last_error(void);
int  mpic_device_count(void);

int  mpic_create(const mpic_config_t *cfg, mpic_run_t **run);
int  mpic_destroy(mpic_run_t *run);
/* Device memory the run holds (bytes); used for HBM admission. */
int  mpic_footprint(const mpic_config_t *cfg, size_t *bytes);

/* sysIO_loadEM / sysIO_save payloads (host buffers, 3*(n+1)^3 doubles). */
int  mpic_set_fields(mpic_run_t *run, const double *E, const double *H);
int  mpic_get_fields(mpic_run_t *run, double *E, double *H);
/* J after jbc_finish, nodes [-2,n+2], interleaved, k fastest. */
int  mpic_get_J(mpic_run_t *run, double *J);

/* plasma_load: records in file order, before mpic_init (a load after it
 * returns MPIC_ERR_STATE).  plasma_save: live markers in device order; *n
 * in: capacity of 'rec', out: markers written. */
int  mpic_markers_load(mpic_run_t *run, const void *rec, long n, int recsize);
/* plasma_load's buffered read (source/core/plasma_IO.c:227-264: BUFFER_N
 * records at a time, file after file): appends n records after the markers
 * already loaded, so a record larger than host memory streams in pieces.
 * Before mpic_init only; ids default to the position in the whole record. */
int  mpic_markers_append(mpic_run_t *run, const void *rec, long n, int recsize);
int  mpic_markers_save(mpic_run_t *run, void *rec, long *n, int recsize);
long mpic_marker_count(mpic_run_t *run);

/* Focused-laser soft source (source/scpic/em_sources.c, em_mirror.c:667-717):
 * A = field amplitude (2*pi*a0), T0 = peak arrival at focus, delay = travel
 * time from the source plane to the focus, dt = 1/e duration, phi0 = phase,
 * forward = +x propagation, plane = source plane index (5 in the
 * reference).  map: E map then H map, each (n[1]-9)*(n[2]-9) records of
 * {y_cos, y_sin, z_cos, z_sin}, k fastest. */
int  mpic_add_laser(mpic_run_t *run, double A, double T0, double delay,
                    double dt, double phi0, int forward, int plane,
                    const double *map);

/* Hard Gaussian spot source (source/core/em_gaussSpot.c:42-156): p = the 11
 * values of .gaussSpot.cfg in file order {ampl, frequency [omega_0], x, y, z,
 * width, duration, front offset, plateau, Ey flag, Ez flag}; Ey, Ez on the
 * plane x/h + 1/2 are set at the end of every E step. */
int  mpic_add_gauss_spot(mpic_run_t *run, const double *p);

/* TFSF interface playback, replacing TFSF_init/TFSF_turnPlay and the
 * post-step source terms (source/core/em_TFSF.c:400-545, 666-860) for one
 * GPU: reg = the interface box of TFSF.cfg {imin, imax, jmin, jmax, kmin,
 * kmax}; mode 0 playForward or 2 playBackward; nframes = the recorded frame
 * count (lastFrameTime.bin).  dataE[f] are the frames played onto E face f
 * (the recording of the H face of the same slot), dataH[f] those played onto
 * H face f (the recording of E face f), each nframes x face nodes doubles in
 * the file's order (tangential axis ep outer, eq inner); NULL = face opened
 * (openFaces.cfg).  Slot f = axis*4 + top*2 + comp (mf_face, :79).  The
 * recording is copied to device memory; faces must keep 5 cells from the
 * domain edges.  Add before mpic_init. */
int  mpic_add_tfsf(mpic_run_t *run, const int reg[6], int mode, int nframes,
                   const double *const dataE[12], const double *const dataH[12]);

/* VSP_configure (source/core/plasma_VSP.c:18-27; core.out's --VSP:weight
 * option, source/core/main.c:314-320): the (beta, alpha, beta) convolution,
 * beta = (1 - alpha)/2, of J at the end of mpic_j_finish and of the charge
 * density of mpic_rho / the Gauss test, along z, y, x.  alpha in (-0.1,
 * 1.1]; 1 (the default) switches it off.  Between steps only. */
int  mpic_set_vsp(mpic_run_t *run, double alpha);

/* em_init + plasma_init (caps, Mur buffers, ghost nodes, walls applied to
 * the loaded markers). */
int  mpic_init(mpic_run_t *run);

/* The stages of one time step, in the order of source/core/main.c:365-425. */
int  mpic_em_h_half(mpic_run_t *run);                           /* em_HHalfStep   */
int  mpic_em_energy(mpic_run_t *run, double *W_E, double *W_M); /* em_energy      */
int  mpic_push_deposit(mpic_run_t *run);                        /* plasma_move    */
int  mpic_temperature(mpic_run_t *run, double *WTx, double *WTy, double *WTz);
int  mpic_em_h_step(mpic_run_t *run);                           /* em_HStep       */
int  mpic_em_e_start(mpic_run_t *run);                          /* em_EStep_start */
int  mpic_j_finish(mpic_run_t *run);                            /* jbc_finish     */
int  mpic_em_e_finish(mpic_run_t *run);                         /* em_EStep_finish; Time += tau */

/* A whole main-loop iteration; diag (may be NULL) = {Time, W_E, W_M, WTx,
 * WTy, WTz} as handed to wDensity_addPoint (source/core/diag_WDensity.c:136).
 * With diag == NULL nothing is copied back and the step stays asynchronous. */
int  mpic_step(mpic_run_t *run, double *diag);
/* n steps enqueued on the run's stream without host synchronisation; with
 * graph replay on, as replays of the captured step. */
int  mpic_advance(mpic_run_t *run, int nsteps);
/* The energy history of the steps run by mpic_step(run, NULL),
 * mpic_advance and mpic_batch_advance, which compute em_energy and
 * plasma_temperature on the device every step (source/core/main.c:371,383)
 * into a ring of 1024 rows: the last min(max_rows, 1024) of those steps,
 * oldest first, each {Time, W_E, W_M, WTx, WTy, WTz} as wDensity_addPoint
 * records it (source/core/diag_WDensity.c:136); *nrows rows written. */
int  mpic_diag_history(mpic_run_t *run, double *rows, int max_rows, int *nrows);
/* Graph replay (no counterpart in the reference: its loop calls the stages
 * one by one): on = 1 captures one step per marker-buffer parity into HIP
 * graphs at the next advance and replays them; results are those of the
 * stage calls.  Kernel timing (mpic_last_timing) is recorded only with
 * replay off. */
int  mpic_set_graphs(mpic_run_t *run, int on);
/* An ensemble of runs advanced together (no counterpart in the reference,
 * which runs one simulation per job): one step of every run is recorded,
 * the launches of each kernel kind are fused across the runs into one
 * launch over a table of per-run arguments and block ranges, and the fused
 * step is captured in one graph (per marker-buffer phase) and replayed on
 * the ensemble's own stream, ordered after and before the runs' streams.
 * The fused step is kept for the next call with the same runs.  With
 * MPIC_BATCH_STREAMS=1, or runs on several devices, every run replays its
 * own captured step on its own stream instead.  Every run must be between
 * steps (a step begun through the stage calls: MPIC_ERR_ORDER). */
int  mpic_batch_advance(mpic_run_t *const *runs, int nruns, int nsteps);
/* Picks the launch shape of the FP32 push -- markers per work item (a
 * tile's markers split into even items of at most 2048 or 1024, or one item
 * per tile), wavefronts per workgroup (1, 2 or 4 sharing one tile's staged
 * fields and J block) and workgroup count -- by timing 'steps' real steps
 * per candidate (the steps advance the run; the choice changes results only
 * by the summation order of J).  chosen (may be NULL) receives {workgroups,
 * wavefronts per workgroup, markers per item (0: one item per tile)}.  The
 * work-item buffers may grow by up to an eighth of the marker buffers'
 * size.  Between steps only. */
int  mpic_autotune(mpic_run_t *run, int steps, int *chosen);
int  mpic_sync(mpic_run_t *run);
double mpic_time(mpic_run_t *run);

/* TFSF record mode (TFSF_turnRec / TFSF_preHStep / TFSF_writeMesh,
 * source/core/em_TFSF.c:273-330, 573-624): at the start of every step (in
 * mpic_em_h_half, before the boundary caps) the tangential E on the 12 E
 * faces and H on the 12 H faces of the interface box reg = {imin, imax,
 * jmin, jmax, kmin, kmax} are stored on the device, each value times
 * (1-2 top)(1-2 comp) tau/h of the face's axis, for up to max_frames steps.
 * Call before mpic_init; excludes mpic_add_tfsf on the same run. */
int  mpic_tfsf_record(mpic_run_t *run, const int reg[6], int max_frames);
/* The recording so far: *nframes; vol (may be NULL) the node count of face
 * slot q (E faces q = 0..11, H faces 12..23, slot axis*4 + top*2 + comp);
 * out (may be NULL) every slot's frames in slot order, each nframes x vol[q]
 * doubles (the tangential index ep = (axis+1)%3 outer, eq = (axis+2)%3
 * inner): the payloads of the reference's fragment files. */
int  mpic_tfsf_recorded(mpic_run_t *run, int *nframes, long vol[24], double *out);

/* CIC charge density with boundary fold (plasma_rho), nodes [-2,n+2]. */
int  mpic_rho(mpic_run_t *run, double *rho);
/* Tecplot payloads computed on the device (tecIO_saveVec with
 * tecIO_interpE / tecIO_interpH, source/IO_tecplot.c:131-214): what = 0 E,
 * 1 H, 2 J; out: nodes [0,n] (k fastest), 3 interleaved doubles each.  E and
 * J are the mean of the two values along each component's axis, H the mean
 * of the four around the node. */
int  mpic_node_fields(mpic_run_t *run, int what, double *out);
/* Probes (probe_allocate / probe_postData, source/core/diag_probe.c:171-383):
 * the nodes ijk[3p..3p+2], each in [0, n-1]; a sample writes
 * {Ex, Ey, Ez, Hx, Hy, Hz} per probe, interpolated as mpic_node_fields does,
 * from the fields as they are when it is called (the reference samples
 * after em_HHalfStep). */
int  mpic_set_probes(mpic_run_t *run, const int *ijk, int nprobes);
int  mpic_probe_sample(mpic_run_t *run, double *out);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The samples
 * of mpic_set_probes(ijk, nprobes) + mpic_probe_sample(out) in one call
 * that allocates nothing: the nodes and samples pass through a static
 * device scratch of MPIC_PROBE_AT_MAX probes (1 MiB, held by every GPU the
 * library runs on until the process ends).  More probes: MPIC_ERR_CAPACITY.
 * The run's own probe set is left as it is.  All calls on one GPU share the
 * scratch: calls that overlap on one GPU (from several host threads) race. */
#define MPIC_PROBE_AT_MAX 16384
int  mpic_probe_sample_at(mpic_run_t *run, const int *ijk, int nprobes, double *out);
/* gauss_before / gauss_after residuals {max|drho/dt + div J|, max|div E - rho|}:
 * call mpic_gauss_before() before mpic_push_deposit and mpic_gauss_after()
 * after mpic_em_e_finish. */
int  mpic_gauss_before(mpic_run_t *run);
int  mpic_gauss_after(mpic_run_t *run, double *res);

/* Per-step marker sort by cell (tile-major cell keys; the order of markers
 * inside one cell is unspecified); drops deleted markers.  counts (may be
 * NULL): per-cell live-marker counts, n0*n1*n2. */
int  mpic_sort(mpic_run_t *run, int32_t *counts);
/* Cell index of every live marker (device order) for parity checks:
 * (i*n1 + j)*n2 + k of floor(x/h). */
int  mpic_marker_cells(mpic_run_t *run, int32_t *cells, long *n);

/* Kernel timing of the last mpic_advance, from HIP events on the run's
 * stream (three per step: before the push, after it, after the deposit;
 * waits for those steps to finish): milliseconds, summed over the steps, in
 * the push kernel, in the rest of the deposit (the FP32 form's J gather; 0
 * for FP64) and in everything else of the step (from the deposit's end to
 * the next push, scaled to all the steps). */
int  mpic_last_timing(mpic_run_t *run, double *push_ms, double *deposit_ms, double *rest_ms);

#ifdef __cplusplus
}
#endif
#endif
