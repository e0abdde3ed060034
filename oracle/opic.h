/*
 * opic.h -- CPU oracle for the Mandor PIC time step (FP64, serial, one rank).
 *
 * TEST INFRASTRUCTURE ONLY.  This is a restatement of the reference CPU path
 * (cvarin/mandor3, source/core/) used as the checker for the HIP product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  It reproduces the reference arithmetic operation for operation
 * (same association, no FMA contraction, IEEE division and sqrt) and the
 * reference's marker bookkeeping (load-order rotation, core/shell partition),
 * so that on the golden fixtures it matches the reference binaries bit for
 * bit.  Every function cites the reference file:line it restates.
 *
 * Parity status per function (how each one is pinned) is listed in opic.c.
 */
#ifndef OPIC_H
#define OPIC_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct opic_s opic_t;

/* Boundary codes, reference enum source/misc_parameters.h:38-42. */
enum { OPIC_BC_PERIODIC = 0, OPIC_BC_MIRROR = 1, OPIC_BC_OPEN = 2 };

/* n: cells per axis (reference dmn_max); L: box lengths; bc[6] ordered
 * {xmin, ymin, zmin, xmax, ymax, zmax} as in domain.bin. */
opic_t *opic_create(const int n[3], const double h[3], const double L[3],
                    double tau, const int bc[6], double time);
void    opic_destroy(opic_t *s);

/* E, H: nodes [0,n] per axis, interleaved xyz, k fastest (the node block of
 * a sys_E/H record, source/type_mesh.c:285-311).  Ghost nodes are filled
 * with NaN like the reference's MC_FILL_MESH_WITH_NAN build. */
void    opic_set_fields(opic_t *s, const double *E, const double *H);
void    opic_get_fields(const opic_t *s, double *E, double *H);

/* Markers as reference marker_t records (56 B, or 64 B with tracer id), in
 * file order; loading applies the reference's per-chunk reordering. */
int     opic_set_markers(opic_t *s, const void *rec, long n, int recsize);
long    opic_marker_count(const opic_t *s);
/* Markers in the order plasma_save writes them (shell, then core). */
void    opic_get_markers(const opic_t *s, void *rec, int recsize);

/* Focused-laser soft source (source/scpic/em_mirror.c:667-717).  map holds
 * 2 * (n[1]+1-10) * (n[2]+1-10) records of 4 doubles {y_cos,y_sin,z_cos,z_sin}:
 * first the E map (drives H), then the H map (drives E). */
int     opic_add_laser(opic_t *s, double A, double T0, double delay,
                       double dt, double phi0, int forward, int plane,
                       const double *map);

/* Hard Gaussian spot source: the 11 values of .gaussSpot.cfg in file order
 * {ampl, frequency [omega_0], x, y, z, width, duration, front offset,
 * plateau, Ey flag, Ez flag} (source/core/em_gaussSpot.c:42-55). */
int     opic_add_gauss_spot(opic_t *s, const double p[11]);

/* TFSF interface playback (source/core/em_TFSF.c): reg = TFSF.cfg's box
 * {imin, imax, jmin, jmax, kmin, kmax}, mode 0 playForward / 2 playBackward,
 * nframes = the recorded frame count (lastFrameTime.bin).  dataE[f] holds the
 * frames played onto E face f (the recording of the H face of slot f),
 * dataH[f] those played onto H face f (the recording of E face f), each
 * nframes x face nodes doubles in file order; NULL = face opened.  Slot f =
 * axis*4 + top*2 + comp, as mf_face (source/core/em_TFSF.c:79). */
int     opic_add_tfsf(opic_t *s, const int reg[6], int mode, int nframes,
                      const double *const dataE[12], const double *const dataH[12]);

/* VSP_configure (source/core/plasma_VSP.c:18-27): the (beta, alpha, beta)
 * convolution of J in jbc_finish and of rho; alpha in (-0.1, 1.1], 1 =
 * off (the default).  Returns -1 for other values. */
int     opic_set_vsp(opic_t *s, double alpha);

int     opic_init(opic_t *s);              /* em_init + plasma_init */
/* One main-loop iteration, source/core/main.c:353-476.
 * diag (may be NULL) receives {Time, W_E, W_M, WTx, WTy, WTz} as passed to
 * wDensity_addPoint; gauss (may be NULL) receives the two residuals of
 * gauss_after (computed only when gauss != NULL). */
int     opic_step(opic_t *s, double *diag, double *gauss);
double  opic_time(const opic_t *s);

/* J after the last step (nodes [-2,n+2], interleaved, k fastest). */
void    opic_get_J(const opic_t *s, double *J);
/* CIC charge density incl. boundary fold (nodes [-2,n+2]). */
void    opic_rho(opic_t *s, double *rho);

#ifdef __cplusplus
}
#endif
#endif
