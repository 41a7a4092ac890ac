/*
 * peec_internal.h — internal C interface between the C host code (deck, geometry, MNA, sweep
 * driver, ABI wrappers) and the HIP runtime layer (device mirror of a peec_t, fill kernels,
 * dense LU, assembly).  Not part of the public ABI.
 */
#ifndef PEEC_MI355X_INTERNAL_H
#define PEEC_MI355X_INTERNAL_H

#include <stddef.h>
#include "peec.h"
#include "peec_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------------------------- */
/* Packed cell record as the fill kernels see it (host-built, 256 bytes).                    */
/* Every derived quantity is computed on the host with the reference's own expression, so a  */
/* kernel reading t[] gets bit-for-bit the value the reference recomputes inline.            */
typedef struct {
	double x1[3];     /*  0 */
	double x2[3];     /*  3 */
	double d[3];      /*  6  x2 - x1                      */
	double t[3];      /*  9  (x2 - x1) / len              */
	double c[3];      /* 12  0.5 (x1 + x2)                */
	double wv[3];     /* 15  ribbon transverse unit vector */
	double nv[3];     /* 18  t x wv                        */
	double len;       /* 21 */
	double wid;       /* 22 */
	double thick;     /* 23 */
	double aL;        /* 24 */
	double aP;        /* 25 */
	double papex[3];  /* 26 */
	int    npv;       /* 29 (4 ints in 2 doubles) */
	int    poff;      /*    offset into the polygon pools */
	int    vol;       /* 30 */
	int    kind;      /*    CELL_WIRE / CELL_RIBBON / CELL_BAR / CELL_POLY */
	double pad;       /* 31 */
} dcell_t;

#define CELL_WIRE   0
#define CELL_RIBBON 1
#define CELL_BAR    2
#define CELL_POLY   3

/* per-edge polygon data (edge k from pv[k] to pv[k+1]) for the Wilton edge sum */
typedef struct {
	double a[3];      /* start vertex                    */
	double sh[3];     /* unit edge vector                */
	double mh[3];     /* sh x nh (in-plane edge normal)  */
	double el;        /* edge length                     */
	double b[3];      /* end vertex                      */
	double at;        /* area of fan triangle (papex, a, b) */
} dedge_t;

/* per-polygon unit normal (nh) and Newell norm n2 (n2 <= 0: degenerate) */
typedef struct {
	double nh[3];
	double n2;
} dpnorm_t;

/* host-side packing (fill_pack.c) */
int cells_pack(const seg_t *src, int n, int mirror, double gpz,
	dcell_t **cells, dedge_t **edges, dpnorm_t **pnorm, int *nedge);

/* ---------------------------------------------------------------------------------------- */
/* H-matrix geometry (hmat_host.c): cluster tree over a permutation of the cells, far/near    */
/* block lists, leaf-row grouping.  Frequency independent.                                   */
#define HM_LEAF  32
#define HM_ETA   2.0
#define HM_MAXIT 200
#define HM_PROBE 4
/* column check at ACA termination (not in the reference): the block is taken as not converged
 * when the least-represented unused column's residual exceeds HM_COLCHECK tol ||A_k||_F */
#define HM_COLCHECK 1e3
typedef struct {
	int n;
	int *perm;                       /* cluster order -> cell index */
	int nc, cap;
	int *lo, *hi, *left, *right;     /* cluster [lo, hi) of perm, children (-1: leaf) */
	double *cen, *rad;               /* bounding-box centre [3 nc], half diagonal */
	int nfar, nnear;
	int *fr, *fc;                    /* far blocks (row cluster, column cluster) */
	int *nr, *nc2;                   /* near blocks */
	int nleaf;
	int *leafid, *rfoff, *rfidx, *rnoff, *rnidx;
} hm_tree_t;
int  hm_tree_build(hm_tree_t *t, int n, const double *pts);   /* 0 ok */
void hm_tree_free(hm_tree_t *t);

/* ---------------------------------------------------------------------------------------- */
/* Device mirror of one peec_t (gpu_runtime.hip).                                            */
typedef struct pgpu pgpu_t;
typedef struct dgmres_ws dgmres_ws;   /* GMRES workspace of dgmres.hip */

pgpu_t *pgpu_get(const peec_t *p);          /* find or create; uploads cells on creation */
void    peec_gpu_release(peec_t *p);        /* drop the mirror of p (called by peec_free) */
int     pgpu_device_ok(void);               /* 1 if a HIP device is usable                */
void    pgpu_fatal(const char *what);       /* print and abort: no host fallback exists   */
/* fn runs once at process exit, before the HIP runtime's own teardown (hooks last-registered
 * first, then every device mirror is released); for process-lifetime device state */
void    pgpu_at_exit(void (*fn)(void));

/* Fill (results stay on the device; *_to_host copies them into the peec_t arrays).
 * Return 0 on success. */
int  pgpu_fill_lp(pgpu_t *g, double kw);
int  pgpu_fill_pot(pgpu_t *g, double kw);   /* P (aggregated, scaled) into the device P buffer */
int  pgpu_cmat(pgpu_t *g, int *ising);      /* C = P^-1 on the device; *ising = -1 or row     */
void pgpu_lp_to_host(pgpu_t *g, d_complex_t *lp);
void pgpu_cmat_to_host(pgpu_t *g, d_complex_t *cmat);
void pgpu_lp_from_host(pgpu_t *g, const d_complex_t *lp);
void pgpu_cmat_from_host(pgpu_t *g, const d_complex_t *cmat);

/* k-factored fill (kfill.hip): records for wavenumbers in [klo, khi], then Lp / P at any kw
 * of the band straight into device buffers.  setup returns 1 when the records exceed the
 * memory budget. */
int  pgpu_kf_setup(pgpu_t *g, double klo, double khi);
/* records of the frequency band [f0, f1] for the compressed builds (no-op when static or PEEC_KF=0) */
void pgpu_kf_band(pgpu_t *g, const peec_t *p, double f0, double f1);
void pgpu_kf_fill(pgpu_t *g, double kw, d_complex_t *dlp, d_complex_t *dP);
/* Lp (which 0) or P (which 1) alone at kw into out (cell order); 1 when the band's records do
 * not cover kw or P's sub-cell matrix would not fit in a quarter of free memory */
int  pgpu_kf_fill_kind(pgpu_t *g, int which, double kw, d_complex_t *out);
/* the same for nk frequencies at once (the records read once): dlp[k], dP[k] at kw[k] */
void pgpu_kf_fill_multi(pgpu_t *g, int nk, const double *kw, d_complex_t *const *dlp, d_complex_t *const *dP, int upper);
int  pgpu_kf_direct_pairs(pgpu_t *g);
void pgpu_kf_reset_all(void);     /* every device mirror prepares its records afresh */
/* Lp and P at kw (mode 0 element code, mode 1 k-factored, set up beforehand) to host buffers */
void pgpu_fill_to_host(pgpu_t *g, double kw, int mode, d_complex_t *lp, d_complex_t *P);

/* single entries (lp_entry / pot_entry through the same device code) */
d_complex_t pgpu_lp_entry(pgpu_t *g, int i, int j, double kw);
d_complex_t pgpu_pot_entry(pgpu_t *g, int i, int j, double kw);

typedef peec_stats_t pgpu_stats_t;

/* one point of solve() on the dense path: nr right-hand sides (host, n each) -> solutions x
 * (nr x n); accel 1 keeps an LU across points and solves by GMRES with it (acceleration = 1);
 * q: cell charges of the first nq solutions.  0, 1 (singular P) or 2 (singular A). */
typedef struct { int nlu, ngm; long itsum; } pgpu_accel_stats_t;
int  pgpu_solve_point(pgpu_t *g, const peec_t *p, int ifreq, int nr, const d_complex_t *rhs, d_complex_t *x, int nq,
	d_complex_t *q, int accel, pgpu_accel_stats_t *as);
void pgpu_accel_reset(pgpu_t *g);   /* forget the kept LU (a new solve() starts fresh) */
/* the same on the compressed path (H matrices, leaf/Schur preconditioner, GMRES to 1e-10 per
 * right-hand side); its: GMRES iterations summed; mem_mb: the point's H-matrix + preconditioner
 * memory.  0, 5 (H-matrix build), 6 (preconditioner) or 4 (GMRES did not converge). */
int  pgpu_solve_point_compressed(pgpu_t *g, peec_t *p, int ifreq, int nr, const d_complex_t *rhs, d_complex_t *x, long *its,
	double *mem_mb, pgpu_stats_t *st);

/* sweep of frequency points, device resident: fills, assembles, factors, solves every port.
 * Writes zmat (nport x nport per point), and optionally segi/cellq rows.  Returns 0 or an error
 * code (1 singular P, 2 singular MNA, 3 refinement stalled); precision = 1 factors an FP32 copy and
 * refines in FP64 (PEEC_FORM_PRECISION). */
int pgpu_sweep(pgpu_t *g, const peec_t *p, int ifreq0, int nfreq, d_complex_t *zmat,
	d_complex_t *segi, d_complex_t *cellq, pgpu_stats_t *st, int precision);

/* accelerated sweep (sweep_fast.hip): charge-unknown system, stale FP32-stored LU as GMRES
 * preconditioner, frequencies in lock-step batches of `batch` systems */
int  pgpu_sweep_fast(pgpu_t *g, const peec_t *p, int ifreq0, int nfreq, d_complex_t *zmat,
	pgpu_stats_t *st, int batch);
void pgpu_fast_reset(void);

/* compressed (compression = 1) sweep (hsolve.hip): H matrices, leaf/Schur preconditioner,
 * GMRES to 1e-10 per port.  Returns 0, 4 (GMRES), 5 (H matrix) or 6 (preconditioner). */
int  pgpu_sweep_compressed(pgpu_t *g, peec_t *p, int ifreq0, int nfreq, d_complex_t *zmat,
	pgpu_stats_t *st);
/* compressed sweep with the learned preconditioner correction (hsolve.hip) */
int  pgpu_sweep_learned(pgpu_t *g, peec_t *p, int ifreq0, int nfreq, d_complex_t *zmat, pgpu_stats_t *st, double tol);

/* ---------------------------------------------------------------------------------------- */
/* Dense complex LU on device buffers (lu.hip).  Matrices are row-major, leading dim lda.   */
/* factor in place; dpiv[k] as lu_decomp; dperm (may be NULL) receives the row permutation
 * (row k of the factors is original row dperm[k]).  Returns -1, the singular step, or n when
 * n exceeds the cooperative panel's capacity (256 workgroups x 144 rows on MI355X). */
int  dlu_factor(void *stream, int n, d_complex_t *dA, int lda, int *dpiv, int *dperm, pgpu_stats_t *st);
/* B <- A^-1 B (nrhs columns, row-major, leading dim ldb); dwork holds n * nrhs entries */
void dlu_solve_perm(void *stream, int n, const d_complex_t *dA, int lda, const int *dperm,
	d_complex_t *dB, int ldb, int nrhs, d_complex_t *dwork);
/* explicit inverse of the unit-lower (lower = 1) or upper part of packed LU factors (n x n,
 * leading dimension n out); dinv: ceil(n/64) 64 x 64 blocks, scr: (ceil(n/2) + 64)^2 entries */
void dlu_tri_inverse(void *stream, int n, const d_complex_t *LU, int lda, int lower, d_complex_t *Out,
	d_complex_t *scr, d_complex_t *dinv);
/* the same on FP32 storage (slu_factor's factors); Out, scr, dinv float2 */
void slu_tri_inverse(void *stream, int n, const void *LU, int lda, int lower, void *Out, void *scr, void *dinv);
/* the same with FP32 storage (float2 pairs; FP64 panels, FP32-MFMA trailing updates), and the
 * solve of FP64 right-hand sides with those factors */
int  slu_factor(void *stream, int n, void *dA, int lda, int *dpiv, int *dperm, pgpu_stats_t *st);
void slu_solve_perm(void *stream, int n, const void *dA, int lda, const int *dperm,
	d_complex_t *dB, int ldb, int nrhs, d_complex_t *dwork);
/* inverses of the 64 x 64 diagonal blocks of slu_factor's factors (FP64, ceil(n/64) blocks each),
 * and the solve that uses them for each block step (nrhs <= 4) */
void slu_diag_inverse(void *stream, int n, const void *dA, int lda, d_complex_t *dinvL, d_complex_t *dinvU);
void slu_solve_perm_di(void *stream, int n, const void *dA, int lda, const int *dperm, d_complex_t *dB, int ldb, int nrhs,
	d_complex_t *dwork, const d_complex_t *dinvL, const d_complex_t *dinvU);
/* one right-preconditioned GMRES(60) cycle on a dense device matrix A with the LU factors LU /
 * perm (dlu_factor's) of a nearby matrix as preconditioner (dgmres.hip); b (device) becomes the
 * solution when the cycle reaches tol |b|.  Returns iterations or -1.  ws: dgmres_ws_create(n). */
dgmres_ws *dgmres_ws_create(int n);
void dgmres_ws_free(dgmres_ws *ws);
int  dgmres_cycle(void *stream, dgmres_ws *ws, int n, const d_complex_t *A, int lda, const d_complex_t *LU, const int *perm,
	d_complex_t *b, double tol);
/* C[M x N] -= A[M x K] B[K x N] (complex FP64 MFMA) */
void dlu_zgemm_sub(void *stream, int M, int N, int K, const d_complex_t *A, long lda,
	const d_complex_t *B, long ldb, d_complex_t *C, long ldc);

/* MNA sparse stamps (host, mna_host.c), coalesced in the reference's summation order */
typedef struct {
	int n;                   /* number of distinct entries */
	int *ri, *ci;
	d_complex_t *val;
} stamps_t;
void mna_stamps(const peec_t *p, double f, stamps_t *st);
/* the dense numbering extended by one charge unknown per capacitive cell after the last
 * branch current (the system of the accelerated and batched sweeps) */
void mna_stamps_charge(const peec_t *p, double f, stamps_t *st);
void mna_stamps_free(stamps_t *st);

#ifdef __cplusplus
}
#endif

#endif
