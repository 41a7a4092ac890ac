/*
 * peec_mi355x.h — entry points of libpeec_mi355x.so beyond the reference ABI of peec.h.
 *
 * None of these has a counterpart in OpenPEEC's peec.h; they expose the device-resident sweep
 * (so a caller can time or batch frequency points without the host-array round trips of the
 * fine-grained fill API) and a few conveniences for bindings that cannot pass a FILE*.
 */
#ifndef PEEC_MI355X_EXT_H
#define PEEC_MI355X_EXT_H

#include "peec.h"

#ifdef __cplusplus
extern "C" {
#endif

/* solver forms of the device sweep */
#define PEEC_FORM_PLAIN 0   /* FP64 fill, FP64 LU (MFMA FP64 trailing update)                  */
#define PEEC_FORM_MIXED 1   /* k-factored fill, FP64 LU at a reference frequency applied as FP32
                               inverse factors, FGMRES to a true residual of 1e-10            */
#define PEEC_FORM_PRECISION 2 /* FP64 fill, LU of an FP32 copy (FP64 panels and pivoting, FP32
                               MFMA trailing update), FP64 iterative refinement against the
                               FP64 matrix to a componentwise backward error of 1e-10         */
#define PEEC_FORM_LEARNED 3 /* compressed decks: precond.c's preconditioner kept across
                               frequencies with the learned near-field correction; GMRES true
                               residual 1e-8 (see peec_lpc_*)                                  */

/* per-stage timings of device sweeps (accumulated over points; milliseconds) */
typedef struct {
	int    form;
	double fill_ms, pot_ms, cinv_ms, asm_ms, lu_ms, solve_ms, total_ms;
	double lu_gemm_ms;      /* trailing-update GEMM kernel time inside lu_ms */
	double lu_gemm_flop;    /* algorithmic flops of those GEMMs (8 M N K each) */
	int    lu_gemm_launches;
	double fill_flop;       /* nominal fill flops (reference op count)        */
	int    ising;           /* singular row of the last failure, or -1        */
	int    refine_its;      /* refinement sweeps (precision form)             */
	double berr;            /* worst backward error of the solutions (precision) */
	int    n_lu;            /* factorisations performed (mixed form)          */
	long   gmres_its;       /* GMRES iterations summed over systems (mixed)   */
	double kf_setup_ms;     /* k-factored fill preparation (mixed), in fill_ms */
	long   kf_direct_pairs; /* pairs the k-factored fill evaluates directly   */
	double mv_ms;           /* GMRES products A x: summed time of their dense-block kernel (mixed) */
	double mv_bytes;        /* its algorithmic HBM bytes (blocks read, x, segment sums written) */
	long   mv_launches;
	int    fp64_fallbacks;  /* points the precision form solved in FP64 instead */
	int    pc_builds;       /* preconditioner builds (learned form)           */
	int    lpc_fallbacks;   /* learned solves redone with a fresh M, theta 0  */
	double theta_abs_sum;   /* sum over points of |theta| (learned form)      */
} peec_stats_t;

/* Library identification: "libpeec_mi355x <version> gfx950". */
const char *peec_mi355x_version(void);

/* 1 when a HIP device is present (the library has no CPU compute path). */
int peec_device_available(void);

/* Deck from a file path (input_data on fopen(path)); returns input_data's code. */
int peec_load_deck(const char *path, peec_t *p);

/* Deck from a string (same grammar as a .peec file). */
int peec_load_deck_string(const char *text, peec_t *p);

/* Geometry + capacitive cells + MNA numbering, without any fill (the static fills of the
 * reference's main are only needed for non-retarded sweeps).  fp_log may be NULL. */
int peec_setup(peec_t *p, FILE *fp_log);

/* Device-resident sweep of nfreq points starting at ifreq0 (freq_at indices).  zmat receives
 * nfreq * nport * nport entries ([f][i][j]).  st may be NULL.  Returns 0, 1 (singular P),
 * 2 (singular MNA matrix) or another non-zero code. */
int peec_sweep(peec_t *p, int ifreq0, int nfreq, int form, d_complex_t *zmat, peec_stats_t *st);

/* As peec_sweep, with the mixed form's batch size (frequencies advanced in lock-step sharing one
 * preconditioner; 1..8).  The mixed form keeps its preconditioner factorisation between calls. */
int peec_sweep_ex(peec_t *p, int ifreq0, int nfreq, int form, int batch, d_complex_t *zmat, peec_stats_t *st);

/* Drop the mixed form's cached preconditioner and k-factored fill records (the next mixed sweep
 * factorises and prepares afresh). */
void peec_sweep_reset(void);

/* Nominal FP64 operation count of the dense fills (Lp and charge-cell integrals, upper
 * triangles with ground-plane images) at frequency f: the algorithmic work behind the fill
 * GFLOP/s figure.  Host-side walk of the reference's branch tree; evaluates no integral. */
double peec_fill_flops(const peec_t *p, double f);

/* Lp (nseg x nseg) and P (ncell x ncell, the potential coefficients before inversion) at f,
 * filled on the device and copied to host buffers (P may be NULL).  mode 0: element-by-element
 * evaluation (the plain form's fill); mode 1: the k-factored fill of the mixed form, prepared
 * for the deck's whole frequency band.  Returns 0, or 1 when mode 1 cannot be prepared. */
int peec_fill_matrices(peec_t *p, double f, int mode, d_complex_t *lp, d_complex_t *P);

/* Block structure of a compressed matrix (for tests): counts = {far blocks, near blocks,
 * leaves, n}; perm (n) the cluster order; ranks (far blocks) the ACA ranks, -1 where a far block
 * is stored dense.  perm / ranks may be NULL.  Returns 0, or 1 for a NULL matrix. */
struct hmat_t;
int hmat_info(const struct hmat_t *h, int *counts, int *perm, int *ranks);

/* The H-matrix block partition of n points (host only, no device): counts = {far blocks, near
 * blocks, leaves, n}; perm (n); far / near receive (r0, m, c0, nn) per block when non-NULL
 * (call first with NULL to size them).  Returns 0, or 1 when the leaves do not tile [0, n). */
int hmat_tree_info(int n, const double *pts, int *counts, int *perm, int *far, int *near);

/* lu_decomp (src/lu.c:16-62) with the factors stored in FP32: the matrix is rounded to FP32,
 * factored with FP64 panel arithmetic and pivot search and FP32 MFMA trailing updates, and the
 * factors are returned widened to FP64 in lu_decomp's layout.  The precision form's factorisation;
 * no counterpart in the reference.  Returns -1, or the first zero pivot's row. */
int lu_decomp32(int n, d_complex_t *a, int *piv);

/* Batched sweep of many structures (many decks, each of the reference's solve() loops at
 * once).  Every (structure, frequency) pair is one dense charge-unknown system; frequencies
 * ifreq0 .. ifreq0 + nfreq - 1 of each structure's own sweep.  Fills from k-factored records,
 * factorisation and port solves in persistent batched kernels replayed as hipGraphs; the LU
 * panel width, persistent grid and job order come from a timing autotuner (peec_batch_tune, or
 * the first sweep).  No counterpart in the reference (its callers loop over decks); each
 * structure's Z equals what peec_sweep gives for it.  Structures must be set up
 * (peec_setup), dense (compression 0), with 1..8 ports and at most 4096 charge-form unknowns. */
typedef struct peec_batch peec_batch_t;
typedef struct {
	double total_ms, tune_ms;
	double fill_ms;         /* fill graphs (and element-by-element fallbacks)  */
	double lu_ms;           /* stamp + factor + solve graphs                   */
	double stamp_ms;        /* host stamps and uploads                         */
	double lu_flop;         /* 8/3 N^3 + 8 N^2 per port, summed over jobs      */
	long   jobs;
	int    chunks;
	int    nb, wg_per_cu, order, tune_trials;
	int    tile_cols;       /* trailing-update tile width (32 rows x tile_cols) */
	int    direct_structures; /* filled element by element (no usable records) */
	long   direct_pairs;    /* pair evaluations by the element code (no convergent expansion) */
	long   refine_sweeps;   /* precision form: refinement sweeps over all jobs and ports */
	double berr_max;        /* precision form: worst componentwise backward error of the
	                           port solutions accepted from the FP32 factors            */
	int    fp64_fallbacks;  /* precision form: jobs factored again in FP64             */
} peec_batch_stats_t;
peec_batch_t *peec_batch_create(peec_t **models, int nmodel, int ifreq0, int nfreq);
int  peec_batch_tune(peec_batch_t *b, peec_batch_stats_t *st);
/* PEEC_FORM_PLAIN (default): FP64 factorisation; PEEC_FORM_PRECISION: FP32-stored factorisation
 * (FP64 panels, FP32 MFMA updates) with every port solution refined in FP64 against the FP64
 * system to a componentwise backward error of 1e-10 (FP64 factorisation for a job that does not
 * get there).  Returns 0, or 4 for another form. */
int  peec_batch_set_form(peec_batch_t *b, int form);
/* zmat[s] receives nfreq x nport x nport of structure s; returns 0, 2 (a singular system) or 4 */
int  peec_batch_sweep(peec_batch_t *b, d_complex_t **zmat, peec_batch_stats_t *st);
/* the autotuner's candidates (panel width, tile width, workgroups per CU, order) and their LU
 * times (ms); returns their count */
int  peec_batch_trials(const peec_batch_t *b, int *nb, int *tile_cols, int *wg_per_cu, int *order, double *ms, int cap);
void peec_batch_free(peec_batch_t *b);
/* The batched factorisation alone (tests): nmat row-major matrices of orders n[j],
 * concatenated in a, factored in place as lu_decomp (src/lu.c:16-66) does, one workgroup each
 * (panel width nb = 16, 32 or 64); piv receives each matrix's interchanges, concatenated.  Returns
 * -1, the first matrix with a pivot below 1e-30, or -2 for bad arguments. */
int  peec_batch_lu_decomp(int nmat, const int *n, d_complex_t *a, int *piv, int nb);

/* Learned near-field correction of the compressed solver's preconditioner (no counterpart in
 * the reference; precond.c's method is the fallback): P^-1 = M^-1 - theta M^-1 N M^-1 with N
 * the near-field blocks off the block diagonal and theta from a small network (8 features ->
 * 16 -> 16 -> 2).  The parameters are set per process; without them theta = 0. */
int  peec_lpc_nparams(void);
int  peec_lpc_nfeatures(void);
void peec_lpc_set_model(const double *W);      /* NULL clears */
int  peec_lpc_get_model(double *W);            /* 1 when set  */
/* training samples generated by the solver on a compressed deck: for each (fref, f) pair the
 * features X (8) and the label of the iteration-optimal theta among 0, 0.1, .., 0.6 measured by
 * GMRES, in the network's quadratic loss form coef (4: Re p, Im p, c, d); its (may be NULL)
 * receives the 7 iteration counts of each pair */
int  peec_lpc_samples(peec_t *p, const double *fref, const double *f, int npairs, double *X, double *coef, int *its);
/* the network on the device: forward (theta[2 s], theta[2 s + 1]), mean loss and gradient, and
 * full-batch Adam from a seeded start (loss_hist: steps + 1 values) */
int  peec_lpc_mlp_forward(const double *W, const double *X, int S, double *theta);
int  peec_lpc_mlp_grad(const double *W, const double *X, const double *coef, int S, double *grad, double *loss);
int  peec_lpc_fit(const double *X, const double *coef, int S, int steps, double lr, unsigned long long seed, double *W,
                  double *loss_hist);

/* lu_decomp on a device-resident matrix, allocating nothing.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP stream capture, serving loops,
 * harnesses that forbid hipMalloc).  dA (n x n, row-major, leading dimension lda >= n) and dpiv
 * (n ints) are device pointers; the matrix is factored in place exactly as lu_decomp
 * (src/lu.c:16-66) factors it, dpiv in lu_decomp's convention; stream is a hipStream_t (NULL:
 * the default stream) and is synchronised before returning.  The panel's exchange buffers and
 * the permutation scratch come from one static device array per GPU (668 KiB, held until the
 * process ends), so n is at most 32768 (-2 with a message beyond it); calls that overlap on one
 * GPU race on that array.  Returns -1, the first zero pivot's row, or -2 for invalid arguments. */
int lu_decomp_device_noalloc(int n, d_complex_t *dA, int lda, int *dpiv, void *stream);

/* Sizes for bindings */
size_t peec_sizeof_peec(void);
size_t peec_sizeof_seg(void);

#ifdef __cplusplus
}
#endif

#endif
