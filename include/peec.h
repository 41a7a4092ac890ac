/*
 * peec.h — C ABI of the MI355X PEEC full-wave path (drop-in for OpenPEEC's include/peec.h).
 *
 * Every struct below is byte-identical to the reference layout (OpenPEEC include/peec.h:37-231),
 * so callers written against the reference (its main.c, mna.c, output.c, ...) link against
 * libpeec_mi355x.so unchanged.  Host members stay plain malloc() memory owned by peec_t and
 * released by peec_free(); device copies live in a side table keyed by the peec_t address
 * (see peec_mi355x.h).  Each prototype cites the reference declaration it replaces.
 */
#ifndef PEEC_MI355X_PEEC_H
#define PEEC_MI355X_PEEC_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "complex.h"

#ifdef __cplusplus
extern "C" {
#endif

#define PROGRAM "OpenPEEC-MI355X"
#define VERSION_MAJOR 1
#define VERSION_MINOR 0
#define VERSION_BUILD 0

#define FN_LOG  "peec.log"
#define FN_CSV  "zin.csv"
#define FN_DIST "dist.csv"
#define FN_FAR  "far.csv"
#define FN_PW   "pw.csv"
#define FN_TRAN "tran.csv"

/* Physical constants, evaluated exactly as the reference does (peec.h:31-35):
 * PI = 4 atan(1), MU0 = 4 PI 1e-7, EPS0 = 1 / (C0 C0 MU0). */
#define PI   (4.0 * atan(1.0))
#define C0   (2.99792458e8)
#define MU0  (4.0 * PI * 1e-7)
#define EPS0 (1.0 / (C0 * C0 * MU0))
#define EPS  (1e-12)

#define NAMELEN 32

/* lumped elements; node ids are non-negative, 0 = ground / infinity */
typedef struct { int n1, n2; double val; } rc_t;
typedef struct { char name[NAMELEN]; int n1, n2; double val; } ind_t;
typedef struct { char name1[NAMELEN], name2[NAMELEN]; int l1, l2; double k; } mut_t;
typedef struct { int n1, n2, isvsrc, ibr; double amp, phase; } src_t;
typedef struct { int n1, n2; double z0; } port_t;

/* conductor cross-section kinds */
#define SHAPE_ROUND 0   /* round wire, radius a                        */
#define SHAPE_BAR   1   /* rectangular wire, thin-wire GMD radius       */
#define SHAPE_PLATE 2   /* ribbon / volume cell of a plate or brick     */
#define SHAPE_POLY  3   /* planar polygon cell of a quad / disk panel   */

/* vertex capacity of one polygon cell; part of the ABI (seg_t.pv) */
#define POLY_MAX 64

typedef struct {
	double x1[3], x2[3];
	int    shape;
	double radius, width, thick, sigma;
	int    ndiv;
} wire_t;

#define PANEL_QUAD 0
#define PANEL_DISK 1
typedef struct {
	int    kind;
	double v[12];            /* quad corners, cyclic order      */
	double org[3], nrm[3];   /* disk centre and normal          */
	double radius;
	int    ndiva, ndivb;     /* quad grid / disk (nring, nsec)  */
	double thick, sigma;
	int    grading;
} panel_t;

typedef struct {
	double org[3], ea[3], eb[3];
	double thick, sigma;
	int    ndiva, ndivb, ndivt;
} plate_t;

typedef struct {
	double org[3], ea[3], eb[3];
	double thick, epsr, tand;
	double epsinf, frelax;
	int    ndiva, ndivb, ndivt;
} diel_t;

/* One discretised cell: a current branch (p->seg) or a charge cell (p->chg).
 * wid == 0: thin filament; wid > 0: ribbon of width wid across unit vector wv;
 * vol == 1: rectangular bar wid x thick; npv > 0: planar polygon pv[0..npv) fanned from papex. */
typedef struct {
	int    n1, n2;
	double x1[3], x2[3];
	double len;
	int    shape;
	double radius, width, thick;
	double wid, wv[3];
	int    vol;
	int    npv;
	double papex[3];
	double pv[3 * POLY_MAX];
	double aL, aP;           /* equivalent radii: inductive (GMD) / capacitive */
	double area, perim;
	double sigma;
	double res;              /* DC resistance len / (sigma area) */
	int    diel;
	double cexc;
	double gexc;
	double gfac;
	double epss, epsinf, frelax;
} seg_t;

typedef struct {
	char   title[BUFSIZ];

	int    nres, ncap, nind, nmut, nsrc, nport, nwire, nplate, npanel, ndiel, nnodexyz;
	plate_t *plate;
	panel_t *panel;
	diel_t *diel;
	rc_t   *res, *cap;
	ind_t  *ind;
	mut_t  *mut;
	src_t  *src;
	port_t *port;
	wire_t *wire;
	int    *ncid;
	double *ncxyz;

	int    nseg;
	seg_t  *seg;
	d_complex_t *lp;         /* [nseg*nseg] partial inductance, row-major */
	int    lpwarn;

	int    nchg;
	seg_t  *chg;
	int    *chgnode;
	int    ncell;
	int    *cellid;
	int    *cellof;
	double *carea;
	int    *csoff;
	int    *csidx;
	double *cpt;
	d_complex_t *cmat;       /* [ncell*ncell] C = P^-1 (dense path only) */
	int    clogged;

	int    ngnode;
	int    *gid;
	double *gxyz;

	int    maxid;
	int    nnode;
	int    refnode;
	int    *nodemap;
	int    offL, offV, offS;
	int    offQ;
	int    nunknown;

	double f0, f1;
	int    nfreq;
	int    flog;
	double nodetol, gmin;
	int    skin;
	int    capacitance;
	int    retardation;
	int    accel;
	int    dist;
	int    gp;
	double gpz;
	int    ffnth, ffnph;
	int    pw;
	double pwth, pwph;
	int    pwpol;
	double pwamp, pwphase;
	int    tran;
	double tranatt;
	int    grading;
	int    compress;
	double ctol;
	struct hmat_t *hlp;

	d_complex_t *zin;        /* [nport*nfreq]           */
	d_complex_t *zmat;       /* [nfreq*nport*nport]     */
	d_complex_t *smat;       /* [nfreq*nport*nport]     */
	d_complex_t *segi;       /* [nfreq*nport*nseg]      */
	d_complex_t *cellq;      /* [nfreq*nport*ncell]     */
	d_complex_t *voc;        /* [nport*nfreq]           */
	d_complex_t *segipw;     /* [nfreq*nseg]            */
} peec_t;

#define ZIDX(p, ifreq, i, j) ((size_t)((ifreq) * (p)->nport + (i)) * (p)->nport + (j))
#define DIDX(p, ifreq, j, m) ((size_t)((ifreq) * (p)->nport + (j)) * (p)->nseg + (m))
#define QIDX(p, ifreq, j, m) ((size_t)((ifreq) * (p)->nport + (j)) * (p)->ncell + (m))

/* sweep point ifreq (reference peec.h:241-248): linear, or geometric when flog */
static inline double freq_at(const peec_t *p, int ifreq)
{
	if (p->nfreq <= 1) return p->f0;
	if (p->flog) return p->f0 * pow(p->f1 / p->f0, (double)ifreq / (p->nfreq - 1));
	return p->f0 + (p->f1 - p->f0) * ifreq / (p->nfreq - 1);
}

/* ---- deck and geometry (reference peec.h:251-258; callers of the hot path) ---- */
int  tokenize(char *str, const char *tokensep, char *token[], int maxtoken);   /* utils.c:10       */
int  input_data(FILE *fp, peec_t *p);                                           /* input_data.c:38  */
void peec_free(peec_t *p);                                                      /* input_data.c:615 */
int  wire_build(peec_t *p, FILE *fp_log);                                       /* wire.c:348       */

/* ---- element kernels (reference peec.h:260-275, 298-312); evaluated on the GPU ---- */
double rect_potential(const seg_t *s, const double *pt);                        /* surface.c:76   */
double ribbon_static(const seg_t *s1, const seg_t *s2, int nsub, int nq);       /* surface.c:102  */
d_complex_t ribbon_corr(const seg_t *s1, const seg_t *s2, double kw, int nsub); /* surface.c:134  */
int    bar_use_hl(const seg_t *s1, const seg_t *s2);                            /* volume.c:92    */
double bar_pair(const seg_t *s1, const seg_t *s2);                              /* volume.c:114   */
double poly_area(const seg_t *s);                                               /* polygon.c:60   */
double poly_potential(const seg_t *s, const double *pt);                        /* polygon.c:68   */
double poly_static(const seg_t *s1, const seg_t *s2, int nsub, int n7);         /* polygon.c:283  */
d_complex_t poly_corr(const seg_t *s1, const seg_t *s2, double kw);             /* polygon.c:305  */

/* ---- partial elements (reference peec.h:297-312) ---- */
void   seg_mirror(const seg_t *s, double gpz, seg_t *out);                      /* partial.c:352 */
d_complex_t lp_entry(const peec_t *p, int i, int j, double kw);                 /* partial.c:371 */
double neumann_self(double l, double a);                                        /* partial.c:55  */
double neumann_pair(const seg_t *s1, const seg_t *s2, double a1, double a2);    /* partial.c:102 */
d_complex_t neumann_self_k(const seg_t *s, double a, double kw);                /* partial.c:320 */
d_complex_t neumann_pair_k(const seg_t *s1, const seg_t *s2, double a1, double a2, double kw); /* partial.c:256 */
double lp_self(double l, double a);                                             /* partial.c:331 */
void   lp_fill(peec_t *p, double f, FILE *fp_log);                              /* partial.c:428 */

int    pot_fill(peec_t *p, double f, FILE *fp_log);                             /* potential.c:177 */
int    pot_cells(peec_t *p);                                                    /* potential.c:168 */
d_complex_t pot_entry(const peec_t *p, int i, int j, double kw);                /* potential.c:124 */

/* skin effect (reference peec.h:315-317, skin.c) is not part of this path: the deck reader
 * refuses skineffect = 1 and the zint_* functions are not exported. */

/* ---- H-matrix compression (reference peec.h:277-296); device resident ---- */
struct hmat_t;
typedef d_complex_t (*hmat_entry_fn)(void *ctx, int i, int j, double kw);
/* ACA and the near blocks run on the GPU.  The library's own lp_entry / pot_entry are evaluated
 * there too; any other (re-entrant) callback is evaluated once per entry on the host, in
 * parallel, and its n^2 entries uploaded (NULL when they exceed PEEC_HM_GEN_GB GiB, default 16). */
struct hmat_t *hmat_build_gen(int n, const double *pts, hmat_entry_fn fn, void *ctx, double kw,
	double tol, const char *label, FILE *fp_log);                                /* hmatrix.c:403 */
struct hmat_t *hmat_build(peec_t *p, double f, FILE *fp_log);                   /* hmatrix.c:616 */
struct hmat_t *hmat_build_pot(peec_t *p, double f, FILE *fp_log);               /* potential.c:158 */
void hmat_free(struct hmat_t *h);                                               /* hmatrix.c:636 */
void hmat_matvec(const struct hmat_t *h, const d_complex_t *x, d_complex_t *y); /* hmatrix.c:680 */
int  hmat_nleaf(const struct hmat_t *h);                                        /* hmatrix.c:807 */
const d_complex_t *hmat_leaf_block(const struct hmat_t *h, int r, const int **idx, int *m); /* hmatrix.c:812 */
double hmat_memory_mb(const struct hmat_t *h);                                  /* hmatrix.c:656 */
double hmat_dense_mb(int n);                                                    /* hmatrix.c:675 */

/* ---- MNA (reference peec.h:319-352) ---- */
typedef struct {
	int *ri, *ci;
	d_complex_t *val;
	int  nnz, cap;
} mna_sparse_t;                                                                 /* peec.h:322-326 */

typedef struct {
	int off, n;
	const struct hmat_t *h;
	d_complex_t scale;
} hblock_t;                                                                     /* peec.h:334-338 */

#define HBLK_MAX 2

int  mna_numbering(peec_t *p, FILE *fp_log);                                    /* mna.c:16  */
void mna_sparse(const peec_t *p, double f, mna_sparse_t *sp);                   /* mna.c:317 */
void mna_sparse_free(mna_sparse_t *sp);                                         /* mna.c:326 */
void mna_apply(const peec_t *p, const mna_sparse_t *sp, const hblock_t *blk, int nblk,
	const d_complex_t *x, d_complex_t *y, d_complex_t *work);                     /* mna.c:345 */
void mna_rhs_port(const peec_t *p, int iport, d_complex_t *b);                  /* mna.c:367 */
void mna_rhs_sources(const peec_t *p, d_complex_t *b);                          /* mna.c:471 */
void mna_rhs_planewave(const peec_t *p, double f, d_complex_t *b);              /* mna.c:433 */

/* ---- leaf-block + circuit-Schur preconditioner (reference peec.h:358-364); device resident ---- */
struct precond_t;
struct precond_t *pc_build(const peec_t *p, const mna_sparse_t *sp, const hblock_t *blk, int nblk); /* precond.c:139 */
void pc_free(struct precond_t *pc);                                             /* precond.c:78  */
void pc_apply(const struct precond_t *pc, d_complex_t *b);                      /* precond.c:401 */
double pc_memory_mb(const struct precond_t *pc);                                /* precond.c:435 */

/* ---- iterative solvers (reference peec.h:366-371) ---- */
int  gmres_solve(int n, const d_complex_t *a, const d_complex_t *alu, const int *piv,
	d_complex_t *b, double tol);                                                  /* iterative.c:230 */
int  gmres_hmat(const peec_t *p, const mna_sparse_t *sp, const hblock_t *blk, int nblk,
	const struct precond_t *pc, d_complex_t *b, double tol);                      /* iterative.c:285 */

/* ---- dense solver (reference peec.h:354-356) ---- */
int  lu_decomp(int n, d_complex_t *a, int *piv);                                /* lu.c:16 */
void lu_solve(int n, const d_complex_t *a, const int *piv, d_complex_t *b);     /* lu.c:68 */

/* ---- sweep (reference peec.h:374-375) ---- */
int  solve(peec_t *p, FILE *fp_log);                                            /* solve.c:171 */

/* ---- result files and tables (reference peec.h:377-389) ---- */
void output_zin(const peec_t *p, FILE *fp_log);                                 /* output.c:28   */
int  output_csv(const peec_t *p, const char *fn);                               /* output.c:34   */
void output_spara(const peec_t *p, FILE *fp_log);                               /* output.c:85   */
int  output_touchstone(const peec_t *p);                                        /* output.c:103  */
int  output_dist(const peec_t *p, const char *fn);                              /* output.c:229  */
int  output_pw(const peec_t *p, const char *fn, FILE *fp_log);                  /* output.c:178  */
int  output_far(const peec_t *p, const char *fn, FILE *fp_log);                 /* farfield.c:42 */
int  output_tran(const peec_t *p, const char *fn, FILE *fp_log);                /* transient.c:92 */

#ifdef __cplusplus
}
#endif

#endif
