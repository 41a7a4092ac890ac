/*
 * oracle.h — CPU restatement of OpenPEEC's per-frequency fill + dense solve, in plain C.
 *
 * TEST INFRASTRUCTURE ONLY.  The product (openpeec_amd/libpeec_mi355x.so) never links, loads
 * or calls anything here; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * use liboracle.so, as the checker / CPU baseline.
 *
 * Every function follows the reference's arithmetic in the same order (so that, built with the
 * same compiler and libm, it agrees with the reference bit for bit) and cites the reference
 * source it restates.  It works on the seg_t records of a peec_t (include/peec.h layout) and
 * recomputes every frame inline, independently of the product's host-side packing.
 *
 * Pinning (tests/test_oracle.py): bitwise against oracle/_ref (the reference's own sources)
 * on the sample decks where that build exists; against closed forms and brute-force quadrature
 * (neumann_self, fpar, rect_potential, the thin-wire sum rule, the polygon/rectangle identity,
 * Hoer-Love additivity); and through the sweep against the reference's analytic checks
 * (wire L, dipole Rin, monopole = dipole / 2).
 */
#ifndef PEEC_ORACLE_H
#define PEEC_ORACLE_H

#include "peec.h"

/* element integrals (partial.c, surface.c, volume.c, polygon.c) */
double      orc_neumann_self(double l, double a);
double      orc_neumann_pair(const seg_t *s1, const seg_t *s2, double a1, double a2);
d_complex_t orc_pair_k(const seg_t *s1, const seg_t *s2, double a1, double a2, double kw);
d_complex_t orc_self_k(const seg_t *s, double a, double kw);
double      orc_rect_potential(const seg_t *s, const double *pt);
double      orc_ribbon_static(const seg_t *s1, const seg_t *s2, int nsub, int nq);
d_complex_t orc_ribbon_corr(const seg_t *s1, const seg_t *s2, double kw, int nsub);
int         orc_bar_use_hl(const seg_t *s1, const seg_t *s2);
double      orc_bar_pair(const seg_t *s1, const seg_t *s2);
double      orc_poly_potential(const seg_t *s, const double *pt);
double      orc_poly_static(const seg_t *s1, const seg_t *s2, int nsub, int n7);
d_complex_t orc_poly_corr(const seg_t *s1, const seg_t *s2, double kw);
void        orc_seg_mirror(const seg_t *s, double gpz, seg_t *out);

/* matrix fills: lp[nseg*nseg], c[ncell*ncell] (= P^-1); cell maps built internally */
d_complex_t orc_lp_entry(const peec_t *p, int i, int j, double kw);
void        orc_lp_fill(const peec_t *p, double kw, d_complex_t *lp);
int         orc_pot_fill(const peec_t *p, double kw, d_complex_t *pmat, d_complex_t *c);   /* 0 ok, 1 singular; pmat (P before inversion) and c may be NULL */

/* dense LU (lu.c) */
int  orc_lu_decomp(int n, d_complex_t *a, int *piv);
void orc_lu_solve(int n, const d_complex_t *a, const int *piv, d_complex_t *b);

/* sweep points [ifreq0, ifreq0 + nfreq): zmat[f][i][j]; needs a numbered peec_t
 * (wire_build + mna_numbering) and, for static decks, nothing else.  Returns 0 or 1/2 on a
 * singular P / MNA matrix. */
int  orc_sweep(const peec_t *p, int ifreq0, int nfreq, d_complex_t *zmat);

/* bounded timing sample for a CPU baseline: fills rows [r0, r1) of Lp and of the charge-cell
 * integral matrix at wavenumber kw (no output kept) */
void orc_fill_rows(const peec_t *p, double kw, int r0, int r1, int pot);

/* the learned preconditioner's network (mlp.c; no reference counterpart): parameter count,
 * forward (theta[2 s], theta[2 s + 1]) and mean loss with its gradient (grad may be NULL) */
int    orc_mlp_nparams(void);
void   orc_mlp_forward(const double *W, const double *X, int S, double *theta);
double orc_mlp_loss_grad(const double *W, const double *X, const double *coef, int S, double *grad);

#endif
