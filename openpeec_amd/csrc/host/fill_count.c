/*
 * fill_count.c — nominal FP64 operation count of the dense fills (Lp upper triangle with
 * images, charge-cell integral upper triangle with images) at one frequency.
 *
 * Walks exactly the branch tree of lp_entry / pot_fill (partial.c:256-424, potential.c:217-237):
 * orthogonal skips, far midpoints, closed forms, quadrature orders (nsub, nq, 7/3-point rules,
 * 1:4 splits) and image doubling, without evaluating any integral.  Each evaluated point is
 * charged a fixed nominal cost, counting +, -, *, /, sqrt and every libm call as one operation
 * at source level with loop-invariant geometry hoisted:
 *   thin-wire static Gauss point 19, retarded Gauss point 27 (partial.c:77-89)
 *   (e^{-jkR}-1)/R point pair 34 (surface.c:152-166, polygon.c:315-326)
 *   ribbon static outer point including rect_potential 150 (surface.c:76-91, 115-124)
 *   Hoer-Love primitive 100, 64 per bar pair (volume.c:44-71)
 *   polygon potential 20 + 74 npv (polygon.c:68-131), parallel closed form 40, midpoint 5.
 * This is the algorithmic work the fill GFLOP/s figure is quoted against.
 */
#include "peec.h"
#include "peec_mi355x.h"

#define F_WIRE_STATIC 19.0
#define F_WIRE_RET 27.0
#define F_CORR 34.0
#define F_RIBBON_STATIC 150.0
#define F_HL 100.0
#define F_POLY_BASE 20.0
#define F_POLY_EDGE 74.0
#define F_PARALLEL 40.0
#define F_MIDPOINT 5.0

static void ctr(const seg_t *s, double *m)
{
	for (int c = 0; c < 3; c++) m[c] = 0.5 * (s->x1[c] + s->x2[c]);
}

static double gap(const seg_t *a, const seg_t *b)
{
	double p[3], q[3];
	ctr(a, p);
	ctr(b, q);
	return sqrt(((q[0] - p[0]) * (q[0] - p[0])) + ((q[1] - p[1]) * (q[1] - p[1])) + ((q[2] - p[2]) * (q[2] - p[2])));
}

static double lmw(const seg_t *a, const seg_t *b)
{
	double m = (a->len > b->len) ? a->len : b->len;
	if (a->wid > m) m = a->wid;
	if (b->wid > m) m = b->wid;
	return m;
}

static void tang(const seg_t *s, double *t)
{
	for (int c = 0; c < 3; c++) t[c] = (s->x2[c] - s->x1[c]) / s->len;
}

static int parallel(const seg_t *a, const seg_t *b)
{
	double t1[3], t2[3];
	tang(a, t1);
	tang(b, t2);
	const double cx = (t1[1] * t2[2]) - (t1[2] * t2[1]);
	const double cy = (t1[2] * t2[0]) - (t1[0] * t2[2]);
	const double cz = (t1[0] * t2[1]) - (t1[1] * t2[0]);
	return sqrt((cx * cx) + (cy * cy) + (cz * cz)) < 1e-9;
}

/* quadrature points of a cell (cell_qpts: polygon fan, wire 4 x nsub, ribbon 16 x nsub) */
static double npoints(const seg_t *s, int nsub, int n7)
{
	if (s->npv > 0) return (double)s->npv * ((nsub >= 2) ? 4 : 1) * (n7 ? 7 : 3);
	if (s->wid <= 0) return 4.0 * nsub;
	return 16.0 * nsub;
}

static int aligned_bars(const seg_t *a, const seg_t *b)
{
	if (!(a->vol && b->vol)) return 0;
	double t1[3], t2[3];
	tang(a, t1);
	tang(b, t2);
	const double tt = fabs((t1[0] * t2[0]) + (t1[1] * t2[1]) + (t1[2] * t2[2]));
	const double ww = fabs((a->wv[0] * b->wv[0]) + (a->wv[1] * b->wv[1]) + (a->wv[2] * b->wv[2]));
	return (tt >= 1 - 1e-9) && (ww >= 1 - 1e-9) && (gap(a, b) <= 8.0 * lmw(a, b));
}

static double pair_cost(const seg_t *a, const seg_t *b, double kw, int same)
{
	if ((a->npv > 0) || (b->npv > 0)) {
		const double rc = gap(a, b), lm = lmw(a, b);
		double c;
		if ((rc > 20 * lm) && !same) {
			c = F_MIDPOINT;
		}
		else {
			const int nsub = (rc < 2 * lm) ? 2 : 1, n7 = (rc <= 5.0 * lm) ? 1 : 0;
			const int swap = (b->npv <= 0) && (b->wid <= 0);
			const seg_t *o = swap ? b : a, *in = swap ? a : b;
			const double per = (in->npv > 0) ? (F_POLY_BASE + F_POLY_EDGE * in->npv) : F_RIBBON_STATIC;
			c = npoints(o, nsub, n7) * per;
		}
		if (kw > 0) c += npoints(a, 1, 0) * npoints(b, 1, 0) * F_CORR;
		return c;
	}
	const double rc = gap(a, b), lm = lmw(a, b);
	const double ratio = (lm > 0) ? (rc / lm) : 0;
	int rsub = ((ratio > 2) ? 1 : 2) + (int)(kw * lm);
	if (rsub > 4) rsub = 4;
	if (aligned_bars(a, b)) return 64 * F_HL + ((kw > 0) ? 256.0 * rsub * rsub * F_CORR : 0);
	if ((a->wid > 0) || (b->wid > 0)) {
		const int nq = (ratio <= 2.5) ? 8 : (ratio <= 10) ? 4 : 2;
		return (double)nq * nq * rsub * rsub * F_RIBBON_STATIC + ((kw > 0) ? 256.0 * rsub * rsub * F_CORR : 0);
	}
	const double lmax = (a->len > b->len) ? a->len : b->len;
	if (kw <= 0) {
		if (parallel(a, b)) return F_PARALLEL;
		if (rc > 20 * lmax) return F_MIDPOINT;
		const int ns = (rc > 2 * lmax) ? 1 : 4;
		return 64.0 * ns * ns * F_WIRE_STATIC;
	}
	int nsub = 2 + (int)(kw * lmax);
	if (nsub > 8) nsub = 8;
	return (parallel(a, b) ? F_PARALLEL : 64.0 * nsub * nsub * F_WIRE_STATIC) + 64.0 * nsub * nsub * F_WIRE_RET;
}

static double self_cost(const seg_t *s, double kw)
{
	if (s->wid > 0) return pair_cost(s, s, kw, 1);
	if (kw <= 0) return F_PARALLEL;
	int nsub = 2 + (int)(kw * s->len);
	if (nsub > 8) nsub = 8;
	return F_PARALLEL + 64.0 * nsub * nsub * F_WIRE_RET;
}

double peec_fill_flops(const peec_t *p, double f)
{
	const double kw = p->retardation ? (2 * PI * f / C0) : 0;
	double total = 0;
	int i;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
	for (i = 0; i < p->nseg; i++) {
		const seg_t *a = &p->seg[i];
		double t1[3];
		tang(a, t1);
		seg_t im;
		for (int j = i; j < p->nseg; j++) {
			const seg_t *b = &p->seg[j];
			double t2[3];
			tang(b, t2);
			const double tdot = (t1[0] * t2[0]) + (t1[1] * t2[1]) + (t1[2] * t2[2]);
			const double tdm = (t1[0] * t2[0]) + (t1[1] * t2[1]) - (t1[2] * t2[2]);
			if (i == j) total += self_cost(a, kw);
			else if (tdot != 0) total += pair_cost(a, b, kw, 0);
			if (p->gp && ((i == j) || (tdm != 0))) {
				seg_mirror(b, p->gpz, &im);
				total += pair_cost(a, &im, kw, 0);
			}
		}
	}
	if (p->capacitance && !p->compress) {
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
		for (i = 0; i < p->nchg; i++) {
			const seg_t *a = &p->chg[i];
			seg_t im;
			for (int j = i; j < p->nchg; j++) {
				const seg_t *b = &p->chg[j];
				total += (i == j) ? self_cost(a, kw) : pair_cost(a, b, kw, 0);
				if (p->gp) {
					seg_mirror(b, p->gpz, &im);
					total += pair_cost(a, &im, kw, 0);
				}
			}
		}
	}
	return total;
}
