/*
 * fill_pack.c — packs seg_t cell arrays into the device cell format (dcell_t + polygon pools).
 *
 * Frame vectors, centres and polygon edge data are evaluated once here with exactly the
 * expressions the reference evaluates inline at every call (partial.c:107-108 axis,
 * surface.c:63-73 local frame, volume.c:79-89, polygon.c:46-131 Newell normal and edge terms,
 * polygon.c:154-163 fan-triangle area), so the device kernels read bit-identical inputs.
 * mirror = 1 packs the ground-plane images of the cells (seg_mirror, partial.c:352-362): z of
 * x1, x2, papex and every vertex becomes 2 gpz - z, wv.z flips, len is kept.
 */
#include "peec.h"
#include "peec_internal.h"

static void cross3(const double *a, const double *b, double *c)
{
	c[0] = (a[1] * b[2]) - (a[2] * b[1]);
	c[1] = (a[2] * b[0]) - (a[0] * b[2]);
	c[2] = (a[0] * b[1]) - (a[1] * b[0]);
}

static double dot3(const double *a, const double *b)
{
	return (a[0] * b[0]) + (a[1] * b[1]) + (a[2] * b[2]);
}

void seg_mirror(const seg_t *s, double gpz, seg_t *out)
{
	*out = *s;
	out->x1[2] = (2 * gpz) - s->x1[2];
	out->x2[2] = (2 * gpz) - s->x2[2];
	out->wv[2] = -s->wv[2];
	out->papex[2] = (2 * gpz) - s->papex[2];
	for (int k = 0; k < s->npv; k++) out->pv[(3 * k) + 2] = (2 * gpz) - s->pv[(3 * k) + 2];
}

static void pack_one(const seg_t *s, dcell_t *c, dedge_t *edges, dpnorm_t *pn, int poff)
{
	memset(c, 0, sizeof(*c));
	for (int k = 0; k < 3; k++) {
		c->x1[k] = s->x1[k];
		c->x2[k] = s->x2[k];
		c->d[k] = s->x2[k] - s->x1[k];
		c->t[k] = (s->x2[k] - s->x1[k]) / s->len;
		c->c[k] = 0.5 * (s->x1[k] + s->x2[k]);
		c->wv[k] = s->wv[k];
		c->papex[k] = s->papex[k];
	}
	cross3(c->t, c->wv, c->nv);
	c->len = s->len;
	c->wid = s->wid;
	c->thick = s->thick;
	c->aL = s->aL;
	c->aP = s->aP;
	c->npv = s->npv;
	c->poff = poff;
	c->vol = s->vol;
	c->kind = (s->npv > 0) ? CELL_POLY : s->vol ? CELL_BAR : (s->wid > 0) ? CELL_RIBBON : CELL_WIRE;
	if (s->npv <= 0) return;

	/* Newell normal (its norm is twice the area) and the unit normal */
	double nv[3] = {0, 0, 0};
	for (int k = 0; k < s->npv; k++) {
		double cr[3];
		cross3(&s->pv[3 * k], &s->pv[3 * ((k + 1) % s->npv)], cr);
		for (int i = 0; i < 3; i++) nv[i] += cr[i];
	}
	const double n2 = sqrt(dot3(nv, nv));
	pn->n2 = n2;
	for (int i = 0; i < 3; i++) pn->nh[i] = (n2 > 0) ? (nv[i] / n2) : 0;
	for (int k = 0; k < s->npv; k++) {
		const double *a = &s->pv[3 * k];
		const double *b = &s->pv[3 * ((k + 1) % s->npv)];
		dedge_t *e = &edges[k];
		double ev[3], u[3], v[3], w[3];
		for (int i = 0; i < 3; i++) {
			e->a[i] = a[i];
			e->b[i] = b[i];
			ev[i] = b[i] - a[i];
		}
		e->el = sqrt(dot3(ev, ev));
		for (int i = 0; i < 3; i++) e->sh[i] = (e->el > 0) ? (ev[i] / e->el) : 0;
		cross3(e->sh, pn->nh, e->mh);
		/* fan triangle (papex, a, b) */
		for (int i = 0; i < 3; i++) {
			u[i] = a[i] - s->papex[i];
			v[i] = b[i] - s->papex[i];
		}
		cross3(u, v, w);
		e->at = 0.5 * sqrt(dot3(w, w));
	}
}

int cells_pack(const seg_t *src, int n, int mirror, double gpz,
	dcell_t **cells, dedge_t **edges, dpnorm_t **pnorm, int *nedge)
{
	int ne = 0;
	for (int i = 0; i < n; i++) ne += (src[i].npv > 0) ? src[i].npv : 0;
	*cells = (dcell_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(dcell_t));
	*edges = (dedge_t *)calloc((size_t)(ne > 0 ? ne : 1), sizeof(dedge_t));
	*pnorm = (dpnorm_t *)calloc((size_t)(n > 0 ? n : 1), sizeof(dpnorm_t));
	if ((*cells == NULL) || (*edges == NULL) || (*pnorm == NULL)) return 1;
	int off = 0;
	for (int i = 0; i < n; i++) {
		if (mirror) {
			seg_t m;
			seg_mirror(&src[i], gpz, &m);
			pack_one(&m, &(*cells)[i], &(*edges)[off], &(*pnorm)[i], off);
		}
		else {
			pack_one(&src[i], &(*cells)[i], &(*edges)[off], &(*pnorm)[i], off);
		}
		off += (src[i].npv > 0) ? src[i].npv : 0;
	}
	*nedge = ne;
	return 0;
}
