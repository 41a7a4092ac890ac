/*
 * assemble.c — oracle restatement of the dense MNA matrix (OpenPEEC src/mna.c:175-308).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Uses the peec_t's numbering (nodemap, offsets).
 * Skin-effect decks (skineffect = 1) are outside the oracle's scope.
 */
#include "oracle.h"

static void add(d_complex_t *a, int n, int i, int j, d_complex_t v)
{
	if ((i < 0) || (j < 0)) return;
	a[(size_t)i * n + j] = d_add(a[(size_t)i * n + j], v);
}

/* excess-permittivity branch impedance (mna.c:128-138) */
static d_complex_t z_diel(const seg_t *s, double w, double f)
{
	if (s->frelax > 0) {
		const d_complex_t er = d_add(d_complex(s->epsinf, 0), d_div(d_complex(s->epss - s->epsinf, 0), d_complex(1, f / s->frelax)));
		return d_inv(d_complex(w * s->gfac * (-er.i), w * s->gfac * (er.r - 1)));
	}
	return d_rmul(1 / w, d_inv(d_complex(s->gexc, s->cexc)));
}

void orc_assemble(const peec_t *p, double f, const d_complex_t *lp, const d_complex_t *c, const int *cellid, int ncell,
                  d_complex_t *a)
{
	const int n = p->nunknown;
	const double w = 2 * PI * f;
	const int *map = p->nodemap;
	memset(a, 0, (size_t)n * n * sizeof(d_complex_t));
	if (p->gmin > 0)
		for (int i = 0; i < p->nnode; i++) add(a, n, i, i, d_complex(p->gmin, 0));
	for (int i = 0; i < p->nres; i++) {
		const int u = map[p->res[i].n1], v = map[p->res[i].n2];
		const d_complex_t y = d_complex(1 / p->res[i].val, 0);
		add(a, n, u, u, y);
		add(a, n, v, v, y);
		add(a, n, u, v, d_rmul(-1, y));
		add(a, n, v, u, d_rmul(-1, y));
	}
	for (int i = 0; i < p->ncap; i++) {
		const int u = map[p->cap[i].n1], v = map[p->cap[i].n2];
		const d_complex_t y = d_complex(0, w * p->cap[i].val);
		add(a, n, u, u, y);
		add(a, n, v, v, y);
		add(a, n, u, v, d_rmul(-1, y));
		add(a, n, v, u, d_rmul(-1, y));
	}
	for (int i = 0; i < p->nind; i++) {
		const int u = map[p->ind[i].n1], v = map[p->ind[i].n2], r = p->offL + i;
		add(a, n, u, r, d_complex(1, 0));
		add(a, n, r, u, d_complex(1, 0));
		add(a, n, v, r, d_complex(-1, 0));
		add(a, n, r, v, d_complex(-1, 0));
		add(a, n, r, r, d_complex(0, -w * p->ind[i].val));
	}
	for (int i = 0; i < p->nmut; i++) {
		const int r1 = p->offL + p->mut[i].l1, r2 = p->offL + p->mut[i].l2;
		const double m = p->mut[i].k * sqrt(p->ind[p->mut[i].l1].val * p->ind[p->mut[i].l2].val);
		add(a, n, r1, r2, d_complex(0, -w * m));
		add(a, n, r2, r1, d_complex(0, -w * m));
	}
	for (int i = 0; i < p->nsrc; i++) {
		if (!p->src[i].isvsrc) continue;
		const int u = map[p->src[i].n1], v = map[p->src[i].n2], r = p->offV + p->src[i].ibr;
		add(a, n, u, r, d_complex(1, 0));
		add(a, n, r, u, d_complex(1, 0));
		add(a, n, v, r, d_complex(-1, 0));
		add(a, n, r, v, d_complex(-1, 0));
	}
	for (int k = 0; k < p->nseg; k++) {
		const int u = map[p->seg[k].n1], v = map[p->seg[k].n2], r = p->offS + k;
		add(a, n, u, r, d_complex(1, 0));
		add(a, n, r, u, d_complex(1, 0));
		add(a, n, v, r, d_complex(-1, 0));
		add(a, n, r, v, d_complex(-1, 0));
		const d_complex_t zi = p->seg[k].diel ? z_diel(&p->seg[k], w, f) : d_complex(p->seg[k].res, 0);
		add(a, n, r, r, d_rmul(-1, zi));
		for (int m = 0; m < p->nseg; m++) add(a, n, r, p->offS + m, d_mul(d_complex(0, -w), lp[(size_t)k * p->nseg + m]));
	}
	if (c != NULL)
		for (int i = 0; i < ncell; i++) {
			const int mi = map[cellid[i]];
			for (int j = 0; j < ncell; j++) add(a, n, mi, map[cellid[j]], d_mul(d_complex(0, w), c[(size_t)i * ncell + j]));
		}
}
