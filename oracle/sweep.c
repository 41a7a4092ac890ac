/*
 * sweep.c — oracle restatement of the dense sweep (OpenPEEC src/solve.c:239-363): per point,
 * Lp and C = P^-1 (retarded decks refill them, static decks fill once), the MNA matrix, its
 * LU and one unit-current solve per port giving one column of Z.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include "oracle.h"

void orc_assemble(const peec_t *p, double f, const d_complex_t *lp, const d_complex_t *c, const int *cellid, int ncell,
                  d_complex_t *a);

int orc_sweep(const peec_t *p, int ifreq0, int nfreq, d_complex_t *zmat)
{
	const int n = p->nunknown, ns = p->nseg, np = p->nport;
	const int usecap = p->capacitance && (p->nchg > 0);
	int nc = 0;
	int *cellid = (int *)malloc((size_t)(p->nchg > 0 ? p->nchg : 1) * sizeof(int));
	for (int h = 0; h < p->nchg; h++) {
		int f = -1;
		for (int c = 0; c < nc; c++)
			if (cellid[c] == p->chgnode[h]) f = c;
		if (f < 0) cellid[nc++] = p->chgnode[h];
	}
	d_complex_t *lp = (d_complex_t *)malloc((size_t)(ns > 0 ? ns : 1) * (ns > 0 ? ns : 1) * sizeof(d_complex_t));
	d_complex_t *c = usecap ? (d_complex_t *)malloc((size_t)nc * nc * sizeof(d_complex_t)) : NULL;
	d_complex_t *a = (d_complex_t *)malloc((size_t)n * n * sizeof(d_complex_t));
	d_complex_t *b = (d_complex_t *)malloc((size_t)n * sizeof(d_complex_t));
	int *piv = (int *)malloc((size_t)n * sizeof(int));
	int rc = 0;
	if (!p->retardation) {
		if (ns > 0) orc_lp_fill(p, 0, lp);
		if (usecap && orc_pot_fill(p, 0, NULL, c)) rc = 1;
	}
	for (int q = 0; (q < nfreq) && !rc; q++) {
		const double f = freq_at(p, ifreq0 + q);
		if (p->retardation) {
			const double kw = 2 * PI * f / C0;
			if (ns > 0) orc_lp_fill(p, kw, lp);
			if (usecap && orc_pot_fill(p, kw, NULL, c)) {
				rc = 1;
				break;
			}
		}
		orc_assemble(p, f, lp, c, cellid, nc, a);
		if (orc_lu_decomp(n, a, piv) >= 0) {
			rc = 2;
			break;
		}
		for (int j = 0; j < np; j++) {
			memset(b, 0, (size_t)n * sizeof(d_complex_t));
			const int u = p->nodemap[p->port[j].n1], v = p->nodemap[p->port[j].n2];
			if (u >= 0) b[u] = d_add(b[u], d_complex(1, 0));
			if (v >= 0) b[v] = d_sub(b[v], d_complex(1, 0));
			orc_lu_solve(n, a, piv, b);
			for (int i = 0; i < np; i++) {
				const int ui = p->nodemap[p->port[i].n1], vi = p->nodemap[p->port[i].n2];
				const d_complex_t v1 = (ui < 0) ? d_complex(0, 0) : b[ui];
				const d_complex_t v2 = (vi < 0) ? d_complex(0, 0) : b[vi];
				zmat[((size_t)q * np + i) * np + j] = d_sub(v1, v2);
			}
		}
	}
	free(cellid);
	free(lp);
	free(c);
	free(a);
	free(b);
	free(piv);
	return rc;
}
