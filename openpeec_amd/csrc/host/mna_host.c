/*
 * mna_host.c — MNA numbering and the sparse stamps of the modified nodal analysis.
 *
 * Unknown order [node voltages | inductor currents | vsource currents | cell branch currents]
 * and the reference-node rule follow mna_numbering (OpenPEEC src/mna.c:16-113).  The device
 * assembly writes the two dense kernel blocks (-jw Lp, jw C) itself; everything else (lumped
 * elements, incidence, internal impedance) is produced here as stamps in the reference's
 * stamping order (mna.c:175-297) and coalesced per matrix entry with the same left-to-right
 * summation, so that sparse-sum + dense-block reproduces the reference's matrix bit for bit.
 */
#include "peec.h"
#include "peec_internal.h"

int mna_numbering(peec_t *p, FILE *fp_log)
{
	const int nid = p->maxid + 1;
	int *used = (int *)calloc((size_t)nid, sizeof(int));
	if (used == NULL) return 1;
	for (int i = 0; i < p->nres; i++) used[p->res[i].n1] = used[p->res[i].n2] = 1;
	for (int i = 0; i < p->ncap; i++) used[p->cap[i].n1] = used[p->cap[i].n2] = 1;
	for (int i = 0; i < p->nind; i++) used[p->ind[i].n1] = used[p->ind[i].n2] = 1;
	for (int i = 0; i < p->nsrc; i++) used[p->src[i].n1] = used[p->src[i].n2] = 1;
	for (int i = 0; i < p->nport; i++) used[p->port[i].n1] = used[p->port[i].n2] = 1;
	for (int i = 0; i < p->nseg; i++) used[p->seg[i].n1] = used[p->seg[i].n2] = 1;

	/* With capacitive cells the potential reference is infinity: only node 0 may be eliminated
	 * (mna.c:47-55); otherwise node 0, or port #1's n2 when node 0 is unused. */
	const int usecap = (p->capacitance && (p->ncell > 0));
	p->refnode = (usecap || used[0]) ? 0 : p->port[0].n2;

	p->nodemap = (int *)malloc((size_t)nid * sizeof(int));
	if (p->nodemap == NULL) {
		free(used);
		return 1;
	}
	int idx = 0;
	for (int id = 0; id < nid; id++) {
		if (id == p->refnode) p->nodemap[id] = -1;
		else if (!used[id]) p->nodemap[id] = -2;
		else p->nodemap[id] = idx++;
	}
	p->nnode = idx;
	free(used);

	int nvsrc = 0;
	for (int i = 0; i < p->nsrc; i++)
		if (p->src[i].isvsrc) p->src[i].ibr = nvsrc++;
	p->offL = p->nnode;
	p->offV = p->offL + p->nind;
	p->offS = p->offV + nvsrc;
	p->nunknown = p->offS + p->nseg;

	/* compressed capacitive form: charges become unknowns after the branch currents */
	p->offQ = p->nunknown;
	if (p->compress && p->capacitance && (p->nchg > 0)) {
		if (pot_cells(p)) {
			printf("%s\n", "*** memory allocation error (potential cells)");
			return 1;
		}
		p->nunknown = p->offQ + p->ncell;
	}
	if (fp_log) fprintf(fp_log, "MNA : %d nodes (reference = node %d), %d unknowns\n", p->nnode + 1, p->refnode, p->nunknown);
	return 0;
}

/* series impedance of a dielectric branch (mna.c:128-138): Debye or constant loss tangent */
static d_complex_t zint_dielectric(const seg_t *s, double omega, double f)
{
	if (s->frelax > 0) {
		const d_complex_t er = d_add(d_complex(s->epsinf, 0), d_div(d_complex(s->epss - s->epsinf, 0), d_complex(1, f / s->frelax)));
		return d_inv(d_complex(omega * s->gfac * (-er.i), omega * s->gfac * (er.r - 1)));
	}
	return d_rmul(1 / omega, d_inv(d_complex(s->gexc, s->cexc)));
}

/* ---- stamp accumulation: open-addressing map (row, col) -> slot, fold in emission order ---- */

typedef struct {
	int n;             /* matrix order (key = r * n + c) */
	int cap;           /* power of two */
	long long *key;
	int *slot;
	stamps_t *out;
	int outcap;
} acc_t;

static void acc_init(acc_t *a, int n, int expect, stamps_t *out)
{
	a->n = n;
	a->cap = 1024;
	while (a->cap < 2 * expect) a->cap <<= 1;
	a->key = (long long *)malloc((size_t)a->cap * sizeof(long long));
	a->slot = (int *)malloc((size_t)a->cap * sizeof(int));
	for (int i = 0; i < a->cap; i++) a->key[i] = -1;
	a->out = out;
	a->outcap = (expect > 16) ? expect : 16;
	out->n = 0;
	out->ri = (int *)malloc((size_t)a->outcap * sizeof(int));
	out->ci = (int *)malloc((size_t)a->outcap * sizeof(int));
	out->val = (d_complex_t *)malloc((size_t)a->outcap * sizeof(d_complex_t));
}

static void acc_grow(acc_t *a)
{
	const int oc = a->cap;
	long long *ok = a->key;
	int *os = a->slot;
	a->cap = 2 * oc;
	a->key = (long long *)malloc((size_t)a->cap * sizeof(long long));
	a->slot = (int *)malloc((size_t)a->cap * sizeof(int));
	for (int i = 0; i < a->cap; i++) a->key[i] = -1;
	for (int i = 0; i < oc; i++) {
		if (ok[i] < 0) continue;
		unsigned long long h = ((unsigned long long)ok[i] * 0x9E3779B97F4A7C15ull) >> 20;
		int b = (int)(h & (unsigned long long)(a->cap - 1));
		while (a->key[b] >= 0) b = (b + 1) & (a->cap - 1);
		a->key[b] = ok[i];
		a->slot[b] = os[i];
	}
	free(ok);
	free(os);
}

/* a[i][j] += v, skipping the reference node (i or j < 0), as stamp() in mna.c:154-172 */
static void stamp(acc_t *a, int i, int j, d_complex_t v)
{
	if ((i < 0) || (j < 0)) return;
	if (2 * (a->out->n + 1) > a->cap) acc_grow(a);
	const long long k = (long long)i * a->n + j;
	unsigned long long h = ((unsigned long long)k * 0x9E3779B97F4A7C15ull) >> 20;
	int b = (int)(h & (unsigned long long)(a->cap - 1));
	while ((a->key[b] >= 0) && (a->key[b] != k)) b = (b + 1) & (a->cap - 1);
	if (a->key[b] == k) {
		const int s = a->slot[b];
		a->out->val[s] = d_add(a->out->val[s], v);
		return;
	}
	stamps_t *o = a->out;
	if (o->n >= a->outcap) {
		a->outcap *= 2;
		o->ri = (int *)realloc(o->ri, (size_t)a->outcap * sizeof(int));
		o->ci = (int *)realloc(o->ci, (size_t)a->outcap * sizeof(int));
		o->val = (d_complex_t *)realloc(o->val, (size_t)a->outcap * sizeof(d_complex_t));
	}
	a->key[b] = k;
	a->slot[b] = o->n;
	o->ri[o->n] = i;
	o->ci[o->n] = j;
	o->val[o->n] = d_add(d_complex(0, 0), v);
	o->n++;
}

/* stamps of assemble(skiplp = 1) (mna.c:175-297); charge_off >= 0 adds the charge-unknown
 * couplings of the compressed formulation with the charges numbered from charge_off */
static void stamps_any(const peec_t *p, double f, stamps_t *st, int charge_off)
{
	const double omega = 2 * PI * f;
	const int *map = p->nodemap;
	acc_t a;
	acc_init(&a, p->nunknown + p->ncell, p->nnode + 2 * p->nres + 2 * p->ncap + 5 * p->nind + 5 * p->nseg + 4 * p->nsrc + 3 * p->ncell, st);

	if (p->gmin > 0)
		for (int i = 0; i < p->nnode; i++) stamp(&a, i, i, d_complex(p->gmin, 0));
	for (int i = 0; i < p->nres; i++) {
		const int i1 = map[p->res[i].n1], i2 = map[p->res[i].n2];
		const d_complex_t y = d_complex(1 / p->res[i].val, 0);
		stamp(&a, i1, i1, y);
		stamp(&a, i2, i2, y);
		stamp(&a, i1, i2, d_rmul(-1, y));
		stamp(&a, i2, i1, d_rmul(-1, y));
	}
	for (int i = 0; i < p->ncap; i++) {
		const int i1 = map[p->cap[i].n1], i2 = map[p->cap[i].n2];
		const d_complex_t y = d_complex(0, omega * p->cap[i].val);
		stamp(&a, i1, i1, y);
		stamp(&a, i2, i2, y);
		stamp(&a, i1, i2, d_rmul(-1, y));
		stamp(&a, i2, i1, d_rmul(-1, y));
	}
	for (int i = 0; i < p->nind; i++) {
		const int i1 = map[p->ind[i].n1], i2 = map[p->ind[i].n2];
		const int r = p->offL + i;
		stamp(&a, i1, r, d_complex(1, 0));
		stamp(&a, r, i1, d_complex(1, 0));
		stamp(&a, i2, r, d_complex(-1, 0));
		stamp(&a, r, i2, d_complex(-1, 0));
		stamp(&a, r, r, d_complex(0, -omega * p->ind[i].val));
	}
	for (int i = 0; i < p->nmut; i++) {
		const int r1 = p->offL + p->mut[i].l1, r2 = p->offL + p->mut[i].l2;
		const double m = p->mut[i].k * sqrt(p->ind[p->mut[i].l1].val * p->ind[p->mut[i].l2].val);
		stamp(&a, r1, r2, d_complex(0, -omega * m));
		stamp(&a, r2, r1, d_complex(0, -omega * m));
	}
	for (int i = 0; i < p->nsrc; i++) {
		if (!p->src[i].isvsrc) continue;
		const int i1 = map[p->src[i].n1], i2 = map[p->src[i].n2];
		const int r = p->offV + p->src[i].ibr;
		stamp(&a, i1, r, d_complex(1, 0));
		stamp(&a, r, i1, d_complex(1, 0));
		stamp(&a, i2, r, d_complex(-1, 0));
		stamp(&a, r, i2, d_complex(-1, 0));
	}
	for (int k = 0; k < p->nseg; k++) {
		const int i1 = map[p->seg[k].n1], i2 = map[p->seg[k].n2];
		const int rk = p->offS + k;
		stamp(&a, i1, rk, d_complex(1, 0));
		stamp(&a, rk, i1, d_complex(1, 0));
		stamp(&a, i2, rk, d_complex(-1, 0));
		stamp(&a, rk, i2, d_complex(-1, 0));
		/* internal impedance: the DC resistance (skin effect is refused by the deck reader) */
		const d_complex_t zint = p->seg[k].diel ? zint_dielectric(&p->seg[k], omega, f) : d_complex(p->seg[k].res, 0);
		stamp(&a, rk, rk, d_rmul(-1, zint));
	}
	/* compressed capacitive form: KCL + jw q and charge-row -v couplings (mna.c:271-283) */
	if (p->capacitance && (p->ncell > 0) && (charge_off >= 0)) {
		for (int i = 0; i < p->ncell; i++) {
			const int mi = map[p->cellid[i]];
			const int ri = charge_off + i;
			stamp(&a, mi, ri, d_complex(0, omega));
			stamp(&a, ri, mi, d_complex(-1, 0));
		}
	}
	free(a.key);
	free(a.slot);
}

void mna_stamps(const peec_t *p, double f, stamps_t *st)
{
	stamps_any(p, f, st, (p->offQ < p->nunknown) ? p->offQ : -1);
}

/* the dense numbering extended by one charge unknown per capacitive cell after the last
 * branch current (charges at p->nunknown ...): the system of the accelerated sweep */
void mna_stamps_charge(const peec_t *p, double f, stamps_t *st)
{
	stamps_any(p, f, st, p->nunknown);
}

void mna_stamps_free(stamps_t *st)
{
	free(st->ri);
	free(st->ci);
	free(st->val);
	memset(st, 0, sizeof(*st));
}

void mna_rhs_port(const peec_t *p, int iport, d_complex_t *b)
{
	const int n = p->nunknown;
	const int i1 = p->nodemap[p->port[iport].n1];
	const int i2 = p->nodemap[p->port[iport].n2];
	memset(b, 0, (size_t)n * sizeof(d_complex_t));
	if (i1 >= 0) b[i1] = d_add(b[i1], d_complex(1, 0));
	if (i2 >= 0) b[i2] = d_sub(b[i2], d_complex(1, 0));
}

/* Independent sources as a right-hand side (mna.c:471-489): a vsource sets its branch row to
 * amp / phase; an isource drives its current from n1 to n2 through the source (SPICE sign). */
void mna_rhs_sources(const peec_t *p, d_complex_t *b)
{
	memset(b, 0, (size_t)p->nunknown * sizeof(d_complex_t));
	for (int k = 0; k < p->nsrc; k++) {
		const src_t *e = &p->src[k];
		const d_complex_t v = d_polar_deg(e->amp, e->phase);
		if (e->isvsrc) {
			b[p->offV + e->ibr] = v;
			continue;
		}
		const int r1 = p->nodemap[e->n1], r2 = p->nodemap[e->n2];
		if (r1 >= 0) b[r1] = d_sub(b[r1], v);
		if (r2 >= 0) b[r2] = d_add(b[r2], v);
	}
}

/* sin(u) / u, 1 for |u| < 1e-8 */
static double pw_sinc(double u)
{
	return (fabs(u) < 1e-8) ? 1.0 : (sin(u) / u);
}

static double dot3(const double *a, const double *b)
{
	return (a[0] * b[0]) + (a[1] * b[1]) + (a[2] * b[2]);
}

/* Line integral of a uniform plane wave E0 e exp(+j k dir . r) along one current cell, exact for
 * a straight cell of constant direction: (E0 e . t) len exp(j k dir . c) sinc(k len (dir . t) / 2),
 * with the cross-width average sinc(k wid (dir . w) / 2) for ribbons and bars (mna.c:402-431). */
static d_complex_t pw_line_integral(const seg_t *s, const double *dir, const double *e, d_complex_t e0, double kw)
{
	double t[3], c[3];
	for (int q = 0; q < 3; q++) {
		t[q] = (s->x2[q] - s->x1[q]) / s->len;
		c[q] = 0.5 * (s->x1[q] + s->x2[q]);
	}
	double amp = dot3(e, t) * s->len * pw_sinc(0.5 * kw * s->len * dot3(dir, t));
	const double wn = sqrt(dot3(s->wv, s->wv));
	if ((s->wid > 0) && (wn > 0)) amp *= pw_sinc(0.5 * kw * s->wid * (dot3(dir, s->wv) / wn));
	const double ph = kw * dot3(dir, c);
	return d_rmul(amp, d_mul(e0, d_complex(cos(ph), sin(ph))));
}

/* Plane-wave excitation as a right-hand side (mna.c:433-469): each branch row receives minus
 * the incident field's line integral along its cell.  The wave arrives from (pwth, pwph)
 * degrees with theta (pwpol 1) or phi (pwpol 2) polarisation; over a ground plane the reflected
 * wave (arrival mirrored in z, horizontal field reversed, phase matched on the plane) is added. */
void mna_rhs_planewave(const peec_t *p, double f, d_complex_t *b)
{
	memset(b, 0, (size_t)p->nunknown * sizeof(d_complex_t));
	if (!p->pw) return;
	const double kw = 2 * PI * f / C0;
	const double th = p->pwth * PI / 180, ph = p->pwph * PI / 180;
	const double st = sin(th), ct = cos(th), sp = sin(ph), cp = cos(ph);
	const double dir[3] = {st * cp, st * sp, ct};
	double e[3];
	if (p->pwpol == 2) {
		e[0] = -sp;
		e[1] = cp;
		e[2] = 0;
	}
	else {
		e[0] = ct * cp;
		e[1] = ct * sp;
		e[2] = -st;
	}
	const d_complex_t e0 = d_polar_deg(p->pwamp, p->pwphase);
	const double dir_r[3] = {dir[0], dir[1], -dir[2]};
	const double e_r[3] = {-e[0], -e[1], e[2]};
	const double psi = 2 * kw * dir[2] * p->gpz;
	const d_complex_t e0_r = d_mul(e0, d_complex(cos(psi), sin(psi)));
	for (int m = 0; m < p->nseg; m++) {
		d_complex_t v = pw_line_integral(&p->seg[m], dir, e, e0, kw);
		if (p->gp) v = d_add(v, pw_line_integral(&p->seg[m], dir_r, e_r, e0_r, kw));
		b[p->offS + m] = d_rmul(-1, v);
	}
}
