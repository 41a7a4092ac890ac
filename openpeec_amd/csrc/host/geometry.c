/*
 * geometry.c — discretisation of the deck's conductors into current cells (p->seg) and charge
 * cells (p->chg): the input of the fill.  The cell records must come out bit-identical to the
 * reference's wire_build (OpenPEEC src/wire.c:348-913) because every partial element is a
 * function of them; the arithmetic below therefore keeps the reference's expression order
 * (e.g. x1 + (x2 - x1) * k / ndiv) while the code structure is this project's own.
 *
 * Node merging follows the reference rule (wire.c:36-63): a point within nodetol of an
 * existing geometric node reuses that node (first registered wins), otherwise it gets the next
 * automatic id after the deck's largest id.  The lookup uses a uniform hash grid instead of the
 * reference's O(N^2) scan; it returns the same (lowest-index) match.
 */
#include "peec.h"
#include "peec_internal.h"

/* ------------------------------------------------------------------------------------------ */
/* small vector helpers                                                                        */

static double vdist(const double *a, const double *b)
{
	const double dx = a[0] - b[0];
	const double dy = a[1] - b[1];
	const double dz = a[2] - b[2];
	return sqrt((dx * dx) + (dy * dy) + (dz * dz));
}

static double vnorm(const double *a)
{
	return sqrt((a[0] * a[0]) + (a[1] * a[1]) + (a[2] * a[2]));
}

static void vmid(const double *a, const double *b, double *m)
{
	for (int k = 0; k < 3; k++) m[k] = 0.5 * (a[k] + b[k]);
}

static void vcross(const double *a, const double *b, double *c)
{
	c[0] = (a[1] * b[2]) - (a[2] * b[1]);
	c[1] = (a[2] * b[0]) - (a[0] * b[2]);
	c[2] = (a[0] * b[1]) - (a[1] * b[0]);
}

/* ------------------------------------------------------------------------------------------ */
/* geometric node table with a hash grid (cell size = nodetol)                                 */

typedef struct {
	peec_t *p;
	int     autoid;
	int     nbucket;
	int    *head;      /* [nbucket] first node index in bucket, -1 = empty */
	int    *next;      /* [capacity] chain                                 */
} nodes_t;

static long long cell_coord(double x, double h)
{
	return (long long)floor(x / h);
}

static unsigned bucket_of(const nodes_t *nt, long long ix, long long iy, long long iz)
{
	unsigned long long h = (unsigned long long)ix * 0x9E3779B97F4A7C15ull;
	h ^= (unsigned long long)iy * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2);
	h ^= (unsigned long long)iz * 0x165667B19E3779F9ull + (h << 6) + (h >> 2);
	return (unsigned)(h % (unsigned long long)nt->nbucket);
}

/* lowest-index node within nodetol of x, or -1 (same result as a linear scan) */
static int nodes_find(const nodes_t *nt, const double *x)
{
	const peec_t *p = nt->p;
	const double h = p->nodetol;
	const long long cx = cell_coord(x[0], h), cy = cell_coord(x[1], h), cz = cell_coord(x[2], h);
	int best = -1;
	for (long long dx = -1; dx <= 1; dx++)
	for (long long dy = -1; dy <= 1; dy++)
	for (long long dz = -1; dz <= 1; dz++) {
		for (int i = nt->head[bucket_of(nt, cx + dx, cy + dy, cz + dz)]; i >= 0; i = nt->next[i]) {
			if ((best >= 0) && (i >= best)) continue;
			if (vdist(&p->gxyz[3 * i], x) <= h) best = i;
		}
	}
	return best;
}

static void nodes_add(nodes_t *nt, int id, const double *x)
{
	peec_t *p = nt->p;
	const int i = p->ngnode;
	p->gid[i] = id;
	for (int k = 0; k < 3; k++) p->gxyz[(3 * i) + k] = x[k];
	const double h = p->nodetol;
	const unsigned b = bucket_of(nt, cell_coord(x[0], h), cell_coord(x[1], h), cell_coord(x[2], h));
	nt->next[i] = nt->head[b];
	nt->head[b] = i;
	p->ngnode++;
}

static int nodes_at(nodes_t *nt, const double *x)
{
	const int found = nodes_find(nt, x);
	if (found >= 0) return nt->p->gid[found];
	const int id = ++nt->autoid;
	nodes_add(nt, id, x);
	return id;
}

/* ------------------------------------------------------------------------------------------ */
/* cell constructors                                                                            */

static void push_charge(peec_t *p, int node, const double *x1, const double *x2,
	double wid, const double *wv, const seg_t *proto)
{
	seg_t *c = &p->chg[p->nchg];
	*c = *proto;
	for (int k = 0; k < 3; k++) {
		c->x1[k] = x1[k];
		c->x2[k] = x2[k];
		c->wv[k] = (wv != NULL) ? wv[k] : 0;
	}
	c->len = vdist(x1, x2);
	c->wid = wid;
	p->chgnode[p->nchg] = node;
	p->nchg++;
}

static void build_wires(peec_t *p, nodes_t *nt)
{
	for (int iw = 0; iw < p->nwire; iw++) {
		const wire_t *w = &p->wire[iw];
		const int ndiv = w->ndiv;
		int prev = -1;
		for (int k = 0; k <= ndiv; k++) {
			double x[3];
			for (int c = 0; c < 3; c++) x[c] = w->x1[c] + (w->x2[c] - w->x1[c]) * k / ndiv;
			const int id = nodes_at(nt, x);
			if (k > 0) {
				seg_t *s = &p->seg[p->nseg];
				memset(s, 0, sizeof(seg_t));
				s->n1 = prev;
				s->n2 = id;
				for (int c = 0; c < 3; c++) {
					s->x1[c] = w->x1[c] + (w->x2[c] - w->x1[c]) * (k - 1) / ndiv;
					s->x2[c] = x[c];
				}
				s->len = vdist(s->x1, s->x2);
				s->shape = w->shape;
				s->radius = w->radius;
				s->width = w->width;
				s->thick = w->thick;
				s->sigma = w->sigma;
				s->wid = 0;
				if (w->shape == SHAPE_BAR) {
					/* GMD radii of a rectangular section (Grover): aL = 0.2235 (w + t), aP = (w + t) / 4 */
					const double wt = w->width + w->thick;
					s->aL = 0.2235 * wt;
					s->aP = 0.25 * wt;
					s->area = w->width * w->thick;
					s->perim = 2 * wt;
				}
				else {
					s->aL = w->radius;
					s->aP = w->radius;
					s->area = PI * w->radius * w->radius;
					s->perim = 2 * PI * w->radius;
				}
				s->res = (w->sigma > 0) ? s->len / (w->sigma * s->area) : 0;
				/* two half-segment charge cells, one per end node */
				double mid[3];
				vmid(s->x1, s->x2, mid);
				push_charge(p, s->n1, s->x1, mid, 0, NULL, s);
				push_charge(p, s->n2, mid, s->x2, 0, NULL, s);
				p->nseg++;
			}
			prev = id;
		}
	}
}

static int build_plates(peec_t *p, nodes_t *nt)
{
	for (int ip = 0; ip < p->nplate; ip++) {
		const plate_t *pl = &p->plate[ip];
		const int na = pl->ndiva, nb = pl->ndivb;
		const double la = vnorm(pl->ea), lb = vnorm(pl->eb);
		const double ha = la / na, hb = lb / nb;
		double ta[3], tb[3], nv[3];
		for (int c = 0; c < 3; c++) {
			ta[c] = pl->ea[c] / la;
			tb[c] = pl->eb[c] / lb;
		}
		double *xa = (double *)malloc((size_t)(na + 1) * sizeof(double));
		double *xb = (double *)malloc((size_t)(nb + 1) * sizeof(double));
		int *nid = (int *)malloc((size_t)(na + 1) * (nb + 1) * sizeof(int));
		if ((xa == NULL) || (xb == NULL) || (nid == NULL)) {
			printf("%s\n", "*** memory allocation error (plate)");
			free(xa);
			free(xb);
			free(nid);
			return 1;
		}
		for (int ia = 0; ia <= na; ia++) xa[ia] = p->grading ? (0.5 * la * (1 - cos(PI * ia / na))) : (ia * ha);
		for (int ib = 0; ib <= nb; ib++) xb[ib] = p->grading ? (0.5 * lb * (1 - cos(PI * ib / nb))) : (ib * hb);
		for (int ia = 0; ia <= na; ia++) {
			for (int ib = 0; ib <= nb; ib++) {
				double x[3];
				for (int c = 0; c < 3; c++) x[c] = pl->org[c] + (xa[ia] * ta[c]) + (xb[ib] * tb[c]);
				nid[(ia * (nb + 1)) + ib] = nodes_at(nt, x);
			}
		}
		seg_t proto;
		memset(&proto, 0, sizeof(seg_t));
		proto.shape = SHAPE_PLATE;
		proto.thick = pl->thick;
		proto.sigma = pl->sigma;

		const int nlay = (pl->ndivt > 1) ? pl->ndivt : 1;
		const double tl = pl->thick / nlay;
		vcross(ta, tb, nv);

		/* current cells: ribbons (or volume bars when layered) along both grid directions;
		 * layers are stacked symmetrically about the plate plane */
		for (int it = 0; it < nlay; it++) {
			const double zoff = (it - (0.5 * (nlay - 1))) * tl;
			for (int dir = 0; dir < 2; dir++) {
				const int nrun = dir ? nb : na;
				const int nacr = dir ? na : nb;
				const double *xrun = dir ? xb : xa;
				const double *xacr = dir ? xa : xb;
				const double *trun = dir ? tb : ta;
				const double *tacr = dir ? ta : tb;
				for (int j = 0; j <= nacr; j++) {
					const double lo = (j == 0) ? xacr[0] : (0.5 * (xacr[j - 1] + xacr[j]));
					const double hi = (j == nacr) ? xacr[nacr] : (0.5 * (xacr[j] + xacr[j + 1]));
					const double wid = hi - lo;
					const double cen = 0.5 * (lo + hi);
					for (int k = 0; k < nrun; k++) {
						seg_t *s = &p->seg[p->nseg];
						proto.wid = wid;
						*s = proto;
						const int ia1 = dir ? j : k, ib1 = dir ? k : j;
						const int ia2 = dir ? j : (k + 1), ib2 = dir ? (k + 1) : j;
						s->n1 = nid[(ia1 * (nb + 1)) + ib1];
						s->n2 = nid[(ia2 * (nb + 1)) + ib2];
						for (int c = 0; c < 3; c++) {
							const double base = pl->org[c] + (cen * tacr[c]) + (zoff * nv[c]);
							s->x1[c] = base + (xrun[k] * trun[c]);
							s->x2[c] = base + (xrun[k + 1] * trun[c]);
							s->wv[c] = tacr[c];
						}
						s->len = xrun[k + 1] - xrun[k];
						s->vol = (nlay > 1);
						s->thick = (nlay > 1) ? tl : pl->thick;
						s->area = wid * s->thick;
						s->perim = 2 * (wid + s->thick);
						s->width = wid;
						s->aL = 0.2235 * (wid + s->thick);
						s->aP = 0.25 * (wid + s->thick);
						s->res = (pl->sigma > 0) ? s->len / (pl->sigma * s->area) : 0;
						p->nseg++;
					}
				}
			}
		}
		/* charge cells: the dual rectangle of every grid node (half at edges, quarter at corners) */
		for (int ia = 0; ia <= na; ia++) {
			for (int ib = 0; ib <= nb; ib++) {
				const double loa = (ia == 0) ? xa[0] : (0.5 * (xa[ia - 1] + xa[ia]));
				const double hia = (ia == na) ? xa[na] : (0.5 * (xa[ia] + xa[ia + 1]));
				const double lob = (ib == 0) ? xb[0] : (0.5 * (xb[ib - 1] + xb[ib]));
				const double hib = (ib == nb) ? xb[nb] : (0.5 * (xb[ib] + xb[ib + 1]));
				double c0[3], x1[3], x2[3];
				for (int c = 0; c < 3; c++) {
					c0[c] = pl->org[c] + (0.5 * (loa + hia) * ta[c]) + (0.5 * (lob + hib) * tb[c]);
					x1[c] = c0[c] - (0.5 * (hia - loa) * ta[c]);
					x2[c] = c0[c] + (0.5 * (hia - loa) * ta[c]);
				}
				push_charge(p, nid[(ia * (nb + 1)) + ib], x1, x2, hib - lob, tb, &proto);
			}
		}
		free(xa);
		free(xb);
		free(nid);
	}
	return 0;
}

/* polygon area from Newell's normal, 0.5 |sum v_k x v_k+1| (reference polygon.c:46-65) */
double poly_area(const seg_t *s)
{
	double nv[3] = {0, 0, 0};
	for (int k = 0; k < s->npv; k++) {
		double c[3];
		vcross(&s->pv[3 * k], &s->pv[3 * ((k + 1) % s->npv)], c);
		for (int i = 0; i < 3; i++) nv[i] += c[i];
	}
	return 0.5 * sqrt((nv[0] * nv[0]) + (nv[1] * nv[1]) + (nv[2] * nv[2]));
}

/* ---- structured panels (quad / disk) ---- */

static void panel_point(const panel_t *pa, int i, int j, double *x)
{
	if (pa->kind == PANEL_QUAD) {
		const double u = pa->grading ? (0.5 * (1 - cos(PI * i / pa->ndiva))) : ((double)i / pa->ndiva);
		const double v = pa->grading ? (0.5 * (1 - cos(PI * j / pa->ndivb))) : ((double)j / pa->ndivb);
		for (int c = 0; c < 3; c++) {
			x[c] = ((1 - u) * (1 - v) * pa->v[c]) + (u * (1 - v) * pa->v[3 + c]) +
			       (u * v * pa->v[6 + c]) + ((1 - u) * v * pa->v[9 + c]);
		}
		return;
	}
	if (i == 0) {
		for (int c = 0; c < 3; c++) x[c] = pa->org[c];
		return;
	}
	/* in-plane axes from the coordinate axis least aligned with the normal */
	double nn[3], ax[3] = {0, 0, 0}, u[3], v[3];
	const double nl = vnorm(pa->nrm);
	for (int c = 0; c < 3; c++) nn[c] = pa->nrm[c] / nl;
	int imin = 0;
	if (fabs(nn[1]) < fabs(nn[imin])) imin = 1;
	if (fabs(nn[2]) < fabs(nn[imin])) imin = 2;
	ax[imin] = 1;
	vcross(ax, nn, u);
	const double ul = vnorm(u);
	v[0] = ((nn[1] * u[2]) - (nn[2] * u[1])) / ul;
	v[1] = ((nn[2] * u[0]) - (nn[0] * u[2])) / ul;
	v[2] = ((nn[0] * u[1]) - (nn[1] * u[0])) / ul;
	const double r = pa->grading ? (pa->radius * sin(0.5 * PI * i / pa->ndiva)) : (pa->radius * i / pa->ndiva);
	const int jm = ((j % pa->ndivb) + pa->ndivb) % pa->ndivb;
	const double th = 2 * PI * jm / pa->ndivb;
	for (int c = 0; c < 3; c++) x[c] = pa->org[c] + (r * cos(th) * u[c] / ul) + (r * sin(th) * v[c]);
}

static int panel_exists(const panel_t *pa, int i, int j)
{
	if (i < 0) return 0;
	if (pa->kind == PANEL_QUAD) return (i <= pa->ndiva) && (j >= 0) && (j <= pa->ndivb);
	return (i <= pa->ndiva);
}

static void panel_centre(const panel_t *pa, int i, int j, double *x)
{
	double a[3], b[3], c[3], d[3];
	panel_point(pa, i, j, a);
	panel_point(pa, i + 1, j, b);
	panel_point(pa, i + 1, j + 1, c);
	panel_point(pa, i, j + 1, d);
	for (int k = 0; k < 3; k++) x[k] = 0.25 * (a[k] + b[k] + c[k] + d[k]);
}

/* branch cell along grid line (i1,j1)-(i2,j2): strip reaching the neighbouring mid-lines */
static void panel_branch(peec_t *p, const panel_t *pa, int n1, int n2, int i1, int j1, int i2, int j2)
{
	seg_t *s = &p->seg[p->nseg];
	memset(s, 0, sizeof(seg_t));
	s->shape = SHAPE_POLY;
	s->n1 = n1;
	s->n2 = n2;
	panel_point(pa, i1, j1, s->x1);
	panel_point(pa, i2, j2, s->x2);
	s->len = vdist(s->x1, s->x2);
	const int di = j2 - j1, dj = i2 - i1;
	double q[3];
	int n = 0;
	if (panel_exists(pa, i1 - di, j1 - dj) && panel_exists(pa, i2 - di, j2 - dj)) {
		panel_point(pa, i1 - di, j1 - dj, q);
		vmid(s->x1, q, &s->pv[3 * n++]);
		panel_point(pa, i2 - di, j2 - dj, q);
		vmid(s->x2, q, &s->pv[3 * n++]);
	}
	else {
		memcpy(&s->pv[3 * n++], s->x1, 3 * sizeof(double));
		memcpy(&s->pv[3 * n++], s->x2, 3 * sizeof(double));
	}
	if (panel_exists(pa, i2 + di, j2 + dj) && panel_exists(pa, i1 + di, j1 + dj)) {
		panel_point(pa, i2 + di, j2 + dj, q);
		vmid(s->x2, q, &s->pv[3 * n++]);
		panel_point(pa, i1 + di, j1 + dj, q);
		vmid(s->x1, q, &s->pv[3 * n++]);
	}
	else {
		memcpy(&s->pv[3 * n++], s->x2, 3 * sizeof(double));
		memcpy(&s->pv[3 * n++], s->x1, 3 * sizeof(double));
	}
	s->npv = n;
	memcpy(s->papex, &s->pv[0], 3 * sizeof(double));
	const double area = poly_area(s);
	s->wid = area / s->len;
	s->thick = pa->thick;
	s->sigma = pa->sigma;
	s->area = s->wid * pa->thick;
	s->perim = 2 * (s->wid + pa->thick);
	s->width = s->wid;
	s->aL = 0.2235 * (s->wid + pa->thick);
	s->aP = 0.25 * (s->wid + pa->thick);
	s->res = (pa->sigma > 0) ? s->len / (pa->sigma * s->area) : 0;
	p->nseg++;
}

/* sector cells the disk's centre ring is emitted as: 1 while its 2 nsec vertices fit a cell,
 * else the fewest equal groups of nsec / k sectors with 2 nsec / k + 2 <= POLY_MAX vertices.
 * PEEC_DISK_CENTRE_CELLS=k asks for k groups where k divides nsec and they fit (tests compare
 * a split centre with the reference's single cell). */
static int centre_split(int nsec)
{
	const char *e = getenv("PEEC_DISK_CENTRE_CELLS");
	const int want = e ? atoi(e) : 0;
	if ((want >= 1) && (nsec % want == 0) && ((want == 1 ? 2 * nsec : 2 * (nsec / want) + 2) <= POLY_MAX)) return want;
	if (2 * nsec <= POLY_MAX) return 1;
	for (int k = 2; k <= nsec; k++)
		if ((nsec % k == 0) && (2 * (nsec / k) + 2 <= POLY_MAX)) return k;
	return nsec;
}

static void poly_charge_finish(peec_t *p, seg_t *c, const panel_t *pa, const double *v0, int node)
{
	memcpy(c->papex, v0, 3 * sizeof(double));
	memcpy(c->x1, v0, 3 * sizeof(double));
	memcpy(c->x2, v0, 3 * sizeof(double));
	const double area = poly_area(c);
	c->len = sqrt(area);
	c->wid = sqrt(area);
	c->thick = pa->thick;
	c->sigma = pa->sigma;
	c->aP = 0.25 * (c->wid + pa->thick);
	p->chgnode[p->nchg] = node;
	p->nchg++;
}

/* dual polygon of grid node (i,j): ring E, NE, N, NW, W, SW, S, SE of edge midpoints and cell
 * centres; a run of missing neighbours (boundary) contributes the node itself once */
static void panel_charge(peec_t *p, const panel_t *pa, int node, int i, int j)
{
	static const int ring[8][2] = {{0, 1}, {1, 1}, {1, 0}, {1, -1}, {0, -1}, {-1, -1}, {-1, 0}, {-1, 1}};
	seg_t *c = &p->chg[p->nchg];
	memset(c, 0, sizeof(seg_t));
	c->shape = SHAPE_POLY;
	double v0[3], q[3];
	panel_point(pa, i, j, v0);
	int n = 0, gap = 0;
	for (int k = 0; k < 8; k++) {
		const int di = ring[k][0], dj = ring[k][1];
		int ok;
		if ((di != 0) && (dj != 0)) {
			const int ci = (di > 0) ? i : (i - 1);
			const int cj = (dj > 0) ? j : (j - 1);
			ok = panel_exists(pa, ci, cj) && panel_exists(pa, ci + 1, cj + 1);
			if (ok) panel_centre(pa, ci, cj, q);
		}
		else {
			ok = panel_exists(pa, i + di, j + dj);
			if (ok) {
				panel_point(pa, i + di, j + dj, q);
				vmid(v0, q, q);
			}
		}
		if (ok) {
			memcpy(&c->pv[3 * n++], q, 3 * sizeof(double));
		}
		else if (!gap) {
			memcpy(&c->pv[3 * n++], v0, 3 * sizeof(double));
			gap = 1;
		}
	}
	c->npv = n;
	poly_charge_finish(p, c, pa, v0, node);
}

static int build_panels(peec_t *p, nodes_t *nt)
{
	for (int ip = 0; ip < p->npanel; ip++) {
		p->panel[ip].grading = p->grading;
		const panel_t *pa = &p->panel[ip];
		const int na = pa->ndiva, nb = pa->ndivb;
		if (pa->kind == PANEL_QUAD) {
			int *nid = (int *)malloc((size_t)(na + 1) * (nb + 1) * sizeof(int));
			if (nid == NULL) {
				printf("%s\n", "*** memory allocation error (panel)");
				return 1;
			}
			for (int i = 0; i <= na; i++) {
				for (int j = 0; j <= nb; j++) {
					double x[3];
					panel_point(pa, i, j, x);
					nid[(i * (nb + 1)) + j] = nodes_at(nt, x);
				}
			}
			for (int j = 0; j <= nb; j++)
				for (int i = 0; i < na; i++)
					panel_branch(p, pa, nid[(i * (nb + 1)) + j], nid[((i + 1) * (nb + 1)) + j], i, j, i + 1, j);
			for (int i = 0; i <= na; i++)
				for (int j = 0; j < nb; j++)
					panel_branch(p, pa, nid[(i * (nb + 1)) + j], nid[(i * (nb + 1)) + j + 1], i, j, i, j + 1);
			for (int i = 0; i <= na; i++)
				for (int j = 0; j <= nb; j++)
					panel_charge(p, pa, nid[(i * (nb + 1)) + j], i, j);
			free(nid);
			continue;
		}
		/* disk: node [0] is the centre, then ring i (1..nring) sector j at 1 + (i-1) nsec + j */
		int *nid = (int *)malloc((size_t)(1 + (na * nb)) * sizeof(int));
		if (nid == NULL) {
			printf("%s\n", "*** memory allocation error (panel)");
			return 1;
		}
		{
			double x[3];
			panel_point(pa, 0, 0, x);
			nid[0] = nodes_at(nt, x);
		}
		for (int i = 1; i <= na; i++) {
			for (int j = 0; j < nb; j++) {
				double x[3];
				panel_point(pa, i, j, x);
				nid[1 + ((i - 1) * nb) + j] = nodes_at(nt, x);
			}
		}
		for (int j = 0; j < nb; j++) {
			for (int i = 0; i < na; i++) {
				const int m1 = (i == 0) ? 0 : (1 + ((i - 1) * nb) + j);
				panel_branch(p, pa, nid[m1], nid[1 + (i * nb) + j], i, j, i + 1, j);
			}
		}
		for (int i = 1; i <= na; i++) {
			for (int j = 0; j < nb; j++) {
				const int m1 = 1 + ((i - 1) * nb) + j;
				const int m2 = 1 + ((i - 1) * nb) + ((j + 1) % nb);
				panel_branch(p, pa, nid[m1], nid[m2], i, j, i, j + 1);
			}
		}
		{
			/* centre charge cell: ring of spoke midpoints and first-ring cell centres, fanned
			 * from the centre.  A ring of more than POLY_MAX vertices (nsec > 32, beyond the
			 * reference) is emitted as k congruent sector cells on the centre node, each the
			 * fan of nsec / k sectors closed through the centre: their fan triangles are the
			 * ring's own, and with equal areas the per-cell normalisation (len = sqrt(area),
			 * cell area = sum of len) gives the whole cell's potential coefficients exactly */
			double v0[3], q[3];
			panel_point(pa, 0, 0, v0);
			const int k = centre_split(nb);
			const int per = nb / k;
			for (int part = 0; part < k; part++) {
				seg_t *c = &p->chg[p->nchg];
				memset(c, 0, sizeof(seg_t));
				c->shape = SHAPE_POLY;
				int n = 0;
				if (k > 1) memcpy(&c->pv[3 * n++], v0, 3 * sizeof(double));
				for (int j = part * per; j < (part + 1) * per; j++) {
					panel_point(pa, 1, j, q);
					vmid(v0, q, &c->pv[3 * n++]);
					panel_centre(pa, 0, j, &c->pv[3 * n++]);
				}
				if (k > 1) {
					panel_point(pa, 1, ((part + 1) * per) % nb, q);
					vmid(v0, q, &c->pv[3 * n++]);
				}
				c->npv = n;
				poly_charge_finish(p, c, pa, v0, nid[0]);
			}
		}
		for (int i = 1; i <= na; i++)
			for (int j = 0; j < nb; j++)
				panel_charge(p, pa, nid[1 + ((i - 1) * nb) + j], i, j);
		free(nid);
	}
	return 0;
}

/* ---- dielectric bricks (excess-capacitance volume cells) ---- */

static double dual_width(int i, int n, double h)
{
	return ((i == 0) || (i == n)) ? (0.5 * h) : h;
}

static double dual_shift(int i, int n, double h)
{
	return (i == 0) ? (0.25 * h) : (i == n) ? (-0.25 * h) : 0;
}

static void diel_branch(peec_t *p, const double *x1, const double *x2, int n1, int n2,
	const double *wv, double wid, double thk, const diel_t *dl)
{
	seg_t *s = &p->seg[p->nseg];
	memset(s, 0, sizeof(seg_t));
	s->shape = SHAPE_PLATE;
	s->vol = 1;
	s->diel = 1;
	s->n1 = n1;
	s->n2 = n2;
	for (int c = 0; c < 3; c++) {
		s->x1[c] = x1[c];
		s->x2[c] = x2[c];
		s->wv[c] = wv[c];
	}
	s->len = vdist(x1, x2);
	s->wid = wid;
	s->thick = thk;
	s->width = wid;
	s->area = wid * thk;
	s->perim = 2 * (wid + thk);
	s->aL = 0.2235 * (wid + thk);
	s->aP = 0.25 * (wid + thk);
	s->sigma = 0;
	s->res = 0;
	s->cexc = EPS0 * (dl->epsr - 1) * s->area / s->len;
	s->gexc = EPS0 * dl->epsr * dl->tand * s->area / s->len;
	s->gfac = EPS0 * s->area / s->len;
	s->epss = dl->epsr;
	s->epsinf = dl->epsinf;
	s->frelax = dl->frelax;
	p->nseg++;
}

static int build_dielectrics(peec_t *p, nodes_t *nt)
{
	for (int idl = 0; idl < p->ndiel; idl++) {
		const diel_t *dl = &p->diel[idl];
		const int na = dl->ndiva, nb = dl->ndivb, nl = dl->ndivt;
		const double la = vnorm(dl->ea), lb = vnorm(dl->eb);
		const double ha = la / na, hb = lb / nb, ht = dl->thick / nl;
		double ta[3], tb[3], nv[3];
		for (int c = 0; c < 3; c++) {
			ta[c] = dl->ea[c] / la;
			tb[c] = dl->eb[c] / lb;
		}
		vcross(ta, tb, nv);
		int *nid = (int *)malloc((size_t)(na + 1) * (nb + 1) * (nl + 1) * sizeof(int));
		if (nid == NULL) {
			printf("%s\n", "*** memory allocation error (dielectric)");
			return 1;
		}
#define NODE3(ia, ib, it) nid[((((ia) * (nb + 1)) + (ib)) * (nl + 1)) + (it)]
		for (int ia = 0; ia <= na; ia++)
			for (int ib = 0; ib <= nb; ib++)
				for (int it = 0; it <= nl; it++) {
					double x[3];
					for (int c = 0; c < 3; c++)
						x[c] = dl->org[c] + (ia * ha * ta[c]) + (ib * hb * tb[c]) + (it * ht * nv[c]);
					NODE3(ia, ib, it) = nodes_at(nt, x);
				}
		/* branches along a (section b x t), b (a x t), t (a x b) */
		for (int ib = 0; ib <= nb; ib++)
			for (int it = 0; it <= nl; it++) {
				const double wb = dual_width(ib, nb, hb), ob = dual_shift(ib, nb, hb);
				const double wt = dual_width(it, nl, ht), ot = dual_shift(it, nl, ht);
				for (int k = 0; k < na; k++) {
					double x1[3], x2[3];
					for (int c = 0; c < 3; c++) {
						const double base = dl->org[c] + (((ib * hb) + ob) * tb[c]) + (((it * ht) + ot) * nv[c]);
						x1[c] = base + (k * ha * ta[c]);
						x2[c] = base + ((k + 1) * ha * ta[c]);
					}
					diel_branch(p, x1, x2, NODE3(k, ib, it), NODE3(k + 1, ib, it), tb, wb, wt, dl);
				}
			}
		for (int ia = 0; ia <= na; ia++)
			for (int it = 0; it <= nl; it++) {
				const double wa = dual_width(ia, na, ha), oa = dual_shift(ia, na, ha);
				const double wt = dual_width(it, nl, ht), ot = dual_shift(it, nl, ht);
				for (int k = 0; k < nb; k++) {
					double x1[3], x2[3];
					for (int c = 0; c < 3; c++) {
						const double base = dl->org[c] + (((ia * ha) + oa) * ta[c]) + (((it * ht) + ot) * nv[c]);
						x1[c] = base + (k * hb * tb[c]);
						x2[c] = base + ((k + 1) * hb * tb[c]);
					}
					diel_branch(p, x1, x2, NODE3(ia, k, it), NODE3(ia, k + 1, it), ta, wa, wt, dl);
				}
			}
		for (int ia = 0; ia <= na; ia++)
			for (int ib = 0; ib <= nb; ib++) {
				const double wa = dual_width(ia, na, ha), oa = dual_shift(ia, na, ha);
				const double wb = dual_width(ib, nb, hb), ob = dual_shift(ib, nb, hb);
				for (int k = 0; k < nl; k++) {
					double x1[3], x2[3];
					for (int c = 0; c < 3; c++) {
						const double base = dl->org[c] + (((ia * ha) + oa) * ta[c]) + (((ib * hb) + ob) * tb[c]);
						x1[c] = base + (k * ht * nv[c]);
						x2[c] = base + ((k + 1) * ht * nv[c]);
					}
					diel_branch(p, x1, x2, NODE3(ia, ib, k), NODE3(ia, ib, k + 1), ta, wa, wb, dl);
				}
			}
		/* bound-charge cells: a-b dual rectangle of every node, on the node's t plane */
		seg_t proto;
		memset(&proto, 0, sizeof(seg_t));
		proto.shape = SHAPE_PLATE;
		proto.thick = ht;
		for (int ia = 0; ia <= na; ia++)
			for (int ib = 0; ib <= nb; ib++) {
				const double wa = dual_width(ia, na, ha), oa = dual_shift(ia, na, ha);
				const double wb = dual_width(ib, nb, hb), ob = dual_shift(ib, nb, hb);
				for (int it = 0; it <= nl; it++) {
					double x1[3], x2[3];
					for (int c = 0; c < 3; c++) {
						const double c0 = dl->org[c] + (((ia * ha) + oa) * ta[c]) + (((ib * hb) + ob) * tb[c]) + (it * ht * nv[c]);
						x1[c] = c0 - (0.5 * wa * ta[c]);
						x2[c] = c0 + (0.5 * wa * ta[c]);
					}
					push_charge(p, NODE3(ia, ib, it), x1, x2, wb, tb, &proto);
				}
			}
#undef NODE3
		free(nid);
	}
	return 0;
}

/* all cells must lie on or above the ground plane (reference wire.c:862-894) */
static int check_ground(peec_t *p, FILE *fp_log)
{
	for (int i = 0; i < p->nseg + p->nchg; i++) {
		const seg_t *s = (i < p->nseg) ? &p->seg[i] : &p->chg[i - p->nseg];
		double mz;
		if (s->npv > 0) {
			mz = s->pv[2];
			for (int k = 1; k < s->npv; k++)
				if (s->pv[(3 * k) + 2] < mz) mz = s->pv[(3 * k) + 2];
		}
		else {
			mz = (s->x1[2] < s->x2[2]) ? s->x1[2] : s->x2[2];
			if (s->wid > 0) {
				double lo = fabs(0.5 * s->wid * s->wv[2]);
				if (s->vol) {
					double t[3];
					for (int c = 0; c < 3; c++) t[c] = (s->x2[c] - s->x1[c]) / s->len;
					lo += fabs(0.5 * s->thick * ((t[0] * s->wv[1]) - (t[1] * s->wv[0])));
				}
				mz -= lo;
			}
		}
		if (mz < p->gpz - p->nodetol) {
			printf("*** conductor below ground plane (z = %.3e < %.3e)\n", mz, p->gpz);
			if (fp_log) fprintf(fp_log, "*** conductor below ground plane (z = %.3e < %.3e)\n", mz, p->gpz);
			return 1;
		}
	}
	if (fp_log) fprintf(fp_log, "ground plane : z = %g (image method)\n", p->gpz);
	return 0;
}

int wire_build(peec_t *p, FILE *fp_log)
{
	if ((p->nwire <= 0) && (p->nplate <= 0) && (p->npanel <= 0) && (p->ndiel <= 0)) return 0;

	/* capacity bounds (reference wire.c:352-391) */
	int npoint = p->nnodexyz, nseg = 0, nchg = 0;
	for (int i = 0; i < p->nwire; i++) {
		npoint += p->wire[i].ndiv + 1;
		nseg += p->wire[i].ndiv;
		nchg += 2 * p->wire[i].ndiv;
	}
	for (int i = 0; i < p->nplate; i++) {
		const int na = p->plate[i].ndiva, nb = p->plate[i].ndivb;
		const int nl = (p->plate[i].ndivt > 1) ? p->plate[i].ndivt : 1;
		npoint += (na + 1) * (nb + 1);
		nseg += nl * ((na * (nb + 1)) + ((na + 1) * nb));
		nchg += (na + 1) * (nb + 1);
	}
	for (int i = 0; i < p->npanel; i++) {
		const int na = p->panel[i].ndiva, nb = p->panel[i].ndivb;
		if (p->panel[i].kind == PANEL_QUAD) {
			npoint += (na + 1) * (nb + 1);
			nseg += (na * (nb + 1)) + ((na + 1) * nb);
			nchg += (na + 1) * (nb + 1);
		}
		else {
			npoint += 1 + (na * nb);
			nseg += 2 * na * nb;
			nchg += centre_split(nb) + (na * nb);
		}
	}
	for (int i = 0; i < p->ndiel; i++) {
		const int na = p->diel[i].ndiva, nb = p->diel[i].ndivb, nl = p->diel[i].ndivt;
		npoint += (na + 1) * (nb + 1) * (nl + 1);
		nseg += (na * (nb + 1) * (nl + 1)) + ((na + 1) * nb * (nl + 1)) + ((na + 1) * (nb + 1) * nl);
		nchg += (na + 1) * (nb + 1) * (nl + 1);
	}
	p->gid = (int *)malloc((size_t)npoint * sizeof(int));
	p->gxyz = (double *)malloc((size_t)npoint * 3 * sizeof(double));
	p->seg = (seg_t *)malloc((size_t)(nseg > 0 ? nseg : 1) * sizeof(seg_t));
	p->chg = (seg_t *)malloc((size_t)(nchg > 0 ? nchg : 1) * sizeof(seg_t));
	p->chgnode = (int *)malloc((size_t)(nchg > 0 ? nchg : 1) * sizeof(int));
	nodes_t nt;
	nt.p = p;
	nt.nbucket = 2 * npoint + 1;
	nt.head = (int *)malloc((size_t)nt.nbucket * sizeof(int));
	nt.next = (int *)malloc((size_t)npoint * sizeof(int));
	if ((p->gid == NULL) || (p->gxyz == NULL) || (p->seg == NULL) || (p->chg == NULL) ||
	    (p->chgnode == NULL) || (nt.head == NULL) || (nt.next == NULL)) {
		printf("%s\n", "*** memory allocation error (geometry)");
		free(nt.head);
		free(nt.next);
		return 1;
	}
	for (int b = 0; b < nt.nbucket; b++) nt.head[b] = -1;
	p->ngnode = 0;
	p->nseg = 0;
	p->nchg = 0;

	/* seeds: coordinates bound by "node =" (duplicates warned, first wins) */
	for (int i = 0; i < p->nnodexyz; i++) {
		const int found = nodes_find(&nt, &p->ncxyz[3 * i]);
		if (found >= 0) {
			if (fp_log) fprintf(fp_log, "*** warning : node %d and node %d are within nodetol, node %d is used\n",
				p->gid[found], p->ncid[i], p->gid[found]);
			continue;
		}
		nodes_add(&nt, p->ncid[i], &p->ncxyz[3 * i]);
	}
	nt.autoid = p->maxid;

	int rc = 0;
	build_wires(p, &nt);
	if (!rc) rc = build_plates(p, &nt);
	if (!rc) rc = build_panels(p, &nt);
	if (!rc) rc = build_dielectrics(p, &nt);
	p->maxid = nt.autoid;

	free(nt.head);
	free(nt.next);
	if (rc) return rc;
	if (p->gp && check_ground(p, fp_log)) return 1;

	/* near-duplicate node warning (unmerged nodes closer than 10 nodetol), up to 10 lines */
	if (fp_log) {
		int nwarn = 0;
		const double r = 10 * p->nodetol;
		for (int i = 0; (i < p->ngnode) && (nwarn < 10); i++) {
			for (int j = i + 1; (j < p->ngnode) && (nwarn < 10); j++) {
				const double d = vdist(&p->gxyz[3 * i], &p->gxyz[3 * j]);
				if (d < r) {
					fprintf(fp_log, "*** warning : node %d and node %d are separated by only %.3e m\n", p->gid[i], p->gid[j], d);
					nwarn++;
				}
			}
		}
	}
	if (fp_log) fprintf(fp_log, "geometry : %d wires + %d plates + %d panels + %d dielectric bricks -> %d cells, %d charge cells, %d nodes\n",
		p->nwire, p->nplate, p->npanel, p->ndiel, p->nseg, p->nchg, p->ngnode);
	return 0;
}
