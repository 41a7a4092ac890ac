/*
 * api.c — the peec.h entry points of the hot path, backed by the device (no CPU compute path):
 * lp_fill / pot_fill / lp_entry / pot_entry (partial.c:371-473, potential.c:124-303),
 * pot_cells (potential.c:29-114, 168-175) and solve (solve.c:171-430), plus the extension
 * entry points of peec_mi355x.h.  Host arrays keep the reference's ownership: p->lp and
 * p->cmat are malloc'ed here on first use and released by peec_free.
 */
#include "peec.h"
#include "peec_mi355x.h"
#include "peec_internal.h"

/* ---------------------------------------------------------------- capacitive cells ---- */

/* cells = distinct nodes of the charge cells in order of first appearance; carea = sum of the
 * sub-cell lengths (reference potential.c:29-62) */
static int build_cells(peec_t *p)
{
	const int nh = p->nchg;
	p->cellof = (int *)malloc((size_t)nh * sizeof(int));
	p->cellid = (int *)malloc((size_t)nh * sizeof(int));
	if ((p->cellof == NULL) || (p->cellid == NULL)) return 1;
	/* first-appearance numbering with a node -> cell lookup table */
	int maxnode = 0;
	for (int h = 0; h < nh; h++)
		if (p->chgnode[h] > maxnode) maxnode = p->chgnode[h];
	int *slot = (int *)malloc((size_t)(maxnode + 1) * sizeof(int));
	if (slot == NULL) return 1;
	for (int i = 0; i <= maxnode; i++) slot[i] = -1;
	p->ncell = 0;
	for (int h = 0; h < nh; h++) {
		const int node = p->chgnode[h];
		if (slot[node] < 0) {
			slot[node] = p->ncell;
			p->cellid[p->ncell++] = node;
		}
		p->cellof[h] = slot[node];
	}
	free(slot);
	p->carea = (double *)calloc((size_t)p->ncell, sizeof(double));
	if (p->carea == NULL) return 1;
	for (int h = 0; h < nh; h++) p->carea[p->cellof[h]] += p->chg[h].len;
	return 0;
}

/* cell -> sub-cell CSR (ascending sub-cell index) and length-weighted centroids
 * (reference potential.c:72-114) */
static int build_cellsub(peec_t *p)
{
	const int nh = p->nchg, nc = p->ncell;
	p->csoff = (int *)calloc((size_t)(nc + 1), sizeof(int));
	p->csidx = (int *)malloc((size_t)(nh > 0 ? nh : 1) * sizeof(int));
	p->cpt = (double *)malloc((size_t)(nc > 0 ? nc : 1) * 3 * sizeof(double));
	if ((p->csoff == NULL) || (p->csidx == NULL) || (p->cpt == NULL)) return 1;
	for (int h = 0; h < nh; h++) p->csoff[p->cellof[h] + 1]++;
	for (int i = 0; i < nc; i++) p->csoff[i + 1] += p->csoff[i];
	int *fill = (int *)malloc((size_t)(nc + 1) * sizeof(int));
	if (fill == NULL) return 1;
	memcpy(fill, p->csoff, (size_t)(nc + 1) * sizeof(int));
	for (int h = 0; h < nh; h++) p->csidx[fill[p->cellof[h]]++] = h;
	free(fill);
	for (int i = 0; i < nc; i++) {
		double w = 0, x[3] = {0, 0, 0};
		for (int e = p->csoff[i]; e < p->csoff[i + 1]; e++) {
			const seg_t *s = &p->chg[p->csidx[e]];
			for (int c = 0; c < 3; c++) x[c] += s->len * 0.5 * (s->x1[c] + s->x2[c]);
			w += s->len;
		}
		for (int c = 0; c < 3; c++) p->cpt[(3 * i) + c] = (w > 0) ? (x[c] / w) : 0;
	}
	return 0;
}

int pot_cells(peec_t *p)
{
	if (p->cellof != NULL) return 0;
	if (p->nchg <= 0) return 1;
	if (build_cells(p)) return 1;
	return build_cellsub(p);
}

/* ---------------------------------------------------------------- fills ---------------- */

void lp_fill(peec_t *p, double f, FILE *fp_log)
{
	const int n = p->nseg;
	if (n <= 0) return;
	if (p->lp == NULL) {
		p->lp = (d_complex_t *)malloc((size_t)n * n * sizeof(d_complex_t));
		if (p->lp == NULL) pgpu_fatal("host allocation of p->lp failed");
	}
	const double kw = p->retardation ? (2 * PI * f / C0) : 0;
	pgpu_t *g = pgpu_get(p);
	pgpu_fill_lp(g, kw);
	pgpu_lp_to_host(g, p->lp);

	/* one-time report of unphysical coupling |k| > 1 (overlapping conductors) */
	if (p->lpwarn) return;
	p->lpwarn = 1;
	int nwarn = 0;
	for (int i = 0; (i < n) && (nwarn < 10); i++) {
		for (int j = i + 1; (j < n) && (nwarn < 10); j++) {
			const double m = fabs(p->lp[(size_t)i * n + j].r);
			const double s = sqrt(fabs(p->lp[(size_t)i * n + i].r * p->lp[(size_t)j * n + j].r));
			if (m > s) {
				if (fp_log) fprintf(fp_log, "*** warning : |k| > 1 between segment %d and %d (overlapping conductors?)\n", i, j);
				nwarn++;
			}
		}
	}
}

int pot_fill(peec_t *p, double f, FILE *fp_log)
{
	const char errmem[] = "*** memory allocation error (potential)";
	if (!p->capacitance || (p->nchg <= 0)) return 0;
	if (pot_cells(p)) {
		printf("%s\n", errmem);
		return 1;
	}
	if (p->compress) {
		if (!p->clogged) {
			p->clogged = 1;
			printf("PEEC: capacitive cells = %d (compressed : charge unknowns)\n", p->ncell);
			if (fp_log) {
				fprintf(fp_log, "PEEC: capacitive cells = %d (compressed : charge unknowns)\n", p->ncell);
				fflush(fp_log);
			}
		}
		return 0;
	}
	const int nc = p->ncell;
	const double kw = p->retardation ? (2 * PI * f / C0) : 0;
	pgpu_t *g = pgpu_get(p);
	pgpu_fill_pot(g, kw);
	int ising = -1;
	if (pgpu_cmat(g, &ising)) {
		printf("*** singular potential coefficient matrix (row %d)\n", ising);
		if (fp_log) fprintf(fp_log, "*** singular potential coefficient matrix (row %d)\n", ising);
		return 1;
	}
	if (p->cmat == NULL) {
		p->cmat = (d_complex_t *)malloc((size_t)nc * nc * sizeof(d_complex_t));
		if (p->cmat == NULL) {
			printf("%s\n", errmem);
			return 1;
		}
	}
	pgpu_cmat_to_host(g, p->cmat);
	if (!p->clogged) {
		p->clogged = 1;
		double ctotal = 0;
		for (int i = 0; i < nc; i++)
			for (int j = 0; j < nc; j++) ctotal += p->cmat[(size_t)i * nc + j].r;
		printf("PEEC: capacitive cells = %d, total capacitance = %.6e F\n", nc, ctotal);
		if (fp_log) {
			fprintf(fp_log, "PEEC: capacitive cells = %d, total capacitance = %.6e F\n", nc, ctotal);
			fflush(fp_log);
		}
	}
	return 0;
}

d_complex_t lp_entry(const peec_t *p, int i, int j, double kw)
{
	return pgpu_lp_entry(pgpu_get(p), i, j, kw);
}

d_complex_t pot_entry(const peec_t *p, int i, int j, double kw)
{
	return pgpu_pot_entry(pgpu_get(p), i, j, kw);
}

double lp_self(double l, double a)
{
	return MU0 / (4 * PI) * neumann_self(l, a);
}

/* ---------------------------------------------------------------- sweep ---------------- */

/* S = sqrt(z0_j / z0_i) (Z - Z0)(Z + Z0)^-1 on the nport x nport host block (solve.c:127-168);
 * the port block is tiny, so it is eliminated here with the reference's pivoting rule. */
static int z_to_s(peec_t *p, int ifreq)
{
	const int np = p->nport;
	d_complex_t *m = (d_complex_t *)malloc((size_t)np * np * sizeof(d_complex_t));
	d_complex_t *inv = (d_complex_t *)malloc((size_t)np * np * sizeof(d_complex_t));
	int *piv = (int *)malloc((size_t)np * sizeof(int));
	int rc = 0;
	for (int i = 0; i < np; i++)
		for (int j = 0; j < np; j++) {
			d_complex_t v = p->zmat[ZIDX(p, ifreq, i, j)];
			if (i == j) v = d_add(v, d_complex(p->port[i].z0, 0));
			m[(size_t)i * np + j] = v;
		}
	for (int k = 0; k < np; k++) {
		int ip = k;
		double amax = fabs(m[(size_t)k * np + k].r) + fabs(m[(size_t)k * np + k].i);
		for (int i = k + 1; i < np; i++) {
			const double v = fabs(m[(size_t)i * np + k].r) + fabs(m[(size_t)i * np + k].i);
			if (v > amax) {
				amax = v;
				ip = i;
			}
		}
		piv[k] = ip;
		if (amax < 1e-30) {
			rc = 1;
			goto out;
		}
		if (ip != k)
			for (int j = 0; j < np; j++) {
				const d_complex_t t = m[(size_t)k * np + j];
				m[(size_t)k * np + j] = m[(size_t)ip * np + j];
				m[(size_t)ip * np + j] = t;
			}
		const d_complex_t pinv = d_inv(m[(size_t)k * np + k]);
		for (int i = k + 1; i < np; i++) {
			const d_complex_t l = d_mul(m[(size_t)i * np + k], pinv);
			m[(size_t)i * np + k] = l;
			for (int j = k + 1; j < np; j++)
				m[(size_t)i * np + j] = d_sub(m[(size_t)i * np + j], d_mul(l, m[(size_t)k * np + j]));
		}
	}
	for (int c = 0; c < np; c++) {
		d_complex_t *b = &inv[(size_t)c * np];   /* column c, stored contiguously */
		for (int i = 0; i < np; i++) b[i] = d_complex(i == c ? 1 : 0, 0);
		for (int k = 0; k < np; k++)
			if (piv[k] != k) {
				const d_complex_t t = b[k];
				b[k] = b[piv[k]];
				b[piv[k]] = t;
			}
		for (int i = 1; i < np; i++) {
			d_complex_t s = b[i];
			for (int j = 0; j < i; j++) s = d_sub(s, d_mul(m[(size_t)i * np + j], b[j]));
			b[i] = s;
		}
		for (int i = np - 1; i >= 0; i--) {
			d_complex_t s = b[i];
			for (int j = i + 1; j < np; j++) s = d_sub(s, d_mul(m[(size_t)i * np + j], b[j]));
			b[i] = d_mul(s, d_inv(m[(size_t)i * np + i]));
		}
	}
	for (int i = 0; i < np; i++)
		for (int j = 0; j < np; j++) {
			d_complex_t s = d_complex(0, 0);
			for (int k = 0; k < np; k++) {
				d_complex_t nk = p->zmat[ZIDX(p, ifreq, i, k)];
				if (i == k) nk = d_sub(nk, d_complex(p->port[i].z0, 0));
				s = d_add(s, d_mul(nk, inv[(size_t)j * np + k]));
			}
			p->smat[ZIDX(p, ifreq, i, j)] = d_rmul(sqrt(p->port[j].z0 / p->port[i].z0), s);
		}
out:
	free(m);
	free(inv);
	free(piv);
	return rc;
}

/* node voltages of the sources' solution, every used node id (solve.c:101-112) */
static void log_node_voltages(const peec_t *p, const d_complex_t *x, double f, FILE *fp_log)
{
	if (fp_log == NULL) return;
	fprintf(fp_log, "=== node voltage === (f = %.5e Hz)\n", f);
	fprintf(fp_log, "  %s\n", "node      real[V]      imag[V]       |V|[V]");
	for (int id = 0; id <= p->maxid; id++) {
		const int row = p->nodemap[id];
		if (row == -2) continue;
		const d_complex_t v = (row < 0) ? d_complex(0, 0) : x[row];
		fprintf(fp_log, "%6d %12.4e %12.4e %12.4e\n", id, v.r, v.i, d_abs(v));
	}
}

/* port i's terminal voltage V(n1) - V(n2) in the solution x */
static d_complex_t port_voltage(const peec_t *p, int i, const d_complex_t *x)
{
	const int r1 = p->nodemap[p->port[i].n1], r2 = p->nodemap[p->port[i].n2];
	const d_complex_t v1 = (r1 < 0) ? d_complex(0, 0) : x[r1];
	const d_complex_t v2 = (r2 < 0) ? d_complex(0, 0) : x[r2];
	return d_sub(v1, v2);
}

static void say(FILE *fp_log, const char *msg, double f)
{
	printf(msg, f);
	if (fp_log) fprintf(fp_log, msg, f);
}

/* The sweep of the reference's solve() (solve.c:171-430), device resident per point.  Right-hand
 * sides of a point: each port driven by 1 A (Z column j, distributions), then the plane wave
 * (open-circuit port voltages, induced currents) and the independent sources (node voltages to
 * the log), all solved with the point's factorisation -- a fresh LU, the kept LU with GMRES
 * (acceleration = 1) or the compressed operator with its preconditioner (compression = 1). */
int solve(peec_t *p, FILE *fp_log)
{
	const int np = p->nport, n = p->nunknown, nf = p->nfreq;
	const int npw = p->pw ? 1 : 0, nsr = (p->nsrc > 0) ? 1 : 0;
	const int nr = np + npw + nsr;
	const size_t nmat = (size_t)nf * np * np;
	p->zin = (d_complex_t *)malloc((size_t)np * nf * sizeof(d_complex_t));
	p->zmat = (d_complex_t *)malloc(nmat * sizeof(d_complex_t));
	p->smat = (d_complex_t *)malloc(nmat * sizeof(d_complex_t));
	if (p->pw) {
		p->voc = (d_complex_t *)malloc((size_t)np * nf * sizeof(d_complex_t));
		if (p->dist && (p->nseg > 0)) p->segipw = (d_complex_t *)malloc((size_t)nf * p->nseg * sizeof(d_complex_t));
	}
	if ((p->dist || (p->ffnth > 0)) && (p->nseg > 0))
		p->segi = (d_complex_t *)malloc((size_t)nf * np * p->nseg * sizeof(d_complex_t));
	if (p->dist && p->capacitance && (p->ncell > 0))
		p->cellq = (d_complex_t *)malloc((size_t)nf * np * p->ncell * sizeof(d_complex_t));
	d_complex_t *rhs = (d_complex_t *)malloc((size_t)n * (nr > 0 ? nr : 1) * sizeof(d_complex_t));
	d_complex_t *x = (d_complex_t *)malloc((size_t)n * (nr > 0 ? nr : 1) * sizeof(d_complex_t));
	d_complex_t *q = (p->cellq && !p->compress) ? (d_complex_t *)malloc((size_t)p->ncell * np * sizeof(d_complex_t)) : NULL;
	int ret = 1;
	if ((p->zin == NULL) || (p->zmat == NULL) || (p->smat == NULL) || (rhs == NULL) || (x == NULL) ||
	    (p->pw && (p->voc == NULL)) || (p->cellq && !p->compress && (q == NULL))) {
		printf("%s\n", "*** memory allocation error (matrix)");
		goto done;
	}
	pgpu_t *g = pgpu_get(p);
	if (p->compress && (nf > 0)) pgpu_kf_band(g, p, freq_at(p, 0), freq_at(p, nf - 1));
	pgpu_accel_reset(g);
	pgpu_accel_stats_t as = {0, 0, 0};
	long citsum = 0;
	int cnrhs = 0;
	double cmem = 0;
	for (int ifreq = 0; ifreq < nf; ifreq++) {
		const double f = freq_at(p, ifreq);
		for (int j = 0; j < np; j++) mna_rhs_port(p, j, rhs + (size_t)j * n);
		if (npw) mna_rhs_planewave(p, f, rhs + (size_t)np * n);
		if (nsr) mna_rhs_sources(p, rhs + (size_t)(np + npw) * n);
		int rc;
		if (p->compress) {
			rc = pgpu_solve_point_compressed(g, p, ifreq, nr, rhs, x, &citsum, &cmem, NULL);
			cnrhs += nr;
		}
		else {
			rc = pgpu_solve_point(g, p, ifreq, nr, rhs, x, q ? np : 0, q, p->accel, &as);
		}
		if (rc == 1) {
			say(fp_log, "*** singular potential coefficient matrix at f = %.5e Hz\n", f);
			goto done;
		}
		if (rc) {
			say(fp_log, p->compress ? "*** compression : solve failed at f = %.5e Hz\n" : "*** singular matrix at f = %.5e Hz\n", f);
			goto done;
		}
		for (int j = 0; j < np; j++) {
			const d_complex_t *xj = x + (size_t)j * n;
			for (int i = 0; i < np; i++) p->zmat[ZIDX(p, ifreq, i, j)] = port_voltage(p, i, xj);
			p->zin[(size_t)j * nf + ifreq] = p->zmat[ZIDX(p, ifreq, j, j)];
			if (p->segi != NULL) memcpy(&p->segi[DIDX(p, ifreq, j, 0)], xj + p->offS, (size_t)p->nseg * sizeof(d_complex_t));
			if (p->cellq != NULL)
				memcpy(&p->cellq[QIDX(p, ifreq, j, 0)], p->compress ? (xj + p->offQ) : (q + (size_t)j * p->ncell),
				       (size_t)p->ncell * sizeof(d_complex_t));
		}
		if (z_to_s(p, ifreq)) {
			say(fp_log, "*** singular (Z + Z0) at f = %.5e Hz\n", f);
			goto done;
		}
		if (npw) {
			const d_complex_t *xw = x + (size_t)np * n;
			for (int i = 0; i < np; i++) p->voc[(size_t)i * nf + ifreq] = port_voltage(p, i, xw);
			if (p->segipw != NULL) memcpy(&p->segipw[(size_t)ifreq * p->nseg], xw + p->offS, (size_t)p->nseg * sizeof(d_complex_t));
		}
		if (nsr) log_node_voltages(p, x + (size_t)(np + npw) * n, f, fp_log);
	}
	if (p->accel && fp_log)
		fprintf(fp_log, "acceleration : %d LU for %d frequencies, %d GMRES solves (avg %.1f iterations)\n", as.nlu, nf, as.ngm,
		        (as.ngm > 0) ? ((double)as.itsum / as.ngm) : 0.0);
	if (p->compress && (cnrhs > 0) && fp_log) {
		const double dm = hmat_dense_mb(n) + hmat_dense_mb(p->nseg) + (p->capacitance ? hmat_dense_mb(p->ncell) : 0);
		fprintf(fp_log, "compression : %d GMRES solves (avg %.1f iterations), %.2f MB (dense MNA + Lp %.2f MB, %.1fx)\n", cnrhs,
		        (double)citsum / cnrhs, cmem, dm, (cmem > 0) ? (dm / cmem) : 0);
	}
	ret = 0;
done:
	free(rhs);
	free(x);
	free(q);
	return ret;
}

/* ---------------------------------------------------------------- extensions ----------- */

const char *peec_mi355x_version(void)
{
	return "libpeec_mi355x 0.1 gfx950";
}

int peec_device_available(void)
{
	return pgpu_device_ok();
}

int peec_load_deck(const char *path, peec_t *p)
{
	FILE *fp = fopen(path, "r");
	if (fp == NULL) {
		printf("*** %s open error\n", path);
		memset(p, 0, sizeof(peec_t));
		return 1;
	}
	const int rc = input_data(fp, p);
	fclose(fp);
	return rc;
}

int peec_load_deck_string(const char *text, peec_t *p)
{
	FILE *fp = fmemopen((void *)text, strlen(text), "r");
	if (fp == NULL) return 1;
	const int rc = input_data(fp, p);
	fclose(fp);
	return rc;
}

int peec_setup(peec_t *p, FILE *fp_log)
{
	if (wire_build(p, fp_log)) return 1;
	if (!p->retardation) {
		if (!p->compress) lp_fill(p, 0, fp_log);
		if (pot_fill(p, 0, fp_log)) return 1;
	}
	else if (p->capacitance && (p->nchg > 0) && pot_cells(p)) {
		return 1;
	}
	return mna_numbering(p, fp_log);
}

int peec_sweep_ex(peec_t *p, int ifreq0, int nfreq, int form, int batch, d_complex_t *zmat, peec_stats_t *st)
{
	if (st) {
		memset(st, 0, sizeof(*st));
		st->form = form;
		st->ising = -1;
	}
	if (p->compress && (form == PEEC_FORM_LEARNED)) return pgpu_sweep_learned(pgpu_get(p), p, ifreq0, nfreq, zmat, st, 1e-8);
	if (p->compress) return pgpu_sweep_compressed(pgpu_get(p), p, ifreq0, nfreq, zmat, st);
	if (form == PEEC_FORM_PLAIN) return pgpu_sweep(pgpu_get(p), p, ifreq0, nfreq, zmat, NULL, NULL, st, 0);
	if (form == PEEC_FORM_PRECISION) return pgpu_sweep(pgpu_get(p), p, ifreq0, nfreq, zmat, NULL, NULL, st, 1);
	if (form == PEEC_FORM_MIXED) return pgpu_sweep_fast(pgpu_get(p), p, ifreq0, nfreq, zmat, st, batch);
	printf("%s\n", "*** peec_sweep: unknown solver form");
	return 4;
}

int peec_sweep(peec_t *p, int ifreq0, int nfreq, int form, d_complex_t *zmat, peec_stats_t *st)
{
	return peec_sweep_ex(p, ifreq0, nfreq, form, 4, zmat, st);
}

void peec_sweep_reset(void)
{
	pgpu_fast_reset();
	pgpu_kf_reset_all();
}

/* wavenumber band of the deck's sweep together with f */
static void deck_band(const peec_t *p, double f, double *klo, double *khi)
{
	double lo = f, hi = f;
	const double fs[2] = {freq_at(p, 0), freq_at(p, (p->nfreq > 0) ? p->nfreq - 1 : 0)};
	for (int i = 0; i < 2; i++) {
		if (fs[i] < lo) lo = fs[i];
		if (fs[i] > hi) hi = fs[i];
	}
	*klo = 2 * PI * lo / C0;
	*khi = 2 * PI * hi / C0;
}

int peec_fill_matrices(peec_t *p, double f, int mode, d_complex_t *lp, d_complex_t *P)
{
	if (p->capacitance && (p->nchg > 0) && pot_cells(p)) return 2;
	pgpu_t *g = pgpu_get(p);
	const double kw = p->retardation ? (2 * PI * f / C0) : 0;
	if (mode == 1) {
		double klo, khi;
		deck_band(p, f, &klo, &khi);
		if (!p->retardation || pgpu_kf_setup(g, klo, khi)) return 1;
	}
	pgpu_fill_to_host(g, kw, mode, lp, P);
	return 0;
}

size_t peec_sizeof_peec(void) { return sizeof(peec_t); }
size_t peec_sizeof_seg(void) { return sizeof(seg_t); }
