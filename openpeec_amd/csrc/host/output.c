/*
 * output.c — the sweep's result files and log tables, from the arrays solve() leaves in peec_t:
 * input impedance (log table and zin.csv), S parameters (log table and Touchstone peec.sNp),
 * current / charge distributions (dist.csv) and the plane-wave response (pw.csv).  The file
 * layouts are the reference's (OpenPEEC src/output.c), so tools reading its files read these;
 * the far field and the transient synthesis run on the device (postproc.hip).
 */
#include "peec.h"

/* one row of the impedance tables: Z, Y = 1/Z in mS, |Z| */
typedef struct {
	double f;
	d_complex_t z, y;
	double zabs;
} zrow_t;

static zrow_t zrow(const peec_t *p, int iport, int ifreq)
{
	zrow_t r;
	r.f = freq_at(p, ifreq);
	r.z = p->zin[(size_t)iport * p->nfreq + ifreq];
	r.y = d_inv(r.z);
	r.zabs = d_abs(r.z);
	return r;
}

static void zin_table(const peec_t *p, FILE *fp)
{
	if (fp == NULL) return;
	fprintf(fp, "=== input impedance ===\n");
	for (int k = 0; k < p->nport; k++) {
		fprintf(fp, "port #%d (Z0[ohm] = %.2f)\n", k + 1, p->port[k].z0);
		fprintf(fp, "  %s\n", "frequency[Hz] Rin[ohm]   Xin[ohm]    Gin[mS]    Bin[mS]   |Z|[ohm]");
		for (int i = 0; i < p->nfreq; i++) {
			const zrow_t r = zrow(p, k, i);
			fprintf(fp, "%13.5e%11.3f%11.3f%11.3f%11.3f%11.3f\n", r.f, r.z.r, r.z.i, r.y.r * 1e3, r.y.i * 1e3, r.zabs);
		}
	}
	fflush(fp);
}

void output_zin(const peec_t *p, FILE *fp_log)
{
	zin_table(p, stdout);
	zin_table(p, fp_log);
}

static FILE *open_out(const char *fn)
{
	FILE *fp = fopen(fn, "w");
	if (fp == NULL) printf("*** file %s open error.\n", fn);
	return fp;
}

int output_csv(const peec_t *p, const char *fn)
{
	FILE *fp = open_out(fn);
	if (fp == NULL) return 1;
	fprintf(fp, "port,frequency[Hz],Rin[ohm],Xin[ohm],Gin[mS],Bin[mS],Zabs[ohm]\n");
	for (int k = 0; k < p->nport; k++)
		for (int i = 0; i < p->nfreq; i++) {
			const zrow_t r = zrow(p, k, i);
			fprintf(fp, "%d,%.9e,%.9e,%.9e,%.9e,%.9e,%.9e\n", k + 1, r.f, r.z.r, r.z.i, r.y.r * 1e3, r.y.i * 1e3, r.zabs);
		}
	fclose(fp);
	return 0;
}

static void spara_table(const peec_t *p, FILE *fp)
{
	if (fp == NULL) return;
	const int np = p->nport;
	fprintf(fp, "=== S parameters ===\n");
	fprintf(fp, "reference impedance [ohm] :");
	for (int k = 0; k < np; k++) fprintf(fp, " port#%d=%.2f", k + 1, p->port[k].z0);
	fprintf(fp, "\n");
	for (int e = 0; e < np * np; e++) {
		const int i = e / np, j = e % np;
		fprintf(fp, "S%d%d\n", i + 1, j + 1);
		fprintf(fp, "  %s\n", "frequency[Hz]      real      imag   |S|[dB]  phase[deg]");
		for (int q = 0; q < p->nfreq; q++) {
			const d_complex_t v = p->smat[ZIDX(p, q, i, j)];
			const double mag = d_abs(v);
			fprintf(fp, "%13.5e%10.5f%10.5f%10.3f%12.3f\n", freq_at(p, q), v.r, v.i,
			        (mag > 1e-30) ? (20 * log10(mag)) : -600, atan2(v.i, v.r) * 180 / PI);
		}
	}
	fflush(fp);
}

void output_spara(const peec_t *p, FILE *fp_log)
{
	if ((p->nport <= 0) || (p->smat == NULL)) return;
	spara_table(p, stdout);
	spara_table(p, fp_log);
}

/* Touchstone 1.1, real/imaginary pairs, one reference resistance (port #1's; differing z0 are
 * listed in comments).  Two-port files list S11 S21 S12 S22 (the format's column-major rule);
 * three ports and more list rows S_i1 .. S_iN, continuation lines indented 15 columns. */
int output_touchstone(const peec_t *p)
{
	const int np = p->nport;
	if ((np <= 0) || (p->smat == NULL)) return 0;
	char fn[64];
	snprintf(fn, sizeof(fn), "peec.s%dp", np);
	FILE *fp = open_out(fn);
	if (fp == NULL) return 1;
	int uneven = 0;
	for (int k = 1; k < np; k++) uneven |= (p->port[k].z0 != p->port[0].z0);
	fprintf(fp, "! %s Ver.%d.%d.%d\n", PROGRAM, VERSION_MAJOR, VERSION_MINOR, VERSION_BUILD);
	fprintf(fp, "! %s\n", p->title);
	if (uneven) {
		fprintf(fp, "! NOTE: ports have different reference impedances;");
		fprintf(fp, " Touchstone 1.1 can only record one, so port#1 is used.\n");
		for (int k = 0; k < np; k++) fprintf(fp, "!   port#%d z0 = %.6g ohm\n", k + 1, p->port[k].z0);
	}
	fprintf(fp, "# Hz S RI R %.6g\n", p->port[0].z0);
	for (int q = 0; q < p->nfreq; q++) {
		fprintf(fp, "%.9e", freq_at(p, q));
		for (int e = 0; e < np * np; e++) {
			/* entry e of the line: column-major for two ports, row-major otherwise */
			const int i = (np == 2) ? (e % 2) : (e / np);
			const int j = (np == 2) ? (e / 2) : (e % np);
			if ((np != 2) && (j == 0) && (i > 0)) fprintf(fp, "\n%*s", 15, "");
			const d_complex_t v = p->smat[ZIDX(p, q, i, j)];
			fprintf(fp, " %.9e %.9e", v.r, v.i);
		}
		fprintf(fp, "\n");
	}
	fclose(fp);
	return 0;
}

/* Plane-wave response: open-circuit port voltage, effective length |Voc| / E0 and the power
 * available to a conjugate load |Voc|^2 / (8 Re Zin), in the log and in pw.csv. */
int output_pw(const peec_t *p, const char *fn, FILE *fp_log)
{
	if (!p->pw || (p->voc == NULL)) return 0;
	FILE *fp = open_out(fn);
	if (fp == NULL) return 1;
	fprintf(fp, "port,frequency[Hz],Voc_real[V],Voc_imag[V],Voc_abs[V],leff[m],Pav[W]\n");
	if (fp_log)
		fprintf(fp_log, "=== plane wave response === (arrival theta = %.2f deg, phi = %.2f deg, %s pol., E0 = %.4g V/m)\n", p->pwth,
		        p->pwph, (p->pwpol == 2) ? "phi" : "theta", p->pwamp);
	for (int k = 0; k < p->nport; k++) {
		if (fp_log) {
			fprintf(fp_log, "port #%d\n", k + 1);
			fprintf(fp_log, "  %s\n", "frequency[Hz]  |Voc|[V]   leff[m]      Pav[W]");
		}
		for (int i = 0; i < p->nfreq; i++) {
			const double f = freq_at(p, i);
			const d_complex_t v = p->voc[(size_t)k * p->nfreq + i];
			const double va = d_abs(v), rin = p->zin[(size_t)k * p->nfreq + i].r;
			const double pav = (rin > 0) ? (va * va / (8 * rin)) : 0;
			if (fp_log) fprintf(fp_log, "%13.5e%11.4e%11.4e%12.4e\n", f, va, va / p->pwamp, pav);
			fprintf(fp, "%d,%.9e,%.9e,%.9e,%.9e,%.9e,%.9e\n", k + 1, f, v.r, v.i, va, va / p->pwamp, pav);
		}
	}
	if (fp_log) fflush(fp_log);
	fclose(fp);
	return 0;
}

static void dist_row(FILE *fp, const char *type, int port, int index, double f, const double *xc, d_complex_t v)
{
	fprintf(fp, "%s,%d,%d,%.9e,%.9e,%.9e,%.9e,%.9e,%.9e,%.9e\n", type, port, index, f, xc[0], xc[1], xc[2], v.r, v.i, d_abs(v));
}

static void seg_mid(const seg_t *s, double *xc)
{
	for (int d = 0; d < 3; d++) xc[d] = 0.5 * (s->x1[d] + s->x2[d]);
}

/* Distributions (distribution = 1): per frequency the branch currents of each driven port (I,
 * at the cell midpoints), the plane wave's induced currents (Ipw, port 0) and the cell charges
 * of each driven port (Q, at the mean midpoint of the cell's charge sub-cells). */
int output_dist(const peec_t *p, const char *fn)
{
	if (!p->dist) return 0;
	if ((p->segi == NULL) && (p->cellq == NULL) && (p->segipw == NULL)) return 0;
	FILE *fp = open_out(fn);
	if (fp == NULL) return 1;
	fprintf(fp, "type,port,index,frequency[Hz],x[m],y[m],z[m],real,imag,abs\n");
	/* cell positions once: the mean of their sub-cells' midpoints, summed in sub-cell order */
	double *cpos = NULL;
	if (p->cellq != NULL) {
		cpos = (double *)calloc((size_t)p->ncell * 4, sizeof(double));
		if (cpos == NULL) {
			fclose(fp);
			return 1;
		}
		for (int h = 0; h < p->nchg; h++) {
			double *c = &cpos[4 * (size_t)p->cellof[h]];
			for (int d = 0; d < 3; d++) c[d] += 0.5 * (p->chg[h].x1[d] + p->chg[h].x2[d]);
			c[3] += 1;
		}
		for (int m = 0; m < p->ncell; m++)
			if (cpos[4 * (size_t)m + 3] > 0)
				for (int d = 0; d < 3; d++) cpos[4 * (size_t)m + d] /= cpos[4 * (size_t)m + 3];
	}
	double xc[3];
	for (int i = 0; i < p->nfreq; i++) {
		const double f = freq_at(p, i);
		if (p->segi != NULL)
			for (int j = 0; j < p->nport; j++)
				for (int m = 0; m < p->nseg; m++) {
					seg_mid(&p->seg[m], xc);
					dist_row(fp, "I", j + 1, m, f, xc, p->segi[DIDX(p, i, j, m)]);
				}
		if (p->segipw != NULL)
			for (int m = 0; m < p->nseg; m++) {
				seg_mid(&p->seg[m], xc);
				dist_row(fp, "Ipw", 0, m, f, xc, p->segipw[(size_t)i * p->nseg + m]);
			}
		if (p->cellq != NULL)
			for (int j = 0; j < p->nport; j++)
				for (int m = 0; m < p->ncell; m++) dist_row(fp, "Q", j + 1, m, f, &cpos[4 * (size_t)m], p->cellq[QIDX(p, i, j, m)]);
	}
	free(cpos);
	fclose(fp);
	return 0;
}
