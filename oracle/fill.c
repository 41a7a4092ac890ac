/*
 * fill.c — oracle restatement of the partial-element integrals and the dense Lp / P fills.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Reference: OpenPEEC src/partial.c, surface.c,
 * volume.c, polygon.c, potential.c; line numbers cited per function.
 */
#include "oracle.h"

/* ---- quadrature tables: Gauss-Legendre on [-1, 1] and triangle rules (partial.c:32-36,
 *      surface.c:26-37, polygon.c:134-152) ---- */
static const double X8[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290, -0.1834346424956498,
                             0.1834346424956498, 0.5255324099163290, 0.7966664774136267, 0.9602898564975363};
static const double W8[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873, 0.3626837833783620,
                             0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763};
static const double X4[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
static const double W4[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};
static const double X2[2] = {-0.5773502691896257, 0.5773502691896257};
static const double W2[2] = {1.0, 1.0};
static const double T7W[7] = {0.225, 0.1323941527885062, 0.1323941527885062, 0.1323941527885062,
                              0.1259391805448271, 0.1259391805448271, 0.1259391805448271};
static const double T7B[7][3] = {{1.0 / 3, 1.0 / 3, 1.0 / 3},
                                 {0.0597158717897698, 0.4701420641051151, 0.4701420641051151},
                                 {0.4701420641051151, 0.0597158717897698, 0.4701420641051151},
                                 {0.4701420641051151, 0.4701420641051151, 0.0597158717897698},
                                 {0.7974269853530873, 0.1012865073234563, 0.1012865073234563},
                                 {0.1012865073234563, 0.7974269853530873, 0.1012865073234563},
                                 {0.1012865073234563, 0.1012865073234563, 0.7974269853530873}};
static const double T3W[3] = {1.0 / 3, 1.0 / 3, 1.0 / 3};
static const double T3B[3][3] = {{2.0 / 3, 1.0 / 6, 1.0 / 6}, {1.0 / 6, 2.0 / 3, 1.0 / 6}, {1.0 / 6, 1.0 / 6, 2.0 / 3}};
static const double SPLIT[4][3][3] = {{{1, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}},
                                      {{0.5, 0.5, 0}, {0, 1, 0}, {0, 0.5, 0.5}},
                                      {{0.5, 0, 0.5}, {0, 0.5, 0.5}, {0, 0, 1}},
                                      {{0.5, 0.5, 0}, {0, 0.5, 0.5}, {0.5, 0, 0.5}}};

static double dot(const double *a, const double *b)
{
	return (a[0] * b[0]) + (a[1] * b[1]) + (a[2] * b[2]);
}

static void cross(const double *a, const double *b, double *c)
{
	c[0] = (a[1] * b[2]) - (a[2] * b[1]);
	c[1] = (a[2] * b[0]) - (a[0] * b[2]);
	c[2] = (a[0] * b[1]) - (a[1] * b[0]);
}

static void axis(const seg_t *s, double *t)
{
	for (int c = 0; c < 3; c++) t[c] = (s->x2[c] - s->x1[c]) / s->len;
}

static void centre(const seg_t *s, double *m)
{
	for (int c = 0; c < 3; c++) m[c] = 0.5 * (s->x1[c] + s->x2[c]);
}

static double centre_gap(const seg_t *s1, const seg_t *s2)
{
	double a[3], b[3];
	centre(s1, a);
	centre(s2, b);
	const double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
	return sqrt((dx * dx) + (dy * dy) + (dz * dz));
}

static double size_lw(const seg_t *s1, const seg_t *s2)
{
	double m = (s1->len > s2->len) ? s1->len : s2->len;
	if (s1->wid > m) m = s1->wid;
	if (s2->wid > m) m = s2->wid;
	return m;
}

/* ---------------------------------------------------------------- thin wires (partial.c) */

/* antiderivative of the parallel-filament kernel (partial.c:49-52) */
static double fil(double x, double d)
{
	return (x * asinh(x / d)) - sqrt((x * x) + (d * d));
}

/* partial.c:55-58 */
double orc_neumann_self(double l, double a)
{
	return 2 * ((l * asinh(l / a)) - sqrt((l * l) + (a * a)) + a);
}

/* composite 8-point Gauss over nsub x nsub sub-blocks (partial.c:61-99).
 * retarded = 0: sum w / R;  1: sum w (cos kR - 1)/R  and  -sum w sin kR / R */
static d_complex_t wire_quad(const seg_t *s1, const seg_t *s2, double aeff, double kw, int nsub, int retarded)
{
	const double a2 = aeff * aeff;
	double re = 0, im = 0;
	for (int b1 = 0; b1 < nsub; b1++)
		for (int b2 = 0; b2 < nsub; b2++)
			for (int i = 0; i < 8; i++) {
				const double u = (b1 + (0.5 * (1 + X8[i]))) / nsub;
				double q[3];
				for (int c = 0; c < 3; c++) q[c] = s1->x1[c] + ((s1->x2[c] - s1->x1[c]) * u);
				for (int j = 0; j < 8; j++) {
					const double v = (b2 + (0.5 * (1 + X8[j]))) / nsub;
					const double dx = s2->x1[0] + ((s2->x2[0] - s2->x1[0]) * v) - q[0];
					const double dy = s2->x1[1] + ((s2->x2[1] - s2->x1[1]) * v) - q[1];
					const double dz = s2->x1[2] + ((s2->x2[2] - s2->x1[2]) * v) - q[2];
					const double r = sqrt((dx * dx) + (dy * dy) + (dz * dz) + a2);
					const double w = W8[i] * W8[j];
					if (!retarded) {
						re += w / r;
					}
					else {
						re += w * (cos(kw * r) - 1) / r;
						im -= w * sin(kw * r) / r;
					}
				}
			}
	const double jac = s1->len * s2->len * 0.25 / ((double)nsub * nsub);
	return d_complex(jac * re, jac * im);
}

static double tangent_cross(const seg_t *s1, const seg_t *s2)
{
	double t1[3], t2[3], c[3];
	axis(s1, t1);
	axis(s2, t2);
	cross(t1, t2, c);
	return sqrt((c[0] * c[0]) + (c[1] * c[1]) + (c[2] * c[2]));
}

/* static thin-wire pair (partial.c:102-163) */
double orc_neumann_pair(const seg_t *s1, const seg_t *s2, double a1, double a2)
{
	const double aeff = 0.5 * (a1 + a2);
	double t1[3];
	axis(s1, t1);
	const double lmax = (s1->len > s2->len) ? s1->len : s2->len;
	if (tangent_cross(s1, s2) < 1e-9) {
		double r1[3], r2[3], perp[3];
		for (int c = 0; c < 3; c++) {
			r1[c] = s2->x1[c] - s1->x1[c];
			r2[c] = s2->x2[c] - s1->x1[c];
		}
		double lo = dot(r1, t1), hi = dot(r2, t1);
		if (lo > hi) {
			const double t = lo;
			lo = hi;
			hi = t;
		}
		const double along = dot(r1, t1);
		for (int c = 0; c < 3; c++) perp[c] = r1[c] - (along * t1[c]);
		const double d = sqrt(dot(perp, perp) + (aeff * aeff));
		const double L = s1->len;
		return fil(L - lo, d) - fil(L - hi, d) - fil(0 - lo, d) + fil(0 - hi, d);
	}
	const double rc = centre_gap(s1, s2);
	if (rc > 20 * lmax) return s1->len * s2->len / rc;
	return wire_quad(s1, s2, aeff, 0, (rc > 2 * lmax) ? 1 : 4, 0).r;
}

/* ---------------------------------------------------------------- ribbons (surface.c) */

/* surface.c:48-60 */
static double rect_prim(double u, double v, double z)
{
	const double az = fabs(z);
	const double uz = (u * u) + (z * z);
	const double vz = (v * v) + (z * z);
	double t = 0;
	if (uz > 0) t += u * asinh(v / sqrt(uz));
	if (vz > 0) t += v * asinh(u / sqrt(vz));
	if (az > 0) t -= az * atan2(u * v, az * sqrt(uz + (v * v)));
	return t;
}

/* axis t, transverse w, normal n = t x w, centre m (surface.c:63-73) */
static void frame(const seg_t *s, double *t, double *w, double *n, double *m)
{
	axis(s, t);
	for (int c = 0; c < 3; c++) w[c] = s->wv[c];
	centre(s, m);
	cross(t, w, n);
}

/* surface.c:76-91 */
double orc_rect_potential(const seg_t *s, const double *pt)
{
	double t[3], w[3], n[3], m[3], d[3];
	frame(s, t, w, n, m);
	for (int c = 0; c < 3; c++) d[c] = pt[c] - m[c];
	const double u = dot(d, t), v = dot(d, w), z = dot(d, n);
	const double hl = 0.5 * s->len, hw = 0.5 * s->wid;
	return rect_prim(u + hl, v + hw, z) - rect_prim(u - hl, v + hw, z) - rect_prim(u + hl, v - hw, z) + rect_prim(u - hl, v - hw, z);
}

/* surface.c:102-131 */
double orc_ribbon_static(const seg_t *s1, const seg_t *s2, int nsub, int nq)
{
	const double *x = (nq >= 8) ? X8 : (nq >= 4) ? X4 : X2;
	const double *w = (nq >= 8) ? W8 : (nq >= 4) ? W4 : W2;
	const int n = (nq >= 8) ? 8 : (nq >= 4) ? 4 : 2;
	double t[3], v[3], nn[3], m[3];
	frame(s1, t, v, nn, m);
	double acc = 0;
	for (int ba = 0; ba < nsub; ba++)
		for (int bb = 0; bb < nsub; bb++)
			for (int i = 0; i < n; i++) {
				const double a = s1->len * (((ba + (0.5 * (1 + x[i]))) / nsub) - 0.5);
				for (int j = 0; j < n; j++) {
					const double b = s1->wid * (((bb + (0.5 * (1 + x[j]))) / nsub) - 0.5);
					double pt[3];
					for (int c = 0; c < 3; c++) pt[c] = m[c] + (a * t[c]) + (b * v[c]);
					acc += w[i] * w[j] * orc_rect_potential(s2, pt);
				}
			}
	return s1->len * acc / (4 * s2->wid * (double)nsub * nsub);
}

/* (e^{-jkR} - 1)/R term on a pair of point sets, R -> 0 limit -jk (surface.c:160-166) */
static void corr_term(double r, double w, double kw, double *re, double *im)
{
	if (r > 1e-300) {
		*re += w * (cos(kw * r) - 1) / r;
		*im -= w * sin(kw * r) / r;
	}
	else {
		*im -= w * kw;
	}
}

/* surface.c:134-178 */
d_complex_t orc_ribbon_corr(const seg_t *s1, const seg_t *s2, double kw, int nsub)
{
	double t1[3], w1[3], n1[3], m1[3], t2[3], w2[3], n2[3], m2[3];
	frame(s1, t1, w1, n1, m1);
	frame(s2, t2, w2, n2, m2);
	double re = 0, im = 0;
	for (int ba = 0; ba < nsub; ba++)
		for (int bb = 0; bb < nsub; bb++)
			for (int i = 0; i < 4; i++) {
				const double a1 = s1->len * (((ba + (0.5 * (1 + X4[i]))) / nsub) - 0.5);
				for (int j = 0; j < 4; j++) {
					const double b1 = s1->wid * ((0.5 * (1 + X4[j])) - 0.5);
					double p1[3];
					for (int c = 0; c < 3; c++) p1[c] = m1[c] + (a1 * t1[c]) + (b1 * w1[c]);
					for (int k = 0; k < 4; k++) {
						const double a2 = s2->len * (((bb + (0.5 * (1 + X4[k]))) / nsub) - 0.5);
						for (int l = 0; l < 4; l++) {
							const double b2 = s2->wid * ((0.5 * (1 + X4[l])) - 0.5);
							const double dx = m2[0] + (a2 * t2[0]) + (b2 * w2[0]) - p1[0];
							const double dy = m2[1] + (a2 * t2[1]) + (b2 * w2[1]) - p1[1];
							const double dz = m2[2] + (a2 * t2[2]) + (b2 * w2[2]) - p1[2];
							const double r = sqrt((dx * dx) + (dy * dy) + (dz * dz));
							corr_term(r, W4[i] * W4[j] * W4[k] * W4[l], kw, &re, &im);
						}
					}
				}
			}
	const double jac = s1->len * s2->len / (16 * (double)nsub * nsub);
	return d_complex(jac * re, jac * im);
}

/* ---------------------------------------------------------------- Hoer-Love (volume.c) */

/* volume.c:44-71 */
static double hoer_love(double x, double y, double z)
{
	const double x2 = x * x, y2 = y * y, z2 = z * z;
	const double R = sqrt(x2 + y2 + z2);
	const double yz = y2 + z2, xz = x2 + z2, xy = x2 + y2;
	double t = 0;
	if ((x != 0) && (yz > 0)) t += ((y2 * z2 / 4) - (y2 * y2 / 24) - (z2 * z2 / 24)) * x * asinh(x / sqrt(yz));
	if ((y != 0) && (xz > 0)) t += ((x2 * z2 / 4) - (x2 * x2 / 24) - (z2 * z2 / 24)) * y * asinh(y / sqrt(xz));
	if ((z != 0) && (xy > 0)) t += ((x2 * y2 / 4) - (x2 * x2 / 24) - (y2 * y2 / 24)) * z * asinh(z / sqrt(xy));
	t += ((x2 * x2) + (y2 * y2) + (z2 * z2) - 3 * ((x2 * y2) + (y2 * z2) + (z2 * x2))) * R / 60;
	if (z != 0) t -= (x * y * z * z2 / 6) * atan((x * y) / (z * R));
	if (y != 0) t -= (x * y * y2 * z / 6) * atan((x * z) / (y * R));
	if (x != 0) t -= (x * x2 * y * z / 6) * atan((y * z) / (x * R));
	return t;
}

/* volume.c:92-110 */
int orc_bar_use_hl(const seg_t *s1, const seg_t *s2)
{
	double t1[3], w1[3], n1[3], m1[3], t2[3], w2[3], n2[3], m2[3];
	frame(s1, t1, w1, n1, m1);
	frame(s2, t2, w2, n2, m2);
	if (fabs(dot(t1, t2)) < 1 - 1e-9) return 0;
	if (fabs(dot(w1, w2)) < 1 - 1e-9) return 0;
	return centre_gap(s1, s2) <= 8.0 * size_lw(s1, s2);
}

/* volume.c:114-155 */
double orc_bar_pair(const seg_t *s1, const seg_t *s2)
{
	double t1[3], w1[3], n1[3], m1[3], t2[3], w2[3], n2[3], m2[3], rc[3];
	frame(s1, t1, w1, n1, m1);
	frame(s2, t2, w2, n2, m2);
	const double a = s1->thick, b = s1->wid, c = s1->len;
	const double d = s2->thick, e = s2->wid, f = s2->len;
	for (int k = 0; k < 3; k++) rc[k] = m2[k] - m1[k];
	const double E = dot(rc, n1) + ((a - d) / 2);
	const double P = dot(rc, w1) + ((b - e) / 2);
	const double L3 = dot(rc, t1) + ((c - f) / 2);
	const double q[4] = {E - a, E + d - a, E + d, E};
	const double r[4] = {P - b, P + e - b, P + e, P};
	const double s[4] = {L3 - c, L3 + f - c, L3 + f, L3};
	double acc = 0;
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++)
			for (int k = 0; k < 4; k++) acc += ((((i + j + k) % 2) == 0) ? 1.0 : -1.0) * hoer_love(q[i], r[j], s[k]);
	return acc / (a * b * d * e);
}

/* ---------------------------------------------------------------- polygons (polygon.c) */

static void newell(const seg_t *s, double *nv)
{
	nv[0] = nv[1] = nv[2] = 0;
	for (int k = 0; k < s->npv; k++) {
		double c[3];
		cross(&s->pv[3 * k], &s->pv[3 * ((k + 1) % s->npv)], c);
		for (int i = 0; i < 3; i++) nv[i] += c[i];
	}
}

/* Wilton et al. 1984 edge sum (polygon.c:68-131) */
double orc_poly_potential(const seg_t *s, const double *pt)
{
	double nv[3];
	newell(s, nv);
	const double n2 = sqrt(dot(nv, nv));
	if (n2 <= 0) return 0;
	const double nh[3] = {nv[0] / n2, nv[1] / n2, nv[2] / n2};
	double d0[3], rho[3];
	for (int i = 0; i < 3; i++) d0[i] = pt[i] - s->pv[i];
	const double z = dot(d0, nh), az = fabs(z);
	for (int i = 0; i < 3; i++) rho[i] = pt[i] - (z * nh[i]);
	double acc = 0;
	for (int k = 0; k < s->npv; k++) {
		const double *a = &s->pv[3 * k];
		const double *b = &s->pv[3 * ((k + 1) % s->npv)];
		double e[3], sh[3], mh[3], ra[3], rb[3];
		for (int i = 0; i < 3; i++) e[i] = b[i] - a[i];
		const double el = sqrt(dot(e, e));
		if (el <= 0) continue;
		for (int i = 0; i < 3; i++) sh[i] = e[i] / el;
		cross(sh, nh, mh);
		for (int i = 0; i < 3; i++) {
			ra[i] = a[i] - rho[i];
			rb[i] = b[i] - rho[i];
		}
		const double lm = dot(ra, sh), lp = dot(rb, sh), t0 = dot(ra, mh);
		const double r02 = (t0 * t0) + (z * z);
		const double rm = sqrt((lm * lm) + r02), rp = sqrt((lp * lp) + r02);
		if (r02 <= 0) continue;
		const double sp = (lp >= 0) ? (rp + lp) : (r02 / (rp - lp));
		const double sm = (lm >= 0) ? (rm + lm) : (r02 / (rm - lm));
		acc += t0 * log(sp / sm);
		if (az > 0) acc -= az * (atan((t0 * lp) / (r02 + (az * rp))) - atan((t0 * lm) / (r02 + (az * rm))));
	}
	return fabs(acc);
}

#define QMAX (POLY_MAX * 4 * 7)

static double triangle_area(const double *a, const double *b, const double *c)
{
	double u[3], v[3], w[3];
	for (int i = 0; i < 3; i++) {
		u[i] = b[i] - a[i];
		v[i] = c[i] - a[i];
	}
	cross(u, v, w);
	return 0.5 * sqrt(dot(w, w));
}

/* quadrature points and weights of a cell (polygon.c:170-268) */
static int points(const seg_t *s, int nsub, int n7, double *px, double *wt)
{
	int np = 0;
	if (s->npv > 0) {
		const int nq = n7 ? 7 : 3;
		const int ns = (nsub >= 2) ? 4 : 1;
		for (int k = 0; k < s->npv; k++) {
			const double *a = s->papex;
			const double *b = &s->pv[3 * k];
			const double *c = &s->pv[3 * ((k + 1) % s->npv)];
			const double at = triangle_area(a, b, c);
			if (at <= 0) continue;
			for (int t = 0; t < ns; t++) {
				double v[3][3];
				for (int i = 0; i < 3; i++)
					for (int q = 0; q < 3; q++) {
						const double ba = (ns == 4) ? SPLIT[t][q][0] : (q == 0);
						const double bb = (ns == 4) ? SPLIT[t][q][1] : (q == 1);
						const double bc = (ns == 4) ? SPLIT[t][q][2] : (q == 2);
						v[q][i] = (ba * a[i]) + (bb * b[i]) + (bc * c[i]);
					}
				for (int q = 0; q < nq; q++) {
					const double *bw = n7 ? T7B[q] : T3B[q];
					for (int i = 0; i < 3; i++) px[(3 * np) + i] = (bw[0] * v[0][i]) + (bw[1] * v[1][i]) + (bw[2] * v[2][i]);
					wt[np++] = (n7 ? T7W[q] : T3W[q]) * (at / ns);
				}
			}
		}
		return np;
	}
	if (s->wid <= 0) {
		for (int b = 0; b < nsub; b++)
			for (int q = 0; q < 4; q++) {
				const double u = (b + (0.5 * (1 + X4[q]))) / nsub;
				for (int i = 0; i < 3; i++) px[(3 * np) + i] = s->x1[i] + ((s->x2[i] - s->x1[i]) * u);
				wt[np++] = 0.5 * W4[q] * s->len / nsub;
			}
		return np;
	}
	double t[3], m[3];
	axis(s, t);
	centre(s, m);
	for (int b = 0; b < nsub; b++)
		for (int q = 0; q < 4; q++) {
			const double a = s->len * (((b + (0.5 * (1 + X4[q]))) / nsub) - 0.5);
			for (int j = 0; j < 4; j++) {
				const double bt = s->wid * (0.5 * X4[j]);
				for (int i = 0; i < 3; i++) px[(3 * np) + i] = m[i] + (a * t[i]) + (bt * s->wv[i]);
				wt[np++] = 0.25 * W4[q] * W4[j] * s->len * s->wid / nsub;
			}
		}
	return np;
}

/* polygon.c:283-302 */
double orc_poly_static(const seg_t *s1, const seg_t *s2, int nsub, int n7)
{
	if ((s2->npv <= 0) && (s2->wid <= 0)) {
		if ((s1->npv <= 0) && (s1->wid <= 0)) return 0;
		return orc_poly_static(s2, s1, nsub, n7);
	}
	double *px = (double *)malloc(QMAX * 3 * sizeof(double));
	double *wt = (double *)malloc(QMAX * sizeof(double));
	const int np = points(s1, nsub, n7, px, wt);
	double acc = 0;
	for (int q = 0; q < np; q++)
		acc += wt[q] * ((s2->npv > 0) ? orc_poly_potential(s2, &px[3 * q]) : orc_rect_potential(s2, &px[3 * q]));
	free(px);
	free(wt);
	const double w1 = (s1->wid > 0) ? s1->wid : 1;
	const double w2 = (s2->wid > 0) ? s2->wid : 1;
	return acc / (w1 * w2);
}

/* polygon.c:305-334 */
d_complex_t orc_poly_corr(const seg_t *s1, const seg_t *s2, double kw)
{
	double *p1 = (double *)malloc(QMAX * 3 * sizeof(double)), *w1 = (double *)malloc(QMAX * sizeof(double));
	double *p2 = (double *)malloc(QMAX * 3 * sizeof(double)), *w2 = (double *)malloc(QMAX * sizeof(double));
	const int n1 = points(s1, 1, 0, p1, w1);
	const int n2 = points(s2, 1, 0, p2, w2);
	double re = 0, im = 0;
	for (int i = 0; i < n1; i++)
		for (int j = 0; j < n2; j++) {
			const double dx = p2[3 * j] - p1[3 * i];
			const double dy = p2[(3 * j) + 1] - p1[(3 * i) + 1];
			const double dz = p2[(3 * j) + 2] - p1[(3 * i) + 2];
			corr_term(sqrt((dx * dx) + (dy * dy) + (dz * dz)), w1[i] * w2[j], kw, &re, &im);
		}
	free(p1);
	free(w1);
	free(p2);
	free(w2);
	const double ww1 = (s1->wid > 0) ? s1->wid : 1;
	const double ww2 = (s2->wid > 0) ? s2->wid : 1;
	return d_complex(re / (ww1 * ww2), im / (ww1 * ww2));
}

/* ---------------------------------------------------------------- dispatch (partial.c) */

/* partial.c:256-317; `same` stands for the reference's pointer identity s1 == s2 */
static d_complex_t pair_k(const seg_t *s1, const seg_t *s2, double a1, double a2, double kw, int same)
{
	if ((s1->npv > 0) || (s2->npv > 0)) {
		const double rc = centre_gap(s1, s2);
		const double lm = size_lw(s1, s2);
		const double is = ((rc > 20 * lm) && !same) ? (s1->len * s2->len / rc)
		                : orc_poly_static(s1, s2, (rc < 2 * lm) ? 2 : 1, (rc <= 5.0 * lm) ? 1 : 0);
		if (kw <= 0) return d_complex(is, 0);
		return d_add(d_complex(is, 0), orc_poly_corr(s1, s2, kw));
	}
	const double rc = centre_gap(s1, s2);
	const double lm = size_lw(s1, s2);
	const double ratio = (lm > 0) ? (rc / lm) : 0;
	int rsub = ((ratio > 2) ? 1 : 2) + (int)(kw * lm);
	if (rsub > 4) rsub = 4;
	if (s1->vol && s2->vol && orc_bar_use_hl(s1, s2)) {
		const double is = orc_bar_pair(s1, s2);
		if (kw <= 0) return d_complex(is, 0);
		return d_add(d_complex(is, 0), orc_ribbon_corr(s1, s2, kw, rsub));
	}
	if ((s1->wid > 0) || (s2->wid > 0)) {
		const int nq = (ratio <= 2.5) ? 8 : (ratio <= 10) ? 4 : 2;
		const double is = orc_ribbon_static(s1, s2, rsub, nq);
		if (kw <= 0) return d_complex(is, 0);
		return d_add(d_complex(is, 0), orc_ribbon_corr(s1, s2, kw, rsub));
	}
	if (kw <= 0) return d_complex(orc_neumann_pair(s1, s2, a1, a2), 0);
	const double aeff = 0.5 * (a1 + a2);
	const double lmax = (s1->len > s2->len) ? s1->len : s2->len;
	int nsub = 2 + (int)(kw * lmax);
	if (nsub > 8) nsub = 8;
	const double is = (tangent_cross(s1, s2) < 1e-9) ? orc_neumann_pair(s1, s2, a1, a2) : wire_quad(s1, s2, aeff, 0, nsub, 0).r;
	return d_add(d_complex(is, 0), wire_quad(s1, s2, aeff, kw, nsub, 1));
}

d_complex_t orc_pair_k(const seg_t *s1, const seg_t *s2, double a1, double a2, double kw)
{
	return pair_k(s1, s2, a1, a2, kw, s1 == s2);
}

/* partial.c:320-328 */
d_complex_t orc_self_k(const seg_t *s, double a, double kw)
{
	if (s->wid > 0) return pair_k(s, s, a, a, kw, 1);
	const d_complex_t is = d_complex(orc_neumann_self(s->len, a), 0);
	if (kw <= 0) return is;
	int nsub = 2 + (int)(kw * s->len);
	if (nsub > 8) nsub = 8;
	return d_add(is, wire_quad(s, s, a, kw, nsub, 1));
}

/* partial.c:352-362 */
void orc_seg_mirror(const seg_t *s, double gpz, seg_t *out)
{
	*out = *s;
	out->x1[2] = (2 * gpz) - s->x1[2];
	out->x2[2] = (2 * gpz) - s->x2[2];
	out->wv[2] = -s->wv[2];
	out->papex[2] = (2 * gpz) - s->papex[2];
	for (int k = 0; k < s->npv; k++) out->pv[(3 * k) + 2] = (2 * gpz) - s->pv[(3 * k) + 2];
}

/* ---------------------------------------------------------------- matrices */

/* partial.c:371-424 */
d_complex_t orc_lp_entry(const peec_t *p, int i, int j, double kw)
{
	const double mu = MU0 / (4 * PI);
	if (i > j) {
		const int t = i;
		i = j;
		j = t;
	}
	const seg_t *a = &p->seg[i];
	const seg_t *b = &p->seg[j];
	double ta[3];
	axis(a, ta);
	seg_t img;
	if (i == j) {
		d_complex_t v = orc_self_k(a, a->aL, kw);
		if (p->gp) {
			orc_seg_mirror(a, p->gpz, &img);
			const double tm = (ta[0] * ta[0]) + (ta[1] * ta[1]) - (ta[2] * ta[2]);
			v = d_sub(v, d_rmul(tm, pair_k(a, &img, a->aL, a->aL, kw, 0)));
		}
		return d_rmul(mu, v);
	}
	double tb[3];
	axis(b, tb);
	const double td = dot(ta, tb);
	const double tm = (ta[0] * tb[0]) + (ta[1] * tb[1]) - (ta[2] * tb[2]);
	d_complex_t v = d_complex(0, 0);
	if (td != 0) v = d_rmul(td, pair_k(a, b, a->aL, b->aL, kw, 0));
	if (p->gp && (tm != 0)) {
		orc_seg_mirror(b, p->gpz, &img);
		v = d_sub(v, d_rmul(tm, pair_k(a, &img, a->aL, b->aL, kw, 0)));
	}
	return d_rmul(mu, v);
}

void orc_lp_fill(const peec_t *p, double kw, d_complex_t *lp)
{
	const int n = p->nseg;
	int i;
#pragma omp parallel for schedule(dynamic)
	for (i = 0; i < n; i++)
		for (int j = i; j < n; j++) lp[(size_t)i * n + j] = orc_lp_entry(p, i, j, kw);
	for (i = 0; i < n; i++)
		for (int j = 0; j < i; j++) lp[(size_t)i * n + j] = lp[(size_t)j * n + i];
}

/* charge-cell integral with the image subtracted (potential.c:217-237) */
static d_complex_t charge_pair(const peec_t *p, int h1, int h2, double kw)
{
	const seg_t *a = &p->chg[h1];
	const seg_t *b = &p->chg[h2];
	d_complex_t v = (h1 == h2) ? orc_self_k(a, a->aP, kw) : pair_k(a, b, a->aP, b->aP, kw, 0);
	if (p->gp) {
		seg_t img;
		orc_seg_mirror(b, p->gpz, &img);
		v = d_sub(v, pair_k(a, &img, a->aP, b->aP, kw, 0));
	}
	return v;
}

void orc_fill_rows(const peec_t *p, double kw, int r0, int r1, int pot)
{
	const int n = pot ? p->nchg : p->nseg;
	volatile double sink = 0;
	int i;
#pragma omp parallel for schedule(dynamic) reduction(+ : sink)
	for (i = r0; i < r1; i++)
		for (int j = i; j < n; j++) {
			const d_complex_t v = pot ? charge_pair(p, i, j, kw) : orc_lp_entry(p, i, j, kw);
			sink += v.r;
		}
	(void)sink;
}

/* cells = distinct nodes of the charge cells in first-appearance order (potential.c:29-62) */
static int cell_maps(const peec_t *p, int **cellof, double **carea)
{
	const int nh = p->nchg;
	int *of = (int *)malloc((size_t)nh * sizeof(int));
	int *id = (int *)malloc((size_t)nh * sizeof(int));
	int nc = 0;
	for (int h = 0; h < nh; h++) {
		int f = -1;
		for (int c = 0; c < nc; c++)
			if (id[c] == p->chgnode[h]) {
				f = c;
				break;
			}
		if (f < 0) {
			f = nc;
			id[nc++] = p->chgnode[h];
		}
		of[h] = f;
	}
	double *ar = (double *)calloc((size_t)(nc > 0 ? nc : 1), sizeof(double));
	for (int h = 0; h < nh; h++) ar[of[h]] += p->chg[h].len;
	free(id);
	*cellof = of;
	*carea = ar;
	return nc;
}

/* P = aggregated, scaled charge-cell integrals; C = P^-1 by LU + unit columns
 * (potential.c:238-281).  pmat (may be NULL) receives P before factorisation. */
int orc_pot_fill(const peec_t *p, double kw, d_complex_t *pmat, d_complex_t *c)
{
	const int nh = p->nchg;
	int *of;
	double *ar;
	const int nc = cell_maps(p, &of, &ar);
	d_complex_t *g = (d_complex_t *)malloc((size_t)nh * nh * sizeof(d_complex_t));
	d_complex_t *P = (d_complex_t *)calloc((size_t)nc * nc, sizeof(d_complex_t));
	int h1;
#pragma omp parallel for schedule(dynamic)
	for (h1 = 0; h1 < nh; h1++)
		for (int h2 = h1; h2 < nh; h2++) g[(size_t)h1 * nh + h2] = charge_pair(p, h1, h2, kw);
	for (h1 = 0; h1 < nh; h1++)
		for (int h2 = 0; h2 < h1; h2++) g[(size_t)h1 * nh + h2] = g[(size_t)h2 * nh + h1];
	for (h1 = 0; h1 < nh; h1++)
		for (int h2 = 0; h2 < nh; h2++) {
			const size_t id = (size_t)of[h1] * nc + of[h2];
			P[id] = d_add(P[id], g[(size_t)h1 * nh + h2]);
		}
	for (int i = 0; i < nc; i++)
		for (int j = 0; j < nc; j++)
			P[(size_t)i * nc + j] = d_rmul(1 / (4 * PI * EPS0 * ar[i] * ar[j]), P[(size_t)i * nc + j]);
	if (pmat) memcpy(pmat, P, (size_t)nc * nc * sizeof(d_complex_t));
	int *piv = (int *)malloc((size_t)nc * sizeof(int));
	int rc = 0;
	if (orc_lu_decomp(nc, P, piv) >= 0) {
		rc = 1;
	}
	else if (c != NULL) {
		d_complex_t *b = (d_complex_t *)malloc((size_t)nc * sizeof(d_complex_t));
		for (int j = 0; j < nc; j++) {
			memset(b, 0, (size_t)nc * sizeof(d_complex_t));
			b[j] = d_complex(1, 0);
			orc_lu_solve(nc, P, piv, b);
			for (int i = 0; i < nc; i++) c[(size_t)i * nc + j] = b[i];
		}
		free(b);
	}
	free(piv);
	free(g);
	free(P);
	free(of);
	free(ar);
	return rc;
}
