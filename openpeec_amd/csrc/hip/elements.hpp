// elements.hpp — FP64 partial-element evaluators on gfx950 (device code).
//
// One call evaluates one (cell, cell) integral exactly as the reference defines it: the same
// branch rules, integer truncations, quadrature tables and FP64 operation order, so a dense
// entry agrees with the reference to a few ulps of libm (ocml vs glibc) rather than to a
// quadrature tolerance.  Compiled with -ffp-contract=off: the reference's x86-64 build never
// fuses multiply-add, and fused roundings would move entries with cancellation (fpar sums,
// phi_rect combinations, Hoer-Love) by more than the 1e-12 parity budget.
//
// Cells arrive pre-packed (dcell_t, fill_pack.c) with their frames evaluated on the host by
// the reference's own expressions.  Reference map (OpenPEEC src/):
//   fpar / neumann_self / gauss_kernel / neumann_pair ......... partial.c:49-163
//   corr_nsub / is_parallel / ribbon_ratio / nsub / nq ......... partial.c:166-249
//   neumann_pair_k dispatcher / neumann_self_k ................. partial.c:256-328
//   phi_rect / rect_potential / ribbon_static / ribbon_corr .... surface.c:48-178
//   hl_F / bar_use_hl / bar_pair ............................... volume.c:44-155
//   poly_potential / poly_qpts / cell_qpts / poly_static / corr  polygon.c:68-334
#pragma once
#include <hip/hip_runtime.h>
#include "../peec_internal.h"
#include "fill_kernels.hpp"

namespace pe {

struct cplx {
	double r, i;
};

__device__ __forceinline__ cplx mk(double r, double i) { return cplx{r, i}; }
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return cplx{a.r + b.r, a.i + b.i}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return cplx{a.r - b.r, a.i - b.i}; }
__device__ __forceinline__ cplx cscale(double s, cplx z) { return cplx{s * z.r, s * z.i}; }

// Gauss-Legendre abscissae/weights on [-1, 1] (8, 4, 2 points) and the triangle rules
// (7-point degree-5 and 3-point degree-2, weights normalised to sum 1).  Standard published
// values, written with the same decimal digits the reference uses so that the sums match.
__constant__ const double GX8[8] = {-0.9602898564975363, -0.7966664774136267, -0.5255324099163290, -0.1834346424956498,
                                     0.1834346424956498,  0.5255324099163290,  0.7966664774136267,  0.9602898564975363};
__constant__ const double GW8[8] = {0.1012285362903763, 0.2223810344533745, 0.3137066458778873, 0.3626837833783620,
                                    0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763};
__constant__ const double GX4[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
__constant__ const double GW4[4] = {0.3478548451374538, 0.6521451548625461, 0.6521451548625461, 0.3478548451374538};
__constant__ const double GX2[2] = {-0.5773502691896257, 0.5773502691896257};
__constant__ const double GW2[2] = {1.0, 1.0};
__constant__ const double TW7[7] = {0.225, 0.1323941527885062, 0.1323941527885062, 0.1323941527885062,
                                    0.1259391805448271, 0.1259391805448271, 0.1259391805448271};
__constant__ const double TB7[7][3] = {{1.0 / 3, 1.0 / 3, 1.0 / 3},
                                       {0.0597158717897698, 0.4701420641051151, 0.4701420641051151},
                                       {0.4701420641051151, 0.0597158717897698, 0.4701420641051151},
                                       {0.4701420641051151, 0.4701420641051151, 0.0597158717897698},
                                       {0.7974269853530873, 0.1012865073234563, 0.1012865073234563},
                                       {0.1012865073234563, 0.7974269853530873, 0.1012865073234563},
                                       {0.1012865073234563, 0.1012865073234563, 0.7974269853530873}};
__constant__ const double TW3[3] = {1.0 / 3, 1.0 / 3, 1.0 / 3};
__constant__ const double TB3[3][3] = {{2.0 / 3, 1.0 / 6, 1.0 / 6}, {1.0 / 6, 2.0 / 3, 1.0 / 6}, {1.0 / 6, 1.0 / 6, 2.0 / 3}};
// barycentric corners of the 4 similar sub-triangles of a 1:4 split
__constant__ const double SUB4[4][3][3] = {{{1, 0, 0}, {0.5, 0.5, 0}, {0.5, 0, 0.5}},
                                           {{0.5, 0.5, 0}, {0, 1, 0}, {0, 0.5, 0.5}},
                                           {{0.5, 0, 0.5}, {0, 0.5, 0.5}, {0, 0, 1}},
                                           {{0.5, 0.5, 0}, {0, 0.5, 0.5}, {0.5, 0, 0.5}}};

__device__ __forceinline__ double dot3(const double *a, const double *b)
{
	return (a[0] * b[0]) + (a[1] * b[1]) + (a[2] * b[2]);
}

__device__ __forceinline__ double lmax_lw(const dcell_t &s1, const dcell_t &s2)
{
	double lm = (s1.len > s2.len) ? s1.len : s2.len;
	if (s1.wid > lm) lm = s1.wid;
	if (s2.wid > lm) lm = s2.wid;
	return lm;
}

__device__ __forceinline__ double centre_dist(const dcell_t &s1, const dcell_t &s2)
{
	const double dx = s2.c[0] - s1.c[0];
	const double dy = s2.c[1] - s1.c[1];
	const double dz = s2.c[2] - s1.c[2];
	return sqrt((dx * dx) + (dy * dy) + (dz * dz));
}

// ---------------------------------------------------------------- thin wire (partial.c) ----

__device__ __forceinline__ double fpar(double x, double d)
{
	return (x * asinh(x / d)) - sqrt((x * x) + (d * d));
}

__device__ __forceinline__ double wire_self_static(double l, double a)
{
	return 2 * ((l * asinh(l / a)) - sqrt((l * l) + (a * a)) + a);
}

// nsub^2 x 8x8 composite Gauss; mode 0: sum w/R, mode 1: sum w (cos kR - 1)/R - j sum w sin kR/R
template <int MODE>
__device__ cplx wire_gauss(const dcell_t &s1, const dcell_t &s2, double aeff, double kw, int nsub)
{
	const double a2 = aeff * aeff;
	double sumr = 0, sumi = 0;
	for (int i1 = 0; i1 < nsub; i1++) {
		for (int i2 = 0; i2 < nsub; i2++) {
			for (int i = 0; i < 8; i++) {
				const double u = (i1 + (0.5 * (1 + GX8[i]))) / nsub;
				const double q0 = s1.x1[0] + (s1.d[0] * u);
				const double q1 = s1.x1[1] + (s1.d[1] * u);
				const double q2 = s1.x1[2] + (s1.d[2] * u);
				for (int j = 0; j < 8; j++) {
					const double v = (i2 + (0.5 * (1 + GX8[j]))) / nsub;
					const double dx = s2.x1[0] + (s2.d[0] * v) - q0;
					const double dy = s2.x1[1] + (s2.d[1] * v) - q1;
					const double dz = s2.x1[2] + (s2.d[2] * v) - q2;
					const double r = sqrt((dx * dx) + (dy * dy) + (dz * dz) + a2);
					const double w = GW8[i] * GW8[j];
					if (MODE == 0) {
						sumr += w / r;
					}
					else {
						double sn, cs;
						sincos(kw * r, &sn, &cs);
						sumr += w * (cs - 1) / r;
						sumi -= w * sn / r;
					}
				}
			}
		}
	}
	const double fac = s1.len * s2.len * 0.25 / ((double)nsub * nsub);
	return mk(fac * sumr, fac * sumi);
}

__device__ __forceinline__ double cross_norm(const double *t1, const double *t2)
{
	const double cx = (t1[1] * t2[2]) - (t1[2] * t2[1]);
	const double cy = (t1[2] * t2[0]) - (t1[0] * t2[2]);
	const double cz = (t1[0] * t2[1]) - (t1[1] * t2[0]);
	return sqrt((cx * cx) + (cy * cy) + (cz * cz));
}

// static thin-wire pair: parallel closed form, far midpoint, or 8-pt Gauss (nsub 1 or 4)
__device__ double wire_pair_static(const dcell_t &s1, const dcell_t &s2, double a1, double a2)
{
	const double aeff = 0.5 * (a1 + a2);
	const double lmax = (s1.len > s2.len) ? s1.len : s2.len;
	if (cross_norm(s1.t, s2.t) < 1e-9) {
		const double p1 = 0;
		const double p2 = s1.len;
		double r1[3], r2[3];
		for (int c = 0; c < 3; c++) {
			r1[c] = s2.x1[c] - s1.x1[c];
			r2[c] = s2.x2[c] - s1.x1[c];
		}
		double b1 = dot3(r1, s1.t);
		double b2 = dot3(r2, s1.t);
		if (b1 > b2) {
			const double t = b1;
			b1 = b2;
			b2 = t;
		}
		const double pr = dot3(r1, s1.t);
		double perp[3];
		for (int c = 0; c < 3; c++) perp[c] = r1[c] - (pr * s1.t[c]);
		const double d = sqrt(dot3(perp, perp) + (aeff * aeff));
		return fpar(p2 - b1, d) - fpar(p2 - b2, d) - fpar(p1 - b1, d) + fpar(p1 - b2, d);
	}
	const double rc = centre_dist(s1, s2);
	if (rc > 20 * lmax) return s1.len * s2.len / rc;
	return wire_gauss<0>(s1, s2, aeff, 0, (rc > 2 * lmax) ? 1 : 4).r;
}

// ---------------------------------------------------------------- ribbons (surface.c) ------

__device__ __forceinline__ double phi_rect(double u, double v, double z)
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

__device__ __forceinline__ double rect_potential(const dcell_t &s, const double *pt)
{
	double d[3];
	for (int i = 0; i < 3; i++) d[i] = pt[i] - s.c[i];
	const double u = dot3(d, s.t);
	const double v = dot3(d, s.wv);
	const double z = dot3(d, s.nv);
	const double hl = 0.5 * s.len;
	const double hw = 0.5 * s.wid;
	return phi_rect(u + hl, v + hw, z) - phi_rect(u - hl, v + hw, z) - phi_rect(u + hl, v - hw, z) + phi_rect(u - hl, v - hw, z);
}

__device__ __forceinline__ double ribbon_ratio(const dcell_t &s1, const dcell_t &s2)
{
	const double rc = centre_dist(s1, s2);
	const double lm = lmax_lw(s1, s2);
	return (lm > 0) ? (rc / lm) : 0;
}

__device__ __forceinline__ int ribbon_nsub(const dcell_t &s1, const dcell_t &s2, double kw)
{
	const double lm = lmax_lw(s1, s2);
	int nsub = (ribbon_ratio(s1, s2) > 2) ? 1 : 2;
	nsub += (int)(kw * lm);
	if (nsub > 4) nsub = 4;
	return nsub;
}

__device__ __forceinline__ int ribbon_nq(const dcell_t &s1, const dcell_t &s2)
{
	const double r = ribbon_ratio(s1, s2);
	return (r <= 2.5) ? 8 : (r <= 10) ? 4 : 2;
}

template <int N>
__device__ double ribbon_static_n(const dcell_t &s1, const dcell_t &s2, int nsub)
{
	const double *xg = (N == 8) ? GX8 : (N == 4) ? GX4 : GX2;
	const double *wg = (N == 8) ? GW8 : (N == 4) ? GW4 : GW2;
	double sum = 0;
	for (int ia = 0; ia < nsub; ia++) {
		for (int ib = 0; ib < nsub; ib++) {
			for (int i = 0; i < N; i++) {
				const double a = s1.len * (((ia + (0.5 * (1 + xg[i]))) / nsub) - 0.5);
				for (int j = 0; j < N; j++) {
					const double b = s1.wid * (((ib + (0.5 * (1 + xg[j]))) / nsub) - 0.5);
					double pt[3];
					for (int k = 0; k < 3; k++) pt[k] = s1.c[k] + (a * s1.t[k]) + (b * s1.wv[k]);
					sum += wg[i] * wg[j] * rect_potential(s2, pt);
				}
			}
		}
	}
	return s1.len * sum / (4 * s2.wid * (double)nsub * nsub);
}

__device__ __forceinline__ double ribbon_static(const dcell_t &s1, const dcell_t &s2, int nsub, int nq)
{
	if (nq >= 8) return ribbon_static_n<8>(s1, s2, nsub);
	if (nq >= 4) return ribbon_static_n<4>(s1, s2, nsub);
	return ribbon_static_n<2>(s1, s2, nsub);
}

// (1/(w1 w2)) int int (exp(-jkR) - 1)/R, 4-D 4-point Gauss with nsub^2 axial sub-blocks
__device__ cplx ribbon_corr(const dcell_t &s1, const dcell_t &s2, double kw, int nsub)
{
	double sr = 0, si = 0;
	for (int ia = 0; ia < nsub; ia++) {
		for (int ja = 0; ja < nsub; ja++) {
			for (int i = 0; i < 4; i++) {
				const double a1 = s1.len * (((ia + (0.5 * (1 + GX4[i]))) / nsub) - 0.5);
				for (int j = 0; j < 4; j++) {
					const double b1 = s1.wid * ((0.5 * (1 + GX4[j])) - 0.5);
					double pt[3];
					for (int k = 0; k < 3; k++) pt[k] = s1.c[k] + (a1 * s1.t[k]) + (b1 * s1.wv[k]);
					const double wij = GW4[i] * GW4[j];
					for (int m = 0; m < 4; m++) {
						const double a2 = s2.len * (((ja + (0.5 * (1 + GX4[m]))) / nsub) - 0.5);
						const double wijm = wij * GW4[m];
						const double e0 = s2.c[0] + (a2 * s2.t[0]);
						const double e1 = s2.c[1] + (a2 * s2.t[1]);
						const double e2 = s2.c[2] + (a2 * s2.t[2]);
						for (int q = 0; q < 4; q++) {
							const double b2 = s2.wid * ((0.5 * (1 + GX4[q])) - 0.5);
							const double dx = e0 + (b2 * s2.wv[0]) - pt[0];
							const double dy = e1 + (b2 * s2.wv[1]) - pt[1];
							const double dz = e2 + (b2 * s2.wv[2]) - pt[2];
							const double r = sqrt((dx * dx) + (dy * dy) + (dz * dz));
							const double w = wijm * GW4[q];
							if (r > 1e-300) {
								double sn, cs;
								sincos(kw * r, &sn, &cs);
								sr += w * (cs - 1) / r;
								si -= w * sn / r;
							}
							else {
								si -= w * kw;
							}
						}
					}
				}
			}
		}
	}
	const double fac = s1.len * s2.len / (16 * (double)nsub * nsub);
	return mk(fac * sr, fac * si);
}

// ---------------------------------------------------------------- Hoer-Love bars (volume.c) -

__device__ double hl_F(double x, double y, double z)
{
	const double x2 = x * x;
	const double y2 = y * y;
	const double z2 = z * z;
	const double R = sqrt(x2 + y2 + z2);
	const double yz = y2 + z2;
	const double xz = x2 + z2;
	const double xy = x2 + y2;
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

__device__ __forceinline__ int bar_use_hl(const dcell_t &s1, const dcell_t &s2)
{
	if (fabs(dot3(s1.t, s2.t)) < 1 - 1e-9) return 0;
	if (fabs(dot3(s1.wv, s2.wv)) < 1 - 1e-9) return 0;
	return centre_dist(s1, s2) <= 8.0 * lmax_lw(s1, s2);
}

__device__ double bar_pair(const dcell_t &s1, const dcell_t &s2)
{
	const double a = s1.thick, b = s1.wid, c = s1.len;
	const double d = s2.thick, e = s2.wid, f = s2.len;
	double rc[3];
	for (int i = 0; i < 3; i++) rc[i] = s2.c[i] - s1.c[i];
	const double E = dot3(rc, s1.nv) + ((a - d) / 2);
	const double P = dot3(rc, s1.wv) + ((b - e) / 2);
	const double l3 = dot3(rc, s1.t) + ((c - f) / 2);
	const double q[4] = {E - a, E + d - a, E + d, E};
	const double r[4] = {P - b, P + e - b, P + e, P};
	const double s[4] = {l3 - c, l3 + f - c, l3 + f, l3};
	double sum = 0;
	for (int i = 0; i < 4; i++)
		for (int j = 0; j < 4; j++)
			for (int k = 0; k < 4; k++) {
				const double sgn = (((i + j + k) % 2) == 0) ? 1.0 : -1.0;
				sum += sgn * hl_F(q[i], r[j], s[k]);
			}
	return sum / (a * b * d * e);
}

// ---------------------------------------------------------------- polygons (polygon.c) -----

// Wilton et al. edge sum for a uniform planar polygon seen from pt
__device__ double poly_potential(const dcell_t &s, const polys &pp, int cell, const double *pt)
{
	const dpnorm_t pn = pp.pn[cell];
	if (pn.n2 <= 0) return 0;
	const dedge_t *ed = pp.edge + s.poff;
	double d0[3];
	for (int i = 0; i < 3; i++) d0[i] = pt[i] - ed[0].a[i];
	const double z = dot3(d0, pn.nh);
	const double az = fabs(z);
	double rho[3];
	for (int i = 0; i < 3; i++) rho[i] = pt[i] - (z * pn.nh[i]);
	double sum = 0;
	for (int k = 0; k < s.npv; k++) {
		const dedge_t &e = ed[k];
		if (e.el <= 0) continue;
		double ra[3], rb[3];
		for (int i = 0; i < 3; i++) {
			ra[i] = e.a[i] - rho[i];
			rb[i] = e.b[i] - rho[i];
		}
		const double lm = dot3(ra, e.sh);
		const double lp = dot3(rb, e.sh);
		const double t0 = dot3(ra, e.mh);
		const double r02 = (t0 * t0) + (z * z);
		if (r02 <= 0) continue;
		const double rm = sqrt((lm * lm) + r02);
		const double rp = sqrt((lp * lp) + r02);
		const double sp = (lp >= 0) ? (rp + lp) : (r02 / (rp - lp));
		const double sm = (lm >= 0) ? (rm + lm) : (r02 / (rm - lm));
		sum += t0 * log(sp / sm);
		if (az > 0) sum -= az * (atan((t0 * lp) / (r02 + (az * rp))) - atan((t0 * lm) / (r02 + (az * rm))));
	}
	return fabs(sum);
}

__device__ __forceinline__ double cell_potential(const dcell_t &s, const polys &pp, int cell, const double *pt)
{
	return (s.npv > 0) ? poly_potential(s, pp, cell, pt) : rect_potential(s, pt);
}

// Visit the quadrature points of a cell in the reference's order (cell_qpts / poly_qpts):
// polygon fan x (1:4 split when nsub >= 2) x 7- or 3-point rule; wire 4-pt Gauss x nsub;
// ribbon 4 x nsub axial x 4 transverse.  f(pt, weight) is called once per point.
template <typename F>
__device__ void for_each_qpt(const dcell_t &s, const polys &pp, int nsub, int n7, F &&f)
{
	if (s.npv > 0) {
		const int nq = n7 ? 7 : 3;
		const int ns = (nsub >= 2) ? 4 : 1;
		const dedge_t *ed = pp.edge + s.poff;
		for (int k = 0; k < s.npv; k++) {
			const double *a = s.papex;
			const double *b = ed[k].a;
			const double *c = ed[k].b;
			const double at = ed[k].at;
			if (at <= 0) continue;
			for (int t = 0; t < ns; t++) {
				double v0[3], v1[3], v2[3];
				for (int i = 0; i < 3; i++) {
					if (ns == 4) {
						v0[i] = (SUB4[t][0][0] * a[i]) + (SUB4[t][0][1] * b[i]) + (SUB4[t][0][2] * c[i]);
						v1[i] = (SUB4[t][1][0] * a[i]) + (SUB4[t][1][1] * b[i]) + (SUB4[t][1][2] * c[i]);
						v2[i] = (SUB4[t][2][0] * a[i]) + (SUB4[t][2][1] * b[i]) + (SUB4[t][2][2] * c[i]);
					}
					else {
						v0[i] = (1.0 * a[i]) + (0.0 * b[i]) + (0.0 * c[i]);
						v1[i] = (0.0 * a[i]) + (1.0 * b[i]) + (0.0 * c[i]);
						v2[i] = (0.0 * a[i]) + (0.0 * b[i]) + (1.0 * c[i]);
					}
				}
				const double sat = at / ns;
				for (int q = 0; q < nq; q++) {
					const double b0 = n7 ? TB7[q][0] : TB3[q][0];
					const double b1 = n7 ? TB7[q][1] : TB3[q][1];
					const double b2 = n7 ? TB7[q][2] : TB3[q][2];
					double px[3];
					for (int i = 0; i < 3; i++) px[i] = (b0 * v0[i]) + (b1 * v1[i]) + (b2 * v2[i]);
					f(px, (n7 ? TW7[q] : TW3[q]) * sat);
				}
			}
		}
		return;
	}
	if (s.wid <= 0) {
		for (int is = 0; is < nsub; is++) {
			for (int q = 0; q < 4; q++) {
				const double u = (is + (0.5 * (1 + GX4[q]))) / nsub;
				double px[3];
				for (int i = 0; i < 3; i++) px[i] = s.x1[i] + (s.d[i] * u);
				f(px, 0.5 * GW4[q] * s.len / nsub);
			}
		}
		return;
	}
	for (int is = 0; is < nsub; is++) {
		for (int q = 0; q < 4; q++) {
			const double a = s.len * (((is + (0.5 * (1 + GX4[q]))) / nsub) - 0.5);
			for (int j = 0; j < 4; j++) {
				const double b = s.wid * (0.5 * GX4[j]);
				double px[3];
				for (int i = 0; i < 3; i++) px[i] = s.c[i] + (a * s.t[i]) + (b * s.wv[i]);
				f(px, 0.25 * GW4[q] * GW4[j] * s.len * s.wid / nsub);
			}
		}
	}
}

// outer quadrature on s1 x analytic inner potential of s2 (a wire s2 swaps sides)
__device__ double poly_static(const dcell_t &s1, const polys &p1, int c1, const dcell_t &s2, const polys &p2, int c2,
                              int nsub, int n7)
{
	const bool swap = (s2.npv <= 0) && (s2.wid <= 0);
	if (swap && (s1.npv <= 0) && (s1.wid <= 0)) return 0;
	const dcell_t &o = swap ? s2 : s1;   // outer (quadrature) cell
	const dcell_t &in = swap ? s1 : s2;  // inner (analytic) cell
	const polys &po = swap ? p2 : p1;
	const polys &pi = swap ? p1 : p2;
	const int ci = swap ? c1 : c2;
	double sum = 0;
	for_each_qpt(o, po, nsub, n7, [&](const double *px, double w) { sum += w * cell_potential(in, pi, ci, px); });
	const double w1 = (o.wid > 0) ? o.wid : 1;
	const double w2 = (in.wid > 0) ? in.wid : 1;
	return sum / (w1 * w2);
}

__device__ cplx poly_corr(const dcell_t &s1, const polys &p1, const dcell_t &s2, const polys &p2, double kw)
{
	double sr = 0, si = 0;
	for_each_qpt(s1, p1, 1, 0, [&](const double *x1, double wt1) {
		for_each_qpt(s2, p2, 1, 0, [&](const double *x2, double wt2) {
			const double dx = x2[0] - x1[0];
			const double dy = x2[1] - x1[1];
			const double dz = x2[2] - x1[2];
			const double r = sqrt((dx * dx) + (dy * dy) + (dz * dz));
			const double w = wt1 * wt2;
			if (r > 1e-300) {
				double sn, cs;
				sincos(kw * r, &sn, &cs);
				sr += w * (cs - 1) / r;
				si -= w * sn / r;
			}
			else {
				si -= w * kw;
			}
		});
	});
	const double w1 = (s1.wid > 0) ? s1.wid : 1;
	const double w2 = (s2.wid > 0) ? s2.wid : 1;
	return mk(sr / (w1 * w2), si / (w1 * w2));
}

// ---------------------------------------------------------------- dispatcher --------------

// Geometry classes a fill launch may specialise on (bit mask of CELL_* kinds present).
enum { KM_WIRE = 1 << CELL_WIRE, KM_RIBBON = 1 << CELL_RIBBON, KM_BAR = 1 << CELL_BAR, KM_POLY = 1 << CELL_POLY };

// neumann_pair_k: polygon -> Hoer-Love -> ribbon -> wire.  `same` is true only when s1 and s2
// are the same cell object (a self term, never an image), which the reference tests by pointer.
template <int KM>
__device__ cplx pair_k(const dcell_t &s1, const polys &p1, int c1, const dcell_t &s2, const polys &p2, int c2,
                       double a1, double a2, double kw, bool same)
{
	if ((KM & KM_POLY) && ((s1.npv > 0) || (s2.npv > 0))) {
		const double rc = centre_dist(s1, s2);
		const double lm = lmax_lw(s1, s2);
		double is;
		if ((rc > 20 * lm) && !same) is = s1.len * s2.len / rc;
		else is = poly_static(s1, p1, c1, s2, p2, c2, (rc < 2 * lm) ? 2 : 1, (rc <= 5.0 * lm) ? 1 : 0);
		if (kw <= 0) return mk(is, 0);
		return cadd(mk(is, 0), poly_corr(s1, p1, s2, p2, kw));
	}
	if ((KM & KM_BAR) && s1.vol && s2.vol && bar_use_hl(s1, s2)) {
		const double is = bar_pair(s1, s2);
		if (kw <= 0) return mk(is, 0);
		return cadd(mk(is, 0), ribbon_corr(s1, s2, kw, ribbon_nsub(s1, s2, kw)));
	}
	if ((KM & (KM_RIBBON | KM_BAR)) && ((s1.wid > 0) || (s2.wid > 0))) {
		const int nsub = ribbon_nsub(s1, s2, kw);
		const double is = ribbon_static(s1, s2, nsub, ribbon_nq(s1, s2));
		if (kw <= 0) return mk(is, 0);
		return cadd(mk(is, 0), ribbon_corr(s1, s2, kw, nsub));
	}
	if (kw <= 0) return mk(wire_pair_static(s1, s2, a1, a2), 0);
	const double aeff = 0.5 * (a1 + a2);
	const double lmax = (s1.len > s2.len) ? s1.len : s2.len;
	int nsub = 2 + (int)(kw * lmax);
	if (nsub > 8) nsub = 8;
	const double is = (cross_norm(s1.t, s2.t) < 1e-9) ? wire_pair_static(s1, s2, a1, a2)
	                                                  : wire_gauss<0>(s1, s2, aeff, 0, nsub).r;
	return cadd(mk(is, 0), wire_gauss<1>(s1, s2, aeff, kw, nsub));
}

// neumann_self_k: surfaces/volumes/polygons go through the pair path with s1 == s2
template <int KM>
__device__ cplx self_k(const dcell_t &s, const polys &pp, int c, double a, double kw)
{
	if (s.wid > 0) return pair_k<KM>(s, pp, c, s, pp, c, a, a, kw, true);
	const cplx is = mk(wire_self_static(s.len, a), 0);
	if (kw <= 0) return is;
	int nsub = 2 + (int)(kw * s.len);
	if (nsub > 8) nsub = 8;
	return cadd(is, wire_gauss<1>(s, s, a, kw, nsub));
}

}  // namespace pe
