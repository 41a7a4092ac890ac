// kfill.hip — k-factored fill of Lp and the charge-cell integrals for frequency sweeps.
//
// Every retarded entry is  mult * [ I_static + fac * sum_p w_p (exp(-jk R_p) - 1) / R_p ]  over
// the reference's quadrature points p (ribbon_corr surface.c:134-178, gauss_kernel mode 1
// partial.c:61-99, poly_corr polygon.c:307-334).  The points, weights and the static part do not
// depend on k as long as the quadrature orders (nsub, which grow with floor(k lmax)) stay the
// same, so for a sweep over [k_lo, k_hi] each pair is reduced once to k-independent numbers:
//
//   shifted (pair well separated, R_p > 0):   exp(-jkR) = exp(-jkRc) sum_m (-jk (R - Rc))^m / m!
//       N_m = sum_p w_p (R_p - Rc)^m / R_p,  C(k) = exp(-jkRc) sum_m (-jk)^m / m! N_m - N_0
//   unshifted (near and self pairs):          (exp(-jkR) - 1)/R = sum_{m>=1} (-jk)^m R^(m-1) / m!
//       N_m = sum_p w_p R_p^(m-1),           C(k) = sum_{m>=1} (-jk)^m / m! N_m
//   (the reference's R -> 0 limit -jk w is the m = 1 term of the unshifted form).
//
// Rc is the centre distance (wires: sqrt(rc^2 + a^2)); |R_p - Rc| <= rho1 + rho2, the cells'
// bounding radii about their centres.  A pair takes the shifted form when k_hi (rho1 + rho2)
// <= 0.8 and the cells' bounding spheres are apart, else the unshifted form when k_hi R_max <=
// 0.8; with moments to m = 14 every term of the series is at most 1 in size and the truncated
// tail is below 0.8^15 e^0.8 / 15! ~ 6e-14 of the pair's static magnitude sum_p w_p / R_p.  A pair that fits neither, or whose quadrature order changes inside [k_lo, k_hi], is
// listed and evaluated directly at every frequency by the element code.
//
// One record per term (direct pair, ground-plane image): base = mult (I_static - fac N_0),
// Rc, N'_m = mult fac N_m; the entry at k is  sum_terms base + exp(-jkRc) sum_m (-jk)^m/m! N'_m.
// Records are stored structure-of-arrays by slot (slot = tile * 256 + lane of the 16 x 16 tile
// grid of the fill), so the per-frequency evaluation streams (2 + 15) doubles per pair per term,
// coalesced, and spends one sincos per pair instead of 256-1024.
#include <hip/hip_runtime.h>
#include "elements.hpp"
#include "hip_check.hpp"
#include "fill_kernels.hpp"
#include "kf_eval.hpp"

namespace pe {

constexpr double KF_X0 = 0.8;

__device__ __forceinline__ void tri_tile_kf(long L, int nt, int &ti, int &tj)
{
	const double b = 2.0 * nt + 1.0;
	int r = (int)floor((b - sqrt(b * b - 8.0 * (double)L)) * 0.5);
	if (r < 0) r = 0;
	if (r > nt - 1) r = nt - 1;
	auto start = [nt](long rr) { return rr * (long)nt - (rr * (rr - 1)) / 2; };
	while ((r > 0) && (start(r) > L)) r--;
	while ((r < nt - 1) && (start(r + 1) <= L)) r++;
	ti = r;
	tj = r + (int)(L - start(r));
}

struct macc {
	double N[KF_M + 1];
	double Rc;
	bool shifted;

	// pick the expansion for a pair at centre distance rc with bounding radii summing to del
	__device__ bool choose(double rc, double del, double khi, bool wire, double aeff)
	{
#pragma unroll
		for (int m = 0; m <= KF_M; m++) N[m] = 0;
		if ((khi * del <= KF_X0) && (wire || (rc - del > 0.25 * del))) {
			shifted = true;
			Rc = wire ? sqrt((rc * rc) + (aeff * aeff)) : rc;
			return true;
		}
		if (khi * (rc + del + aeff) <= KF_X0) {
			shifted = false;
			Rc = 0;
			return true;
		}
		return false;
	}

	__device__ __forceinline__ void add(double w, double r)
	{
		if (shifted) {
			const double d = r - Rc;
			double t = w / r;
#pragma unroll
			for (int m = 0; m <= KF_M; m++) {
				N[m] += t;
				t *= d;
			}
		}
		else {
			double t = w;
#pragma unroll
			for (int m = 1; m <= KF_M; m++) {
				N[m] += t;
				t *= r;
			}
		}
	}
};

// quadrature points of ribbon_corr (4-D 4-point Gauss, nsub^2 axial sub-blocks); returns fac
__device__ double ribbon_corr_pts(const dcell_t &s1, const dcell_t &s2, int nsub, macc &acc)
{
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
							acc.add(wijm * GW4[q], sqrt((dx * dx) + (dy * dy) + (dz * dz)));
						}
					}
				}
			}
		}
	}
	return s1.len * s2.len / (16 * (double)nsub * nsub);
}

// quadrature points of gauss_kernel (nsub^2 x 8 x 8, reduced-kernel distance); returns fac
__device__ double wire_corr_pts(const dcell_t &s1, const dcell_t &s2, double aeff, int nsub, macc &acc)
{
	const double a2 = aeff * aeff;
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
					acc.add(GW8[i] * GW8[j], sqrt((dx * dx) + (dy * dy) + (dz * dz) + a2));
				}
			}
		}
	}
	return s1.len * s2.len * 0.25 / ((double)nsub * nsub);
}

// quadrature points of poly_corr (3-point rules / wire and ribbon points); returns fac
__device__ double poly_corr_pts(const dcell_t &s1, const polys &p1, const dcell_t &s2, const polys &p2, macc &acc)
{
	for_each_qpt(s1, p1, 1, 0, [&](const double *x1, double wt1) {
		for_each_qpt(s2, p2, 1, 0, [&](const double *x2, double wt2) {
			const double dx = x2[0] - x1[0];
			const double dy = x2[1] - x1[1];
			const double dz = x2[2] - x1[2];
			acc.add(wt1 * wt2, sqrt((dx * dx) + (dy * dy) + (dz * dz)));
		});
	});
	const double w1 = (s1.wid > 0) ? s1.wid : 1;
	const double w2 = (s2.wid > 0) ? s2.wid : 1;
	return 1 / (w1 * w2);
}

// neumann_pair_k split into its static part and the moments of its correction; false when the
// pair must be evaluated directly (order change inside the band, or no convergent expansion)
template <int KM>
__device__ bool pair_split(const dcell_t &s1, const polys &p1, int c1, const dcell_t &s2, const polys &p2, int c2,
                           double a1, double a2, double klo, double khi, bool same, double del, double &stat,
                           double &fac, macc &acc)
{
	const double rc = centre_dist(s1, s2);
	if ((KM & KM_POLY) && ((s1.npv > 0) || (s2.npv > 0))) {
		const double lm = lmax_lw(s1, s2);
		if ((rc > 20 * lm) && !same) stat = s1.len * s2.len / rc;
		else stat = poly_static(s1, p1, c1, s2, p2, c2, (rc < 2 * lm) ? 2 : 1, (rc <= 5.0 * lm) ? 1 : 0);
		if (!acc.choose(rc, del, khi, false, 0)) return false;
		fac = poly_corr_pts(s1, p1, s2, p2, acc);
		return true;
	}
	if ((KM & KM_BAR) && s1.vol && s2.vol && bar_use_hl(s1, s2)) {
		const int nsub = ribbon_nsub(s1, s2, khi);
		if (ribbon_nsub(s1, s2, klo) != nsub) return false;
		stat = bar_pair(s1, s2);
		if (!acc.choose(rc, del, khi, false, 0)) return false;
		fac = ribbon_corr_pts(s1, s2, nsub, acc);
		return true;
	}
	if ((KM & (KM_RIBBON | KM_BAR)) && ((s1.wid > 0) || (s2.wid > 0))) {
		const int nsub = ribbon_nsub(s1, s2, khi);
		if (ribbon_nsub(s1, s2, klo) != nsub) return false;
		stat = ribbon_static(s1, s2, nsub, ribbon_nq(s1, s2));
		if (!acc.choose(rc, del, khi, false, 0)) return false;
		fac = ribbon_corr_pts(s1, s2, nsub, acc);
		return true;
	}
	const double aeff = 0.5 * (a1 + a2);
	const double lmax = (s1.len > s2.len) ? s1.len : s2.len;
	int nsub = 2 + (int)(khi * lmax);
	if (nsub > 8) nsub = 8;
	int nlo = 2 + (int)(klo * lmax);
	if (nlo > 8) nlo = 8;
	if (nlo != nsub) return false;
	stat = (cross_norm(s1.t, s2.t) < 1e-9) ? wire_pair_static(s1, s2, a1, a2) : wire_gauss<0>(s1, s2, aeff, 0, nsub).r;
	if (!acc.choose(rc, del, khi, true, aeff)) return false;
	fac = wire_corr_pts(s1, s2, aeff, nsub, acc);
	return true;
}

// neumann_self_k split
template <int KM>
__device__ bool self_split(const dcell_t &s, const polys &pp, int c, double a, double klo, double khi, double del,
                           double &stat, double &fac, macc &acc)
{
	if (s.wid > 0) return pair_split<KM>(s, pp, c, s, pp, c, a, a, klo, khi, true, del, stat, fac, acc);
	stat = wire_self_static(s.len, a);
	int nsub = 2 + (int)(khi * s.len);
	if (nsub > 8) nsub = 8;
	int nlo = 2 + (int)(klo * s.len);
	if (nlo > 8) nlo = 8;
	if (nlo != nsub) return false;
	if (!acc.choose(0, del, khi, true, a)) return false;
	fac = wire_corr_pts(s, s, a, nsub, acc);
	return true;
}

struct kf_args {
	fill_args f;
	double klo, khi;
	const double *rho;     // bounding radius of each cell about its centre
	double *rec;           // T x KF_Q x nslots
	long nslots;
	int T;
	int2 *list;            // pairs evaluated directly
	unsigned *nlist;
};

__device__ __forceinline__ void put_term(const kf_args &a, long slot, int t, double mult, double stat, double fac,
                                         const macc &acc)
{
	double *r = a.rec + (long)t * KF_Q * a.nslots + slot;
	const double ff = mult * fac;
	r[0] = mult * (stat - (fac * acc.N[0]));
	r[a.nslots] = acc.Rc;
#pragma unroll
	for (int m = 0; m <= KF_M; m++) r[(long)(2 + m) * a.nslots] = ff * acc.N[m];
}

__device__ __forceinline__ void zero_term(const kf_args &a, long slot, int t)
{
	double *r = a.rec + (long)t * KF_Q * a.nslots + slot;
	for (int q = 0; q < KF_Q; q++) r[(long)q * a.nslots] = 0;
}

// a listed pair's records are zero (its evaluation from them gives 0, overwritten by the direct
// value) with Rc = -1 in the first term marking it for other readers
__device__ __forceinline__ void to_list(const kf_args &a, long slot, int i, int j)
{
	for (int t = 0; t < a.T; t++) zero_term(a, slot, t);
	a.rec[slot + a.nslots] = -1.0;
	const unsigned k = atomicAdd(a.nlist, 1u);
	a.list[k] = make_int2(i, j);
}

// records of Lp pairs: lp_entry (partial.c:371-424) term by term
template <int KM>
__global__ __launch_bounds__(256) void k_kf_setup_lp(kf_args a)
{
	const fill_args &f = a.f;
	int ti, tj;
	const long tile = (long)blockIdx.x + f.tile0;
	tri_tile_kf(tile, f.nt, ti, tj);
	const int i = ti * 16 + (threadIdx.x >> 4);
	const int j = tj * 16 + (threadIdx.x & 15);
	if ((i >= f.n) || (j >= f.n) || (j < i)) return;
	const long slot = tile * 256 + threadIdx.x;
	const dcell_t si = f.S[i];
	macc acc;
	double stat = 0, fac = 0;
	if (i == j) {
		if (!self_split<KM>(si, f.pS, i, si.aL, a.klo, a.khi, 2 * a.rho[i], stat, fac, acc)) return to_list(a, slot, i, j);
		put_term(a, slot, 0, f.coef, stat, fac, acc);
		if (f.gp) {
			const dcell_t mi = f.SM[i];
			const double tdm = (si.t[0] * si.t[0]) + (si.t[1] * si.t[1]) - (si.t[2] * si.t[2]);
			if (!pair_split<KM>(si, f.pS, i, mi, f.pSM, i, si.aL, si.aL, a.klo, a.khi, false, 2 * a.rho[i], stat, fac, acc))
				return to_list(a, slot, i, j);
			put_term(a, slot, 1, -f.coef * tdm, stat, fac, acc);
		}
		return;
	}
	const dcell_t sj = f.S[j];
	const double tdot = dot3(si.t, sj.t);
	const double tdm = (si.t[0] * sj.t[0]) + (si.t[1] * sj.t[1]) - (si.t[2] * sj.t[2]);
	const double del = a.rho[i] + a.rho[j];
	if (tdot != 0) {
		if (!pair_split<KM>(si, f.pS, i, sj, f.pS, j, si.aL, sj.aL, a.klo, a.khi, false, del, stat, fac, acc))
			return to_list(a, slot, i, j);
		put_term(a, slot, 0, f.coef * tdot, stat, fac, acc);
	}
	else {
		zero_term(a, slot, 0);
	}
	if (f.gp) {
		if (tdm != 0) {
			const dcell_t mj = f.SM[j];
			if (!pair_split<KM>(si, f.pS, i, mj, f.pSM, j, si.aL, sj.aL, a.klo, a.khi, false, del, stat, fac, acc))
				return to_list(a, slot, i, j);
			put_term(a, slot, 1, -f.coef * tdm, stat, fac, acc);
		}
		else {
			zero_term(a, slot, 1);
		}
	}
}

// records of charge-cell pairs: the G(h1, h2) rule of pot_fill (potential.c:217-237)
template <int KM>
__global__ __launch_bounds__(256) void k_kf_setup_ig(kf_args a)
{
	const fill_args &f = a.f;
	int ti, tj;
	const long tile = (long)blockIdx.x + f.tile0;
	tri_tile_kf(tile, f.nt, ti, tj);
	const int i = ti * 16 + (threadIdx.x >> 4);
	const int j = tj * 16 + (threadIdx.x & 15);
	if ((i >= f.n) || (j >= f.n) || (j < i)) return;
	const long slot = tile * 256 + threadIdx.x;
	const dcell_t c1 = f.S[i];
	macc acc;
	double stat = 0, fac = 0;
	const double del = a.rho[i] + a.rho[j];
	bool ok;
	if (i == j) ok = self_split<KM>(c1, f.pS, i, c1.aP, a.klo, a.khi, del, stat, fac, acc);
	else {
		const dcell_t c2 = f.S[j];
		ok = pair_split<KM>(c1, f.pS, i, c2, f.pS, j, c1.aP, c2.aP, a.klo, a.khi, false, del, stat, fac, acc);
	}
	if (!ok) return to_list(a, slot, i, j);
	put_term(a, slot, 0, 1.0, stat, fac, acc);
	if (f.gp) {
		const dcell_t m2 = f.SM[j];
		if (!pair_split<KM>(c1, f.pS, i, m2, f.pSM, j, c1.aP, m2.aP, a.klo, a.khi, false, del, stat, fac, acc))
			return to_list(a, slot, i, j);
		put_term(a, slot, 1, -1.0, stat, fac, acc);
	}
}

// entries at wavenumber kw from the records (directly evaluated pairs are written afterwards
// by the list kernel).  which 0: Lp; 1: G, or P when f.carea is set (one sub-cell per cell).
__global__ __launch_bounds__(256) void k_kf_eval(kf_args a, int which)
{
	const fill_args &f = a.f;
	int ti, tj;
	const long tile = (long)blockIdx.x + f.tile0;
	tri_tile_kf(tile, f.nt, ti, tj);
	const int i = ti * 16 + (threadIdx.x >> 4);
	const int j = tj * 16 + (threadIdx.x & 15);
	if ((i >= f.n) || (j >= f.n) || (j < i)) return;
	const long slot = tile * 256 + threadIdx.x;
	double vr, vi;
	kf_value(a.rec, a.nslots, slot, a.T, f.kw, vr, vi);
	double2 *out = (double2 *)f.out;
	if ((which == 1) && (f.carea != nullptr)) {
		const double ai = f.carea[i], aj = f.carea[j];
		const double sij = 1 / (f.k4pe * ai * aj);
		out[(long)i * f.ld + j] = make_double2(sij * (0 + vr), sij * (0 + vi));
		if (i != j) {
			const double sji = 1 / (f.k4pe * aj * ai);
			out[(long)j * f.ld + i] = make_double2(sji * (0 + vr), sji * (0 + vi));
		}
		return;
	}
	out[(long)i * f.ld + j] = make_double2(vr, vi);
	if (i != j) out[(long)j * f.ld + i] = make_double2(vr, vi);
}

// bounding radius of each cell about its packed centre c: ribbons/bars half the diagonal of
// their mid-surface, wires half the length, polygons the farthest fan vertex
__global__ void k_kf_rho(const dcell_t *C, polys pp, int n, double *rho)
{
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= n) return;
	const dcell_t s = C[e];
	double r = 0;
	auto upd = [&](const double *v) {
		const double dx = v[0] - s.c[0], dy = v[1] - s.c[1], dz = v[2] - s.c[2];
		const double d = sqrt((dx * dx) + (dy * dy) + (dz * dz));
		if (d > r) r = d;
	};
	if (s.npv > 0) {
		upd(s.papex);
		const dedge_t *ed = pp.edge + s.poff;
		for (int k = 0; k < s.npv; k++) {
			upd(ed[k].a);
			upd(ed[k].b);
		}
	}
	else if (s.wid > 0) {
		r = 0.5 * sqrt((s.len * s.len) + (s.wid * s.wid));
	}
	else {
		upd(s.x1);
		upd(s.x2);
	}
	rho[e] = r * (1 + 1e-12);
}

}  // namespace pe

// ------------------------------------------------------------------------------------------

namespace {
template <int KM>
void kf_setup_km(hipStream_t s, int which, pe::kf_args a)
{
	const long ntiles = (long)a.f.nt * (a.f.nt + 1) / 2;
	const long chunk = 1L << 20;
	for (long t0 = 0; t0 < ntiles; t0 += chunk) {
		a.f.tile0 = t0;
		const unsigned nb = (unsigned)((ntiles - t0 < chunk) ? (ntiles - t0) : chunk);
		if (which == 0) hipLaunchKernelGGL(pe::k_kf_setup_lp<KM>, dim3(nb), dim3(256), 0, s, a);
		else hipLaunchKernelGGL(pe::k_kf_setup_ig<KM>, dim3(nb), dim3(256), 0, s, a);
	}
}

int km_class_kf(int km)
{
	if ((km & ~pe::KM_WIRE) == 0) return pe::KM_WIRE;
	if ((km & ~pe::KM_RIBBON) == 0) return pe::KM_RIBBON;
	if ((km & ~(pe::KM_RIBBON | pe::KM_BAR)) == 0) return pe::KM_RIBBON | pe::KM_BAR;
	if ((km & ~(pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR)) == 0) return pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR;
	return 15;
}
}  // namespace

long kf_slots(int n)
{
	const long nt = (n + 15) / 16;
	return nt * (nt + 1) / 2 * 256;
}

void kf_rho_launch(hipStream_t s, const dcell_t *C, const dedge_t *edge, int n, double *rho)
{
	pe::polys pp;
	pp.edge = edge;
	pp.pn = nullptr;
	hipLaunchKernelGGL(pe::k_kf_rho, dim3((n + 255) / 256), dim3(256), 0, s, C, pp, n, rho);
	HIPCHK(hipGetLastError());
}

void kf_setup_launch(hipStream_t s, int which, int km, const pe::fill_args &f, double klo, double khi, const double *rho,
                     double *rec, int T, int2 *list, unsigned *nlist)
{
	pe::kf_args a;
	a.f = f;
	a.klo = klo;
	a.khi = khi;
	a.rho = rho;
	a.rec = rec;
	a.nslots = kf_slots(f.n);
	a.T = T;
	a.list = list;
	a.nlist = nlist;
	switch (km_class_kf(km)) {
	case pe::KM_WIRE: kf_setup_km<pe::KM_WIRE>(s, which, a); break;
	case pe::KM_RIBBON: kf_setup_km<pe::KM_RIBBON>(s, which, a); break;
	case pe::KM_RIBBON | pe::KM_BAR: kf_setup_km<pe::KM_RIBBON | pe::KM_BAR>(s, which, a); break;
	case pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR: kf_setup_km<pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR>(s, which, a); break;
	default: kf_setup_km<15>(s, which, a); break;
	}
	HIPCHK(hipGetLastError());
}

namespace pe {
// wavenumbers and outputs of a batch of frequencies
struct kf_multi {
	double kw[KF_FMAX];
	void *out[KF_FMAX];
	int nk;
	int upper;   // mirrors only inside the diagonal 64 x 64 blocks (the rest of the lower triangle untouched)
};

// k_kf_eval for nk wavenumbers at once: each record read once, nk entries (and their mirrors)
// written, one output matrix per wavenumber (same ld).  With mk.upper the mirror (j, i) is
// written only when i and j share a 64-row block: a symmetric product that reads the upper
// 64 x 64 tiles, whole diagonal tiles included, needs no more, and the column-strided mirror
// stores are most of the kernel's write traffic.
template <int T>
__global__ __launch_bounds__(256) void k_kf_eval_multi(kf_args a, int which, kf_multi mk)
{
	const fill_args &f = a.f;
	int ti, tj;
	const long tile = (long)blockIdx.x + f.tile0;
	tri_tile_kf(tile, f.nt, ti, tj);
	const int i = ti * 16 + (threadIdx.x >> 4);
	const int j = tj * 16 + (threadIdx.x & 15);
	if ((i >= f.n) || (j >= f.n) || (j < i)) return;
	const long slot = tile * 256 + threadIdx.x;
	// the record's T terms (base, and the ground-plane image when present) held in registers,
	// then one frequency at a time evaluated and stored
	double q[T][KF_Q];
#pragma unroll
	for (int t = 0; t < T; t++)
#pragma unroll
		for (int m = 0; m < KF_Q; m++) q[t][m] = a.rec[(long)(t * KF_Q + m) * a.nslots + slot];
	const bool scaled = (which == 1) && (f.carea != nullptr);
	double sij = 1, sji = 1;
	if (scaled) {
		const double ai = f.carea[i], aj = f.carea[j];
		sij = 1 / (f.k4pe * ai * aj);
		sji = 1 / (f.k4pe * aj * ai);
	}
	const bool mirror = (i != j) && (!mk.upper || ((i >> 6) == (j >> 6)));
	for (int k = 0; k < mk.nk; k++) {
		double vr = 0, vi = 0;
#pragma unroll
		for (int t = 0; t < T; t++) kf_term(q[t], mk.kw[k], vr, vi);
		double2 *out = (double2 *)mk.out[k];
		if (scaled) {
			out[(long)i * f.ld + j] = make_double2(sij * (0 + vr), sij * (0 + vi));
			if (mirror) out[(long)j * f.ld + i] = make_double2(sji * (0 + vr), sji * (0 + vi));
		}
		else {
			out[(long)i * f.ld + j] = make_double2(vr, vi);
			if (mirror) out[(long)j * f.ld + i] = make_double2(vr, vi);
		}
	}
}
}  // namespace pe

void kf_eval_multi_launch(hipStream_t s, int which, const pe::fill_args &f, const double *rec, int T, int nk,
                          const double *kw, void *const *out, bool upper)
{
	pe::kf_args a;
	a.f = f;
	a.klo = a.khi = 0;
	a.rho = nullptr;
	a.rec = (double *)rec;
	a.nslots = kf_slots(f.n);
	a.T = T;
	a.list = nullptr;
	a.nlist = nullptr;
	pe::kf_multi mk{};
	mk.nk = nk;
	mk.upper = upper ? 1 : 0;
	for (int k = 0; k < nk; k++) {
		mk.kw[k] = kw[k];
		mk.out[k] = out[k];
	}
	const long ntiles = (long)f.nt * (f.nt + 1) / 2;
	const long chunk = 1L << 20;
	for (long t0 = 0; t0 < ntiles; t0 += chunk) {
		a.f.tile0 = t0;
		const unsigned nb = (unsigned)((ntiles - t0 < chunk) ? (ntiles - t0) : chunk);
		if (T > 1) hipLaunchKernelGGL(pe::k_kf_eval_multi<2>, dim3(nb), dim3(256), 0, s, a, which, mk);
		else hipLaunchKernelGGL(pe::k_kf_eval_multi<1>, dim3(nb), dim3(256), 0, s, a, which, mk);
	}
	HIPCHK(hipGetLastError());
}

void kf_eval_launch(hipStream_t s, int which, const pe::fill_args &f, const double *rec, int T)
{
	pe::kf_args a;
	a.f = f;
	a.klo = a.khi = 0;
	a.rho = nullptr;
	a.rec = (double *)rec;
	a.nslots = kf_slots(f.n);
	a.T = T;
	a.list = nullptr;
	a.nlist = nullptr;
	const long ntiles = (long)f.nt * (f.nt + 1) / 2;
	const long chunk = 1L << 20;
	for (long t0 = 0; t0 < ntiles; t0 += chunk) {
		a.f.tile0 = t0;
		const unsigned nb = (unsigned)((ntiles - t0 < chunk) ? (ntiles - t0) : chunk);
		hipLaunchKernelGGL(pe::k_kf_eval, dim3(nb), dim3(256), 0, s, a, which);
	}
	HIPCHK(hipGetLastError());
}

int kf_record_doubles(void)
{
	return pe::KF_Q;
}
