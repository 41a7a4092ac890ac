// fill.hip — dense fills of the retarded partial inductance Lp and the charge-cell integral
// matrix (and, fused, the potential coefficients P) on gfx950.
//
// Work decomposition: the upper triangle (i <= j, as lp_fill / pot_fill evaluate it,
// partial.c:447-451, potential.c:217-237) is cut into 16 x 16 tiles, one 256-thread workgroup
// per tile, one element pair per lane (4 rows x 16 columns per wavefront).  Each lane writes
// both (i, j) and the mirrored (j, i).  Cells are packed 256-byte records read through L2
// (a full cell set is a few MB, resident in every XCD's L2), so the fill is bound by FP64
// VALU and the transcendental calls, not by HBM: its only HBM traffic is the 16 B per entry
// written.  Neighbouring cells have similar distances, so lanes of one wavefront take the
// same quadrature branch except near the diagonal.
//
// Entry rules (reference):
//   Lp(i,j) = mu0/4pi [ tdot I(si,sj) - [gp] tdm I(si, M sj) ]   lp_entry, partial.c:371-424
//   G(h1,h2) = I(h1,h2) - [gp] I(h1, M h2)                       pot_fill, potential.c:217-237
//   P(c1,c2) = sum_{h1 in c1} sum_{h2 in c2} G / (4 pi eps0 A1 A2) potential.c:245-256
#include <hip/hip_runtime.h>
#include "elements.hpp"
#include "hip_check.hpp"
#include "fill_kernels.hpp"
#include "fill_entries.hpp"

namespace pe {

// tile index L of the upper triangle of an nt x nt tile grid -> (ti, tj), ti <= tj
__device__ __forceinline__ void tri_tile(long L, int nt, int &ti, int &tj)
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

template <int KM>
__global__ __launch_bounds__(256) void k_fill_lp(fill_args a)
{
	int ti, tj;
	tri_tile((long)blockIdx.x + a.tile0, a.nt, ti, tj);
	const int i = ti * 16 + (threadIdx.x >> 4);
	const int j = tj * 16 + (threadIdx.x & 15);
	if ((i >= a.n) || (j >= a.n) || (j < i)) return;
	const cplx v = lp_entry<KM>(a.S, a.pS, a.SM, a.pSM, i, j, a.kw, a.coef, a.gp);
	double2 *out = (double2 *)a.out;
	out[(long)i * a.ld + j] = make_double2(v.r, v.i);
	if (i != j) out[(long)j * a.ld + i] = make_double2(v.r, v.i);
}

// charge-cell integrals; when a.carea != nullptr the cells map one-to-one onto capacitive
// cells (cellof = identity, one sub-cell each) and P is written directly:
// P(i,j) = (0 + G) / ((k4pe * A_i) * A_j), with the (j,i) scale evaluated in its own order.
template <int KM>
__global__ __launch_bounds__(256) void k_fill_ig(fill_args a)
{
	int ti, tj;
	tri_tile((long)blockIdx.x + a.tile0, a.nt, ti, tj);
	const int i = ti * 16 + (threadIdx.x >> 4);
	const int j = tj * 16 + (threadIdx.x & 15);
	if ((i >= a.n) || (j >= a.n) || (j < i)) return;
	const cplx v = ig_entry<KM>(a.S, a.pS, a.SM, a.pSM, i, j, a.kw, a.gp);
	double2 *out = (double2 *)a.out;
	if (a.carea == nullptr) {
		out[(long)i * a.ld + j] = make_double2(v.r, v.i);
		if (i != j) out[(long)j * a.ld + i] = make_double2(v.r, v.i);
		return;
	}
	const double ai = a.carea[i], aj = a.carea[j];
	const double sij = 1 / (a.k4pe * ai * aj);
	out[(long)i * a.ld + j] = make_double2(sij * (0 + v.r), sij * (0 + v.i));
	if (i != j) {
		const double sji = 1 / (a.k4pe * aj * ai);
		out[(long)j * a.ld + i] = make_double2(sji * (0 + v.r), sji * (0 + v.i));
	}
}

// P(c1,c2) = scale * sum over sub-cell pairs of G, in the reference's accumulation order
__global__ void k_aggregate(const double2 *G, long ldg, const int *csoff, const int *csidx, const double *carea,
                            int nc, double k4pe, double2 *P, long ldp)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)nc * nc) return;
	const int c1 = (int)(e / nc), c2 = (int)(e - (long)c1 * nc);
	double sr = 0, si = 0;
	for (int p = csoff[c1]; p < csoff[c1 + 1]; p++) {
		const long row = (long)csidx[p] * ldg;
		for (int q = csoff[c2]; q < csoff[c2 + 1]; q++) {
			const double2 g = G[row + csidx[q]];
			sr = sr + g.x;
			si = si + g.y;
		}
	}
	const double s = 1 / (k4pe * carea[c1] * carea[c2]);
	P[(long)c1 * ldp + c2] = make_double2(s * sr, s * si);
}

// listed (i <= j) pairs at kw: the pairs a k-factored fill evaluates directly each frequency
template <int KM>
__global__ __launch_bounds__(256) void k_fill_list(fill_args a, int which, const int2 *list, int cnt)
{
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= cnt) return;
	const int i = list[e].x, j = list[e].y;
	double2 *out = (double2 *)a.out;
	if (which == 0) {
		const cplx v = lp_entry<KM>(a.S, a.pS, a.SM, a.pSM, i, j, a.kw, a.coef, a.gp);
		out[(long)i * a.ld + j] = make_double2(v.r, v.i);
		if (i != j) out[(long)j * a.ld + i] = make_double2(v.r, v.i);
		return;
	}
	const cplx v = ig_entry<KM>(a.S, a.pS, a.SM, a.pSM, i, j, a.kw, a.gp);
	if (a.carea == nullptr) {
		out[(long)i * a.ld + j] = make_double2(v.r, v.i);
		if (i != j) out[(long)j * a.ld + i] = make_double2(v.r, v.i);
		return;
	}
	const double ai = a.carea[i], aj = a.carea[j];
	const double sij = 1 / (a.k4pe * ai * aj);
	out[(long)i * a.ld + j] = make_double2(sij * (0 + v.r), sij * (0 + v.i));
	if (i != j) {
		const double sji = 1 / (a.k4pe * aj * ai);
		out[(long)j * a.ld + i] = make_double2(sji * (0 + v.r), sji * (0 + v.i));
	}
}

// single entries for lp_entry / pot_entry calls through the ABI
template <int KM>
__global__ void k_one_lp(fill_args a, int i, int j, double2 *res)
{
	const cplx v = lp_entry<KM>(a.S, a.pS, a.SM, a.pSM, i, j, a.kw, a.coef, a.gp);
	*res = make_double2(v.r, v.i);
}

template <int KM>
__global__ void k_one_pot(fill_args a, int c1, int c2, const int *csoff, const int *csidx, double2 *res)
{
	double sr = 0, si = 0;
	for (int p = csoff[c1]; p < csoff[c1 + 1]; p++) {
		const int h1 = csidx[p];
		for (int q = csoff[c2]; q < csoff[c2 + 1]; q++) {
			const int h2 = csidx[q];
			const int lo = (h1 <= h2) ? h1 : h2, hi = (h1 <= h2) ? h2 : h1;
			const cplx v = ig_entry<KM>(a.S, a.pS, a.SM, a.pSM, lo, hi, a.kw, a.gp);
			sr = sr + v.r;
			si = si + v.i;
		}
	}
	const double s = 1 / (a.k4pe * a.carea[c1] * a.carea[c2]);
	*res = make_double2(s * sr, s * si);
}

}  // namespace pe

// ------------------------------------------------------------------------------------------
// launchers (C++ linkage inside the library)

namespace {
template <int KM>
void launch_fill(hipStream_t s, int which, const pe::fill_args &a)
{
	const long ntiles = (long)a.nt * (a.nt + 1) / 2;
	const long chunk = 1L << 20;   // tiles per launch (grid dimension stays far below 2^31)
	for (long t0 = 0; t0 < ntiles; t0 += chunk) {
		pe::fill_args b = a;
		b.tile0 = t0;
		const unsigned nb = (unsigned)((ntiles - t0 < chunk) ? (ntiles - t0) : chunk);
		if (which == 0) hipLaunchKernelGGL(pe::k_fill_lp<KM>, dim3(nb), dim3(256), 0, s, b);
		else hipLaunchKernelGGL(pe::k_fill_ig<KM>, dim3(nb), dim3(256), 0, s, b);
	}
}

template <int KM>
void launch_one(hipStream_t s, int which, const pe::fill_args &a, int i, int j, const int *csoff, const int *csidx,
                double2 *res)
{
	if (which == 0) hipLaunchKernelGGL(pe::k_one_lp<KM>, dim3(1), dim3(1), 0, s, a, i, j, res);
	else hipLaunchKernelGGL(pe::k_one_pot<KM>, dim3(1), dim3(1), 0, s, a, i, j, csoff, csidx, res);
}
}  // namespace

// kind mask -> smallest compiled specialisation that covers it
static int km_class(int km)
{
	if ((km & ~pe::KM_WIRE) == 0) return pe::KM_WIRE;
	if ((km & ~pe::KM_RIBBON) == 0) return pe::KM_RIBBON;
	if ((km & ~(pe::KM_RIBBON | pe::KM_BAR)) == 0) return pe::KM_RIBBON | pe::KM_BAR;
	if ((km & ~(pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR)) == 0) return pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR;
	return 15;
}

void fill_launch(hipStream_t s, int which, int km, const pe::fill_args &a)
{
	switch (km_class(km)) {
	case pe::KM_WIRE: launch_fill<pe::KM_WIRE>(s, which, a); break;
	case pe::KM_RIBBON: launch_fill<pe::KM_RIBBON>(s, which, a); break;
	case pe::KM_RIBBON | pe::KM_BAR: launch_fill<pe::KM_RIBBON | pe::KM_BAR>(s, which, a); break;
	case pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR: launch_fill<pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR>(s, which, a); break;
	default: launch_fill<15>(s, which, a); break;
	}
	HIPCHK(hipGetLastError());
}

void fill_list_launch(hipStream_t s, int which, int km, const pe::fill_args &a, const int2 *list, int cnt)
{
	if (cnt <= 0) return;
	const dim3 g((unsigned)((cnt + 255) / 256)), b(256);
	switch (km_class(km)) {
	case pe::KM_WIRE: hipLaunchKernelGGL(pe::k_fill_list<pe::KM_WIRE>, g, b, 0, s, a, which, list, cnt); break;
	case pe::KM_RIBBON: hipLaunchKernelGGL(pe::k_fill_list<pe::KM_RIBBON>, g, b, 0, s, a, which, list, cnt); break;
	case pe::KM_RIBBON | pe::KM_BAR:
		hipLaunchKernelGGL(pe::k_fill_list<pe::KM_RIBBON | pe::KM_BAR>, g, b, 0, s, a, which, list, cnt);
		break;
	case pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR:
		hipLaunchKernelGGL(pe::k_fill_list<pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR>, g, b, 0, s, a, which, list, cnt);
		break;
	default: hipLaunchKernelGGL(pe::k_fill_list<15>, g, b, 0, s, a, which, list, cnt); break;
	}
	HIPCHK(hipGetLastError());
}

void fill_one(hipStream_t s, int which, int km, const pe::fill_args &a, int i, int j, const int *csoff,
              const int *csidx, double2 *res)
{
	switch (km_class(km)) {
	case pe::KM_WIRE: launch_one<pe::KM_WIRE>(s, which, a, i, j, csoff, csidx, res); break;
	case pe::KM_RIBBON: launch_one<pe::KM_RIBBON>(s, which, a, i, j, csoff, csidx, res); break;
	case pe::KM_RIBBON | pe::KM_BAR: launch_one<pe::KM_RIBBON | pe::KM_BAR>(s, which, a, i, j, csoff, csidx, res); break;
	case pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR:
		launch_one<pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR>(s, which, a, i, j, csoff, csidx, res);
		break;
	default: launch_one<15>(s, which, a, i, j, csoff, csidx, res); break;
	}
	HIPCHK(hipGetLastError());
}

void aggregate_launch(hipStream_t s, const double2 *G, long ldg, const int *csoff, const int *csidx,
                      const double *carea, int nc, double k4pe, double2 *P, long ldp)
{
	const long tot = (long)nc * nc;
	hipLaunchKernelGGL(pe::k_aggregate, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, G, ldg, csoff, csidx,
	                   carea, nc, k4pe, P, ldp);
	HIPCHK(hipGetLastError());
}
