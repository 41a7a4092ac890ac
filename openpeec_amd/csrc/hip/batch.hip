// batch.hip — many structures x many frequencies in persistent batched kernels.
//
// A caller with many small structures (the reference solves each deck with its own solve(),
// solve.c:171-430, one frequency after the other) hands them over at once.  Every (structure,
// frequency) pair is a job: one dense charge-unknown MNA system (unknowns [nodes | L | V |
// branch currents | cell charges], the reference's compressed formulation mna.c:86-107 with the
// kernel blocks dense: -jw Lp on the branch block, P on the charge block, no P^-1), factored
// and solved for every port.  Per chunk of frequencies (sized to an HBM budget) the device work
// is two hipGraphs:
//
//   fill graph    k_b_zero   clears the chunk's system arena (a kernel, so the graph holds no
//                            memset node)
//                 k_b_lp     -jw Lp(f) of every job from the structure's k-factored records
//                            (one sincos + 15-term Horner per pair), written into the system
//                 k_b_g      sub-cell integrals G(f) from the charge records into a scratch
//                            (tile order, coalesced record reads)
//                 k_b_p      P(f) = sum of sub-cell integrals / (4 pi eps0 A1 A2), in the
//                            reference's accumulation order
//   solve graph   k_b_stamp  the sparse stamps (lumped elements, incidence, internal impedance,
//                            charge couplings), computed on the host per job
//                 k_b_reset  work-queue counter
//                 k_b_lu     persistent: every workgroup takes the next job from the queue
//                            (largest first or interleaved, as tuned), factors it in place
//                            (right-looking blocked LU, |re|+|im| first-maximum row pivoting as
//                            lu.c:16-66, FP64 panel, TRSM, trailing update on FP64 MFMA), solves
//                            every port right-hand side and writes Z
//
// A structure whose k-factored records list pairs for direct evaluation (quadrature order
// changing inside the band) is filled element by element per job between the two graphs.
//
// The LU panel width, the workgroups per CU of the persistent grid and the job order are chosen
// by a timing autotuner on a sample of the first chunk's jobs (peec_batch_tune); the choice is
// kept per batch shape.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <thread>
#include <tuple>
#include <vector>
#include "hip_check.hpp"
#include "fill_kernels.hpp"
#include "kf_eval.hpp"
#include "fill_entries.hpp"
#include "../peec_internal.h"
#include "peec_mi355x.h"

namespace {

constexpr int BMAXN = 4096;    // largest system order of a batched job (LDS holds b and pivots)
constexpr int BMAXPORT = 8;
constexpr int BSW = 8;          // sub-panel width inside a batched panel (s_L holds BSW^2 + BSW NB)

typedef double d4 __attribute__((ext_vector_type(4)));

struct bdesc {                 // one structure
	int N;                     // system order
	int offS, nseg;            // branch block
	int offQ, ncell;           // charge block (ncell 0: none)
	int T;                     // record terms
	int ntS, ntH;              // tile grid sizes of the Lp and sub-cell triangles
	int nh;                    // charge sub-cells
	long nslS, nslH;           // record slots
	const double *rec_lp, *rec_ig;
	const int *csoff, *csidx;
	const double *carea;
	int np;
	int i1[BMAXPORT], i2[BMAXPORT];   // port node rows (-1: reference node)
	int direct;                // filled element by element (records unusable)
	int nd_lp, nd_ig;          // pairs without a convergent expansion, evaluated directly per job
	const int2 *list_lp, *list_ig;
	pe::fill_args fs, fc;      // current / charge cell sets for the element code
};

struct bjob {                  // one structure at one frequency
	int s;
	int N;
	double f, omega, kw;
	long aoff;                 // system offset in the arena (entries)
	long zoff;                 // Z offset in the chunk's output (entries)
	long tS, tP, tG;           // first Lp tile, P tile, sub-cell tile of the job in the flattened grids
	long tL;                   // first directly evaluated pair of the job
	long goff;                 // sub-cell integral block offset in the G scratch (entries)
};

// ---------------------------------------------------------------------------------- fills

__global__ void k_b_zero(double2 *A, long cnt)
{
	const long stride = (long)gridDim.x * blockDim.x;
	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += stride) A[e] = make_double2(0, 0);
}

// job of a flattened tile index: last job whose first tile is <= L (tiles ascend with jobs)
__device__ __forceinline__ int job_of(const bjob *J, int nj, long L, int which)
{
	int lo = 0, hi = nj - 1;
	while (lo < hi) {
		const int mid = (lo + hi + 1) >> 1;
		const long t = (which == 0) ? J[mid].tS : (which == 1) ? J[mid].tP : (which == 2) ? J[mid].tG : J[mid].tL;
		if (t <= L) lo = mid;
		else hi = mid - 1;
	}
	return lo;
}

__device__ __forceinline__ void tri_of(long L, int nt, int &ti, int &tj)
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

// -jw Lp of every job: one 16 x 16 tile of a job's upper triangle per workgroup
__global__ __launch_bounds__(256) void k_b_lp(const bdesc *D, const bjob *J, int nj, double2 *arena)
{
	const long L = blockIdx.x;
	const int jb = job_of(J, nj, L, 0);
	const bjob job = J[jb];
	const bdesc &d = D[job.s];
	if (d.direct) return;
	int ti, tj;
	const long t = L - job.tS;
	tri_of(t, d.ntS, ti, tj);
	const int i = ti * 16 + (threadIdx.x >> 4), j = tj * 16 + (threadIdx.x & 15);
	if ((i >= d.nseg) || (j >= d.nseg) || (j < i)) return;
	double vr, vi;
	pe::kf_value(d.rec_lp, d.nslS, t * 256 + threadIdx.x, d.T, job.kw, vr, vi);
	// d_mul((0, -w), v)
	const double w = job.omega;
	const double2 x = make_double2((0 * vr) - ((-w) * vi), (0 * vi) + ((-w) * vr));
	double2 *A = arena + job.aoff;
	const long N = job.N;
	A[(long)(d.offS + i) * N + d.offS + j] = x;
	if (i != j) A[(long)(d.offS + j) * N + d.offS + i] = x;
}

// sub-cell integrals G(h1, h2) of every job from the charge records: one 16 x 16 tile of a job's
// upper triangle per workgroup (slot order, so the record reads coalesce), mirrored
__global__ __launch_bounds__(256) void k_b_g(const bdesc *D, const bjob *J, int nj, double2 *G)
{
	const long L = blockIdx.x;
	const int jb = job_of(J, nj, L, 2);
	const bjob job = J[jb];
	const bdesc &d = D[job.s];
	if (d.direct || (d.ncell == 0)) return;
	int ti, tj;
	const long t = L - job.tG;
	tri_of(t, d.ntH, ti, tj);
	const int nh = d.nh;
	const int i = ti * 16 + (threadIdx.x >> 4), j = tj * 16 + (threadIdx.x & 15);
	if ((i >= nh) || (j >= nh) || (j < i)) return;
	double vr, vi;
	pe::kf_value(d.rec_ig, d.nslH, t * 256 + threadIdx.x, d.T, job.kw, vr, vi);
	double2 *g = G + job.goff;
	g[(long)i * nh + j] = make_double2(vr, vi);
	if (i != j) g[(long)j * nh + i] = make_double2(vr, vi);
}

// pairs the records list for direct evaluation: the element code at the job's wavenumber,
// overwriting the zero the record evaluation left (-jw Lp into the system, G into the scratch)
template <int KM>
__global__ __launch_bounds__(256) void k_b_list(const bdesc *D, const bjob *J, int nj, long total, double2 *arena,
                                                double2 *G)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= total) return;
	const int jb = job_of(J, nj, e, 3);
	const bjob job = J[jb];
	const bdesc &d = D[job.s];
	const int q = (int)(e - job.tL);
	if (q < d.nd_lp) {
		const int2 ij = d.list_lp[q];
		const pe::fill_args &a = d.fs;
		const pe::cplx v = pe::lp_entry<KM>(a.S, a.pS, a.SM, a.pSM, ij.x, ij.y, job.kw, a.coef, a.gp);
		const double w = job.omega;
		const double2 x = make_double2((0 * v.r) - ((-w) * v.i), (0 * v.i) + ((-w) * v.r));
		double2 *A = arena + job.aoff;
		const long N = job.N;
		A[(long)(d.offS + ij.x) * N + d.offS + ij.y] = x;
		if (ij.x != ij.y) A[(long)(d.offS + ij.y) * N + d.offS + ij.x] = x;
		return;
	}
	const int2 ij = d.list_ig[q - d.nd_lp];
	const pe::fill_args &a = d.fc;
	const pe::cplx v = pe::ig_entry<KM>(a.S, a.pS, a.SM, a.pSM, ij.x, ij.y, job.kw, a.gp);
	double2 *g = G + job.goff;
	g[(long)ij.x * d.nh + ij.y] = make_double2(v.r, v.i);
	if (ij.x != ij.y) g[(long)ij.y * d.nh + ij.x] = make_double2(v.r, v.i);
}

// P of every job: P(c1, c2) = sum over sub-cell pairs of G / (4 pi eps0 A1 A2) in the
// reference's accumulation order (potential.c:245-256); one 16 x 16 tile of a job's cell block
// per workgroup
__global__ __launch_bounds__(256) void k_b_p(const bdesc *D, const bjob *J, int nj, double k4pe, const double2 *G,
                                             double2 *arena)
{
	const long L = blockIdx.x;
	const int jb = job_of(J, nj, L, 1);
	const bjob job = J[jb];
	const bdesc &d = D[job.s];
	if (d.direct || (d.ncell == 0)) return;
	const int ntc = (d.ncell + 15) / 16;
	const long t = L - job.tP;
	const int ti = (int)(t / ntc), tj = (int)(t - (long)ti * ntc);
	const int c1 = ti * 16 + (threadIdx.x >> 4), c2 = tj * 16 + (threadIdx.x & 15);
	if ((c1 >= d.ncell) || (c2 >= d.ncell)) return;
	const double2 *g = G + job.goff;
	const int nh = d.nh;
	double sr = 0, si = 0;
	for (int p = d.csoff[c1]; p < d.csoff[c1 + 1]; p++) {
		const long row = (long)d.csidx[p] * nh;
		for (int q = d.csoff[c2]; q < d.csoff[c2 + 1]; q++) {
			const double2 v = g[row + d.csidx[q]];
			sr = sr + v.x;
			si = si + v.y;
		}
	}
	const double s = 1 / (k4pe * d.carea[c1] * d.carea[c2]);
	arena[job.aoff + (long)(d.offQ + c1) * job.N + d.offQ + c2] = make_double2(s * sr, s * si);
}

// element-by-element blocks of one job (structures with directly evaluated pairs)
__global__ void k_b_scatter(double2 *A, int N, int offS, int nseg, const double2 *lp, double omega, int offQ, int nc,
                            const double2 *P)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	const long ns2 = (long)nseg * nseg, nc2 = (long)nc * nc;
	if (e < ns2) {
		const int r = (int)(e / nseg), c = (int)(e - (long)r * nseg);
		const double2 x = lp[e];
		A[(long)(offS + r) * N + offS + c] = make_double2((0 * x.x) - ((-omega) * x.y), (0 * x.y) + ((-omega) * x.x));
	}
	else if ((P != nullptr) && (e < ns2 + nc2)) {
		const long q = e - ns2;
		const int r = (int)(q / nc), c = (int)(q - (long)r * nc);
		A[(long)(offQ + r) * N + offQ + c] = P[q];
	}
}

__global__ void k_b_stamp(const bjob *J, const int *sj, const int *ri, const int *ci, const double2 *val, long nst,
                          double2 *arena)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nst) return;
	const bjob &job = J[sj[e]];
	double2 *a = &arena[job.aoff + (long)ri[e] * job.N + ci[e]];
	const double2 x = *a, v = val[e];
	*a = make_double2(v.x + x.x, v.y + x.y);
}

__global__ void k_b_reset(unsigned *counter)
{
	if ((blockIdx.x == 0) && (threadIdx.x == 0)) *counter = 0u;
}

// ---------------------------------------------------------------------------------- LU

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
	return make_double2((a.x * b.x) - (a.y * b.y), (a.x * b.y) + (a.y * b.x));
}

__device__ __forceinline__ double2 crecip(double2 a)
{
	const double d = (a.x * a.x) + (a.y * a.y);
	return make_double2(a.x / d, -a.y / d);
}

// x -= l u in two fused multiply-adds per component (the library builds without contraction,
// so the LU's update loops state their fused form)
__device__ __forceinline__ void cfms(double2 &x, double2 l, double2 u)
{
	x.x = fma(-l.x, u.x, x.x);
	x.x = fma(l.y, u.y, x.x);
	x.y = fma(-l.x, u.y, x.y);
	x.y = fma(-l.y, u.x, x.y);
}

// (value, index) maximum over the workgroup; ties keep the smaller index
__device__ __forceinline__ void wg_argmax(double &v, int &ix, double *rv, int *ri)
{
	for (int off = 32; off > 0; off >>= 1) {
		const double ov = __shfl_xor(v, off);
		const int oi = __shfl_xor(ix, off);
		if ((ov > v) || ((ov == v) && (oi < ix))) {
			v = ov;
			ix = oi;
		}
	}
	const int wave = threadIdx.x >> 6;
	if ((threadIdx.x & 63) == 0) {
		rv[wave] = v;
		ri[wave] = ix;
	}
	__syncthreads();
	v = rv[0];
	ix = ri[0];
	for (int w = 1; w < 4; w++)
		if ((rv[w] > v) || ((rv[w] == v) && (ri[w] < ix))) {
			v = rv[w];
			ix = ri[w];
		}
	__syncthreads();
}

// A22 -= L21 U12 (row-major, leading dimension ld, K = nb <= NB): each wave takes 32 x 16 CB tiles
// (2 x CB blocks of 16 x 16), complex as four real FP64 MFMA 16x16x4 products per block, with the
// next k-step's fragments loaded before the current step's products
template <int NB, int CB>
__device__ __forceinline__ void trailing_update(double2 *A, long ld, int k0, int nb, int m)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int li = lane & 15, lk = lane >> 4;
	const int tr = (m + 31) / 32, tc = (m + 16 * CB - 1) / (16 * CB);
	const double2 *L21 = A + (long)(k0 + nb) * ld + k0;        // m x nb
	const double2 *U12 = A + (long)k0 * ld + k0 + nb;          // nb x m
	double2 *C = A + (long)(k0 + nb) * ld + k0 + nb;
#pragma unroll 1
	for (int t = wave; t < tr * tc; t += 4) {
		const int r0 = (t / tc) * 32, c0 = (t % tc) * (16 * CB);
		d4 ar[2][CB], ai[2][CB];
#pragma unroll
		for (int a = 0; a < 2; a++)
#pragma unroll
			for (int b = 0; b < CB; b++) {
				ar[a][b] = d4{0, 0, 0, 0};
				ai[a][b] = d4{0, 0, 0, 0};
			}
		const double2 *arow[2];
		const double2 *bcol[CB];
		bool aok[2], bok[CB];
#pragma unroll
		for (int rb = 0; rb < 2; rb++) {
			const int r = r0 + rb * 16 + li;
			aok[rb] = r < m;
			arow[rb] = L21 + (long)(aok[rb] ? r : 0) * ld;
		}
#pragma unroll
		for (int cb = 0; cb < CB; cb++) {
			const int c = c0 + cb * 16 + li;
			bok[cb] = c < m;
			bcol[cb] = U12 + (bok[cb] ? c : 0);
		}
		double2 fa[2], fb[CB];
#pragma unroll
		for (int rb = 0; rb < 2; rb++) fa[rb] = (aok[rb] && (lk < nb)) ? arow[rb][lk] : make_double2(0, 0);
#pragma unroll
		for (int cb = 0; cb < CB; cb++) fb[cb] = (bok[cb] && (lk < nb)) ? bcol[cb][(long)lk * ld] : make_double2(0, 0);
#pragma unroll 1
		for (int k = 0; k < nb; k += 4) {
			const int kk = k + 4 + lk;
			double2 na[2], nb4[CB];
#pragma unroll
			for (int rb = 0; rb < 2; rb++) na[rb] = (aok[rb] && (kk < nb)) ? arow[rb][kk] : make_double2(0, 0);
#pragma unroll
			for (int cb = 0; cb < CB; cb++) nb4[cb] = (bok[cb] && (kk < nb)) ? bcol[cb][(long)kk * ld] : make_double2(0, 0);
#pragma unroll
			for (int rb = 0; rb < 2; rb++)
#pragma unroll
				for (int cb = 0; cb < CB; cb++) ar[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[rb].x, fb[cb].x, ar[rb][cb], 0, 0, 0);
#pragma unroll
			for (int rb = 0; rb < 2; rb++)
#pragma unroll
				for (int cb = 0; cb < CB; cb++) ai[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[rb].x, fb[cb].y, ai[rb][cb], 0, 0, 0);
#pragma unroll
			for (int rb = 0; rb < 2; rb++)
#pragma unroll
				for (int cb = 0; cb < CB; cb++) ar[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(-fa[rb].y, fb[cb].y, ar[rb][cb], 0, 0, 0);
#pragma unroll
			for (int rb = 0; rb < 2; rb++)
#pragma unroll
				for (int cb = 0; cb < CB; cb++) ai[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[rb].y, fb[cb].x, ai[rb][cb], 0, 0, 0);
#pragma unroll
			for (int q = 0; q < 2; q++) fa[q] = na[q];
#pragma unroll
			for (int q = 0; q < CB; q++) fb[q] = nb4[q];
		}
		// the tile's C entries, every load issued before the first store
		double2 cv[2][CB][4];
#pragma unroll
		for (int rb = 0; rb < 2; rb++)
#pragma unroll
			for (int cb = 0; cb < CB; cb++)
#pragma unroll
				for (int v = 0; v < 4; v++) {
					const int r = r0 + rb * 16 + lk + 4 * v, c = c0 + cb * 16 + li;
					cv[rb][cb][v] = ((r < m) && (c < m)) ? C[(long)r * ld + c] : make_double2(0, 0);
				}
#pragma unroll
		for (int rb = 0; rb < 2; rb++)
#pragma unroll
			for (int cb = 0; cb < CB; cb++)
#pragma unroll
				for (int v = 0; v < 4; v++) {
					const int r = r0 + rb * 16 + lk + 4 * v, c = c0 + cb * 16 + li;
					if ((r < m) && (c < m)) {
						double2 x = cv[rb][cb][v];
						x.x -= ar[rb][cb][v];
						x.y -= ai[rb][cb][v];
						C[(long)r * ld + c] = x;
					}
				}
	}
}

typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double2 ld2(const double2 &v) { return v; }
__device__ __forceinline__ double2 ld2(const float2 &v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ void st2(double2 &d, double2 v) { d = v; }
__device__ __forceinline__ void st2(float2 &d, double2 v) { d = make_float2((float)v.x, (float)v.y); }

// The same update on FP32 storage: FP32 MFMA 16x16x4 products accumulated in FP32 (the
// accumulator holds rows 4 (lane / 16) + v of the block)
template <int NB, int CB>
__device__ __forceinline__ void trailing_update(float2 *A, long ld, int k0, int nb, int m)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	const int li = lane & 15, lk = lane >> 4;
	const int tr = (m + 31) / 32, tc = (m + 16 * CB - 1) / (16 * CB);
	const float2 *L21 = A + (long)(k0 + nb) * ld + k0;
	const float2 *U12 = A + (long)k0 * ld + k0 + nb;
	float2 *C = A + (long)(k0 + nb) * ld + k0 + nb;
#pragma unroll 1
	for (int t = wave; t < tr * tc; t += 4) {
		const int r0 = (t / tc) * 32, c0 = (t % tc) * (16 * CB);
		f32x4 ar[2][CB], ai[2][CB];
#pragma unroll
		for (int a = 0; a < 2; a++)
#pragma unroll
			for (int b = 0; b < CB; b++) {
				ar[a][b] = f32x4{0, 0, 0, 0};
				ai[a][b] = f32x4{0, 0, 0, 0};
			}
		const float2 *arow[2];
		const float2 *bcol[CB];
		bool aok[2], bok[CB];
#pragma unroll
		for (int rb = 0; rb < 2; rb++) {
			const int r = r0 + rb * 16 + li;
			aok[rb] = r < m;
			arow[rb] = L21 + (long)(aok[rb] ? r : 0) * ld;
		}
#pragma unroll
		for (int cb = 0; cb < CB; cb++) {
			const int c = c0 + cb * 16 + li;
			bok[cb] = c < m;
			bcol[cb] = U12 + (bok[cb] ? c : 0);
		}
		float2 cv[2][CB][4];
#pragma unroll
		for (int rb = 0; rb < 2; rb++)
#pragma unroll
			for (int cb = 0; cb < CB; cb++)
#pragma unroll
				for (int v = 0; v < 4; v++) {
					const int r = r0 + rb * 16 + 4 * lk + v, c = c0 + cb * 16 + li;
					cv[rb][cb][v] = ((r < m) && (c < m)) ? C[(long)r * ld + c] : make_float2(0, 0);
				}
#pragma unroll 1
		for (int k = 0; k < nb; k += 4) {
			const int kk = k + lk;
			float2 fa[2], fb[CB];
#pragma unroll
			for (int rb = 0; rb < 2; rb++) fa[rb] = (aok[rb] && (kk < nb)) ? arow[rb][kk] : make_float2(0, 0);
#pragma unroll
			for (int cb = 0; cb < CB; cb++) fb[cb] = (bok[cb] && (kk < nb)) ? bcol[cb][(long)kk * ld] : make_float2(0, 0);
#pragma unroll
			for (int rb = 0; rb < 2; rb++)
#pragma unroll
				for (int cb = 0; cb < CB; cb++) {
					ar[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[rb].x, fb[cb].x, ar[rb][cb], 0, 0, 0);
					ar[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(-fa[rb].y, fb[cb].y, ar[rb][cb], 0, 0, 0);
					ai[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[rb].x, fb[cb].y, ai[rb][cb], 0, 0, 0);
					ai[rb][cb] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[rb].y, fb[cb].x, ai[rb][cb], 0, 0, 0);
				}
		}
#pragma unroll
		for (int rb = 0; rb < 2; rb++)
#pragma unroll
			for (int cb = 0; cb < CB; cb++)
#pragma unroll
				for (int v = 0; v < 4; v++) {
					const int r = r0 + rb * 16 + 4 * lk + v, c = c0 + cb * 16 + li;
					if ((r < m) && (c < m)) {
						const float2 x = cv[rb][cb][v];
						C[(long)r * ld + c] = make_float2(x.x - ar[rb][cb][v], x.y - ai[rb][cb][v]);
					}
				}
	}
}

// a value as the storage type holds it
template <typename TS>
__device__ __forceinline__ double2 rnd(double2 v)
{
	return v;
}
template <>
__device__ __forceinline__ double2 rnd<float2>(double2 v)
{
	return make_double2((double)(float)v.x, (double)(float)v.y);
}

constexpr int BREG = 5;        // sub-panel rows held per thread: systems up to 256 BREG in registers

// The unblocked steps of one sub-panel (columns c0 .. ce-1, ce - c0 <= BSW) with its rows in
// registers: thread t holds rows t + 256 i (i < R).  Pivot search, row interchange, scaling and
// rank-1 update touch no global memory; the pivot row reaches every thread through s_row[0, BSW)
// and the displaced row its new owner through s_row[BSW, 2 BSW).  Values are rounded to the
// storage type after every operation, so the factors are those of the column-by-column form.
// Nothing else writes the columns outside the sub-panel meanwhile, so their interchanges are
// applied after the last step, in step order, each thread down its own columns.
template <int R, typename TS>
__device__ __forceinline__ void subpanel_reg(TS *A, int N, int c0, int ce, int *piv, double2 *s_row, double *s_rv, int *s_ri,
                                             int &info)
{
	const int tid = threadIdx.x;
	double2 *s_y = s_row, *s_x = s_row + BSW;
	double2 sub[R][BSW];
	int swp[BSW];
#pragma unroll
	for (int i = 0; i < R; i++) {
		const int r = tid + 256 * i;
#pragma unroll
		for (int j = 0; j < BSW; j++)
			sub[i][j] = ((r >= c0) && (r < N) && (c0 + j < ce)) ? ld2(A[(long)r * N + c0 + j]) : make_double2(0, 0);
	}
	int done = 0;
#pragma unroll
	for (int jc = 0; jc < BSW; jc++) {
		const int c = c0 + jc;
		swp[jc] = c;
		if ((c >= ce) || (info >= 0)) continue;
		double best = -1;
		int bi = N;
#pragma unroll
		for (int i = 0; i < R; i++) {
			const int r = tid + 256 * i;
			if ((r >= c) && (r < N)) {
				const double a = fabs(sub[i][jc].x) + fabs(sub[i][jc].y);
				if (a > best) {
					best = a;
					bi = r;
				}
			}
		}
		wg_argmax(best, bi, s_rv, s_ri);
		if (best < 1e-30) {
			info = c;
			continue;
		}
		if (tid == 0) piv[c] = bi;
		swp[jc] = bi;
		done = jc + 1;
		const int oc = c & 255, ic = c >> 8, ob = bi & 255, ib = bi >> 8;
		if (tid == ob) {
#pragma unroll
			for (int i = 0; i < R; i++)
				if (i == ib)
#pragma unroll
					for (int j = 0; j < BSW; j++) s_y[j] = sub[i][j];
		}
		if ((tid == oc) && (bi != c)) {
#pragma unroll
			for (int i = 0; i < R; i++)
				if (i == ic)
#pragma unroll
					for (int j = 0; j < BSW; j++) s_x[j] = sub[i][j];
		}
		__syncthreads();
		if (bi != c) {
			if (tid == oc) {
#pragma unroll
				for (int i = 0; i < R; i++)
					if (i == ic)
#pragma unroll
						for (int j = 0; j < BSW; j++) sub[i][j] = s_y[j];
			}
			if (tid == ob) {
#pragma unroll
				for (int i = 0; i < R; i++)
					if (i == ib)
#pragma unroll
						for (int j = 0; j < BSW; j++) sub[i][j] = s_x[j];
			}
		}
		const double2 inv = crecip(s_y[jc]);
		double2 u[BSW];
#pragma unroll
		for (int t = jc + 1; t < BSW; t++) u[t] = s_y[t];
#pragma unroll
		for (int i = 0; i < R; i++) {
			const int r = tid + 256 * i;
			if ((r > c) && (r < N)) {
				const double2 l = rnd<TS>(cmul(sub[i][jc], inv));
				sub[i][jc] = l;
#pragma unroll
				for (int t = jc + 1; t < BSW; t++) {
					if (c0 + t >= ce) break;
					double2 x = sub[i][t];
					cfms(x, l, u[t]);
					sub[i][t] = rnd<TS>(x);
				}
			}
		}
		__syncthreads();
	}
#pragma unroll
	for (int i = 0; i < R; i++) {
		const int r = tid + 256 * i;
		if ((r >= c0) && (r < N))
#pragma unroll
			for (int j = 0; j < BSW; j++)
				if (c0 + j < ce) st2(A[(long)r * N + c0 + j], sub[i][j]);
	}
	for (int col = tid; col < N; col += 256) {
		if ((col >= c0) && (col < ce)) continue;
#pragma unroll
		for (int jc = 0; jc < BSW; jc++) {
			if (jc >= done) break;
			const int c = c0 + jc, p = swp[jc];
			if (p != c) {
				const TS t = A[(long)c * N + col];
				A[(long)c * N + col] = A[(long)p * N + col];
				A[(long)p * N + col] = t;
			}
		}
	}
	__syncthreads();
}

// in-place LU of the N x N row-major system A: right-looking, panel width NB, row pivoting by
// the first maximum of |re| + |im| (lu.c:19-30), full-row interchanges recorded in piv (any
// address space); returns -1, or the column whose pivot is below 1e-30.  All 256 threads call.
// Systems up to 256 BREG factor each sub-panel in registers (subpanel_reg); larger ones step
// column by column through memory.
template <int NB, int CB, typename TS>
__device__ __forceinline__ int blu_factor(TS *A, int N, int *piv, double2 *s_row, double2 *s_L, double *s_rv, int *s_ri)
{
	const int tid = threadIdx.x;
	int info = -1;
	for (int k0 = 0; (k0 < N) && (info < 0); k0 += NB) {
		const int nb = (N - k0 < NB) ? (N - k0) : NB;
		const int kend = k0 + nb;
		// the panel in sub-panels of BSW columns: inside a sub-panel the unblocked steps update
		// only its own columns; then its U rows for the rest of the panel (forward substitution
		// with its unit-lower part) and one rank-BSW update of the rest of the panel's rows, so
		// the panel is swept once per sub-panel instead of once per column
		for (int c0 = k0; (c0 < kend) && (info < 0); c0 += BSW) {
			const int ce = min(c0 + BSW, kend);
			if (N <= 256 * BREG) subpanel_reg<BREG, TS>(A, N, c0, ce, piv, s_row, s_rv, s_ri, info);
			else {
				for (int c = c0; c < ce; c++) {
					double best = -1;
					int bi = N;
					for (int r = c + tid; r < N; r += 256) {
						const double2 v = ld2(A[(long)r * N + c]);
						const double a = fabs(v.x) + fabs(v.y);
						if (a > best) {
							best = a;
							bi = r;
						}
					}
					wg_argmax(best, bi, s_rv, s_ri);
					if (best < 1e-30) {
						info = c;
						break;
					}
					if (tid == 0) piv[c] = bi;
					if (bi != c)
						for (int col = tid; col < N; col += 256) {
							const TS t = A[(long)c * N + col];
							A[(long)c * N + col] = A[(long)bi * N + col];
							A[(long)bi * N + col] = t;
						}
					__syncthreads();
					const double2 inv = crecip(ld2(A[(long)c * N + c]));
					for (int r = c + 1 + tid; r < N; r += 256) st2(A[(long)r * N + c], cmul(ld2(A[(long)r * N + c]), inv));
					const int w = ce - c - 1;
					if (tid < w) s_row[tid] = ld2(A[(long)c * N + c + 1 + tid]);
					__syncthreads();
					if (w > 0) {
						// rank-1 update of the sub-panel's remaining columns, every row below c
						for (int r = c + 1 + tid; r < N; r += 256) {
							TS *ar = A + (long)r * N;
							const double2 l = ld2(ar[c]);
							for (int t = 0; t < w; t++) {
								const double2 u = s_row[t];
								double2 x = ld2(ar[c + 1 + t]);
								x.x -= (l.x * u.x) - (l.y * u.y);
								x.y -= (l.x * u.y) + (l.y * u.x);
								st2(ar[c + 1 + t], x);
							}
						}
					}
					__syncthreads();
				}
			}
			if (info >= 0) break;
			const int wr = kend - ce, sw = ce - c0;
			if (wr > 0) {
				// s_L[0 .. BSW^2): the sub-panel's unit-lower part; s_U = s_L + BSW^2: its U rows
				// for the panel's remaining columns (stride NB)
				double2 *s_U = s_L + BSW * BSW;
				for (int e = tid; e < sw * sw; e += 256) {
					const int i = e / sw, j = e - i * sw;
					s_L[i * BSW + j] = ld2(A[(long)(c0 + i) * N + c0 + j]);
				}
				__syncthreads();
				if (tid < wr) {
					const int col = ce + tid;
					double2 u[BSW];
#pragma unroll
					for (int i = 0; i < BSW; i++) {
						if (i >= sw) break;
						double2 x = ld2(A[(long)(c0 + i) * N + col]);
#pragma unroll
						for (int j = 0; j < BSW; j++) {
							if (j >= i) break;
							cfms(x, s_L[i * BSW + j], u[j]);
						}
						u[i] = x;
						st2(A[(long)(c0 + i) * N + col], x);
						s_U[i * NB + tid] = x;
					}
				}
				__syncthreads();
				// rank-sw update of the rest of the panel, every row below the sub-panel
				for (int r = ce + tid; r < N; r += 256) {
					TS *ar = A + (long)r * N;
					double2 l[BSW];
#pragma unroll
					for (int j = 0; j < BSW; j++) l[j] = (j < sw) ? ld2(ar[c0 + j]) : make_double2(0, 0);
					for (int t = 0; t < wr; t++) {
						double2 x = ld2(ar[ce + t]);
#pragma unroll
						for (int j = 0; j < BSW; j++) {
							if (j >= sw) break;
							cfms(x, l[j], s_U[j * NB + t]);
						}
						st2(ar[ce + t], x);
					}
				}
				__syncthreads();
			}
		}
		if (info >= 0) break;
		const int m = N - kend;
		if (m <= 0) break;
		// U12 = L11^-1 A12 once the panel's interchanges are all applied: every thread owns
		// columns (loads and stores coalesced across threads, row by row)
		for (int e = tid; e < nb * nb; e += 256) {
			const int r = e / nb, cc = e - r * nb;
			s_L[r * NB + cc] = ld2(A[(long)(k0 + r) * N + k0 + cc]);
		}
		__syncthreads();
		// blocked forward substitution down a thread's column, 8 rows in registers at a time:
		// the rows already solved are read back once per block, each row stored once (about
		// 3 nb + nb^2 / 16 accesses per column instead of nb^2 read-modify-writes); rows from nb
		// on, in a short last panel, are computed from stale s_L entries and never stored
		for (int col = kend + tid; col < N; col += 256) {
			for (int i0 = 0; i0 < nb; i0 += 8) {
				double2 u[8];
#pragma unroll
				for (int q = 0; q < 8; q++) u[q] = (i0 + q < nb) ? ld2(A[(long)(k0 + i0 + q) * N + col]) : make_double2(0, 0);
#pragma unroll 1
				for (int c = 0; c < i0; c++) {
					const double2 x = ld2(A[(long)(k0 + c) * N + col]);
#pragma unroll
					for (int q = 0; q < 8; q++) cfms(u[q], s_L[(i0 + q) * NB + c], x);
				}
#pragma unroll
				for (int c = 0; c < 7; c++)
#pragma unroll
					for (int q = c + 1; q < 8; q++) cfms(u[q], s_L[(i0 + q) * NB + i0 + c], u[c]);
#pragma unroll
				for (int q = 0; q < 8; q++)
					if (i0 + q < nb) st2(A[(long)(k0 + i0 + q) * N + col], u[q]);
			}
		}
		__syncthreads();
		trailing_update<NB, CB>(A, N, k0, nb, m);
		__syncthreads();
	}
	return info;
}

// v <- U^-1 L^-1 P v for the packed factors of one job (v in LDS or global memory), all threads
template <typename TS>
__device__ void lu_subst(const TS *A, int N, const int *piv, double2 *v)
{
	const int tid = threadIdx.x;
	if (tid == 0)
		for (int c = 0; c < N; c++) {
			const int pr = piv[c];
			if (pr != c) {
				const double2 t = v[c];
				v[c] = v[pr];
				v[pr] = t;
			}
		}
	__syncthreads();
	for (int c = 0; c < N - 1; c++) {
		const double2 xc = v[c];
		if ((xc.x != 0) || (xc.y != 0))
			for (int r = c + 1 + tid; r < N; r += 256) {
				cfms(v[r], ld2(A[(long)r * N + c]), xc);
			}
		__syncthreads();
	}
	for (int c = N - 1; c >= 0; c--) {
		if (tid == 0) v[c] = cmul(v[c], crecip(ld2(A[(long)c * N + c])));
		__syncthreads();
		const double2 xc = v[c];
		for (int r = tid; r < c; r += 256) {
			cfms(v[r], ld2(A[(long)r * N + c]), xc);
		}
		__syncthreads();
	}
}

constexpr int BRF_MAX = 10;          // refinement sweeps of the FP32 batched factorisation
constexpr double BRF_TARGET = 1e-10;  // componentwise backward error the FP32 form guarantees
constexpr double BRF_STOP = 1e-12;    // refinement ends at this backward error

// persistent: factor and solve jobs from the queue until it is empty.  Dynamic LDS holds the
// right-hand side (N complex) and the pivot rows (N ints) of the current job.  FP64 storage
// (arena, TS = double2) or FP32 storage (TS = float2) with every port solution refined against
// the FP64 system A64 (r = b - A64 x in FP64, x += LU^-1 r) until the componentwise backward
// error max_i |r_i| / (|A64| |x| + |b|)_i is at most BRF_STOP, or stops falling by half, or
// reaches 2 eps64; a job whose error stays above BRF_TARGET is marked (info -2) for an FP64
// factorisation.  rscr holds one residual vector per workgroup.
template <int NB, int CB, typename TS>
__global__ __launch_bounds__(256) void k_b_lu(const bdesc *D, const bjob *J, const int *order, int nj, unsigned *counter,
                                              TS *arena, const double2 *arena64, double2 *rscr, long rstride,
                                              double2 *zout, int *info_out, double *berr_out, int *sweeps_out)
{
	extern __shared__ double2 s_dyn[];
	__shared__ double2 s_row[NB];
	__shared__ double2 s_L[NB * NB];
	__shared__ double s_rv[4];
	__shared__ int s_ri[4];
	__shared__ int s_job;
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	for (;;) {
		if (tid == 0) s_job = (int)atomicAdd(counter, 1u);
		__syncthreads();
		const int q = s_job;
		__syncthreads();
		if (q >= nj) break;
		const int jid = order[q];
		const bjob job = J[jid];
		const bdesc &d = D[job.s];
		const int N = job.N;
		TS *A = arena + job.aoff;
		double2 *b = s_dyn;
		int *piv = (int *)(s_dyn + N);
		const int info = blu_factor<NB, CB, TS>(A, N, piv, s_row, s_L, s_rv, s_ri);
		if (info >= 0) {
			if (tid == 0) info_out[jid] = info;
			__syncthreads();
			continue;
		}
		double worst = 0;
		int sweeps = 0;
		for (int pj = 0; pj < d.np; pj++) {
			const int i1 = d.i1[pj], i2 = d.i2[pj];
			for (int r = tid; r < N; r += 256) b[r] = make_double2(0, 0);
			__syncthreads();
			if (tid == 0) {
				if (i1 >= 0) b[i1].x += 1;
				if (i2 >= 0) b[i2].x -= 1;
			}
			__syncthreads();
			lu_subst<TS>(A, N, piv, b);
			if constexpr (sizeof(TS) == sizeof(float2)) {
				const double2 *A64 = arena64 + job.aoff;
				double2 *r = rscr + (long)blockIdx.x * rstride;
				double be = 1, last = 2;
				for (int it = 0;; it++) {
					double wmax = 0;
					for (int row = wave; row < N; row += 4) {
						const double2 *a = A64 + (long)row * N;
						double sr = 0, si = 0, sa = 0;
						for (int c = lane; c < N; c += 64) {
							const double2 u = a[c], x = b[c];
							sr += (u.x * x.x) - (u.y * x.y);
							si += (u.x * x.y) + (u.y * x.x);
							sa += sqrt((u.x * u.x) + (u.y * u.y)) * sqrt((x.x * x.x) + (x.y * x.y));
						}
						for (int off = 32; off > 0; off >>= 1) {
							sr += __shfl_xor(sr, off);
							si += __shfl_xor(si, off);
							sa += __shfl_xor(sa, off);
						}
						const double rhs = ((row == i1) ? 1.0 : 0.0) - ((row == i2) ? 1.0 : 0.0);
						const double rr = rhs - sr, ri = -si;
						if (lane == 0) r[row] = make_double2(rr, ri);
						const double den = sa + fabs(rhs), num = sqrt((rr * rr) + (ri * ri));
						const double e = (den > 0) ? (num / den) : ((num > 0) ? 1.0 : 0.0);
						wmax = (e > wmax) ? e : wmax;
					}
					if (lane == 0) s_rv[wave] = wmax;
					__syncthreads();
					be = s_rv[0];
					for (int w = 1; w < 4; w++) be = (s_rv[w] > be) ? s_rv[w] : be;
					__syncthreads();
					if ((be <= BRF_STOP) || (be <= 4.5e-16) || (be > 0.5 * last) || (it == BRF_MAX)) break;
					last = be;
					lu_subst<TS>(A, N, piv, r);
					for (int i = tid; i < N; i += 256) b[i] = make_double2(b[i].x + r[i].x, b[i].y + r[i].y);
					__syncthreads();
					sweeps++;
				}
				worst = (be > worst) ? be : worst;
			}
			if (tid < d.np) {
				const double2 v1 = (d.i1[tid] >= 0) ? b[d.i1[tid]] : make_double2(0, 0);
				const double2 v2 = (d.i2[tid] >= 0) ? b[d.i2[tid]] : make_double2(0, 0);
				zout[job.zoff + (long)tid * d.np + pj] = make_double2(v1.x - v2.x, v1.y - v2.y);
			}
			__syncthreads();
		}
		if (tid == 0) {
			info_out[jid] = (worst > BRF_TARGET) ? -2 : -1;
			if (berr_out) berr_out[jid] = worst;
			if (sweeps_out) sweeps_out[jid] = sweeps;
		}
		__syncthreads();
	}
}

// the FP32 copy of every job's system
__global__ void k_b_demote(const double2 *src, float2 *dst, long cnt)
{
	const long stride = (long)gridDim.x * blockDim.x;
	for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += stride)
		dst[e] = make_float2((float)src[e].x, (float)src[e].y);
}

// one matrix per workgroup through blu_factor (the batched factorisation on its own)
template <int NB>
__global__ __launch_bounds__(256) void k_b_lu_only(double2 *A, const long *aoff, const int *n, int *piv, const long *poff,
                                                   int *info)
{
	__shared__ double2 s_row[NB];
	__shared__ double2 s_L[NB * NB];
	__shared__ double s_rv[4];
	__shared__ int s_ri[4];
	const int j = blockIdx.x;
	const int r = blu_factor<NB, 2, double2>(A + aoff[j], n[j], piv + poff[j], s_row, s_L, s_rv, s_ri);
	if (threadIdx.x == 0) info[j] = r;
}

// ---------------------------------------------------------------------------------- host

template <typename T>
struct dbuf {                  // device buffer that grows; true when it moved
	T *p = nullptr;
	long cap = 0;
	bool ensure(long n)
	{
		if (n <= cap) return false;
		if (p) HIPCHK(hipFree(p));
		cap = n + n / 4 + 16;
		HIPCHK(hipMalloc((void **)&p, (size_t)cap * sizeof(T)));
		return true;
	}
	~dbuf()
	{
		if (p) HIPCHK(hipFree(p));
	}
};

template <typename T>
struct hbuf {                  // pinned host buffer that grows
	T *p = nullptr;
	long cap = 0;
	void ensure(long n)
	{
		if (n <= cap) return;
		if (p) HIPCHK(hipHostFree(p));
		cap = n + n / 4 + 16;
		HIPCHK(hipHostMalloc((void **)&p, (size_t)cap * sizeof(T), hipHostMallocDefault));
	}
	~hbuf()
	{
		if (p) HIPCHK(hipHostFree(p));
	}
};

struct tune_t {
	int nb = 32;        // LU panel width
	int cb = 2;         // trailing-update tile width in 16-column blocks (2 or 4)
	int wpc = 2;        // persistent workgroups per CU
	int order = 0;      // 0 largest first, 1 interleaved (structure order)
	double ms = 0;      // LU time of the sample with this choice
};

double freq_of(const peec_t *p, int i)
{
	if (p->nfreq <= 1) return p->f0;
	if (p->flog) return p->f0 * pow(p->f1 / p->f0, (double)i / (p->nfreq - 1));
	return p->f0 + (p->f1 - p->f0) * i / (p->nfreq - 1);
}

struct chunk {                 // jobs of one pass through the arena
	std::vector<bjob> J;
	std::vector<int> ord;
	long arena_n = 0, tilesS = 0, tilesP = 0, tilesG = 0, nlist = 0, gn = 0, zn = 0, nst = 0;
	int Nmax = 0;
};

// everything a captured pair of graphs depends on: launch shapes, the tuned variant, and the
// largest order (the LU kernel's dynamic LDS is sized for it when the graph is recorded)
struct gkey {
	int nj;
	long nst, arena_n, tilesS, tilesP, tilesG, nlist;
	int nb, cb, wpc, fp32, Nmax;
	bool operator<(const gkey &o) const
	{
		return std::tie(nj, nst, arena_n, tilesS, tilesP, tilesG, nlist, nb, cb, wpc, fp32, Nmax) <
		       std::tie(o.nj, o.nst, o.arena_n, o.tilesS, o.tilesP, o.tilesG, o.nlist, o.nb, o.cb, o.wpc, o.fp32, o.Nmax);
	}
};

struct batch {
	int nm = 0;
	std::vector<const peec_t *> P;
	std::vector<pgpu_t *> G;
	std::vector<bdesc> hd;
	std::vector<pgpu_bview> views;
	dbuf<bdesc> dd;
	hipStream_t s = nullptr;
	int ncu = 0;
	dbuf<double2> arena, zdev, gscr;
	dbuf<float2> arena32;          // FP32 copies of the systems (precision form)
	dbuf<double2> rscr;            // one residual vector per persistent workgroup (precision form)
	dbuf<double> dberr;
	dbuf<int> dsweeps;
	int fp32 = 0;                  // 1: FP32-stored factorisation with FP64 refinement
	dbuf<bjob> dj;
	dbuf<int> dorder, dinfo, dsj, dri, dci;
	dbuf<unsigned> dcnt;
	dbuf<double2> dval;
	hbuf<int> hsj, hri, hci, hinfo;
	hbuf<double2> hval, hz;
	std::map<gkey, std::pair<hipGraphExec_t, hipGraphExec_t>> graphs;
	int km_class = 15;          // element kinds present (compiled specialisation of the list kernel)
	bool tuned = false;
	tune_t tune;
	std::vector<tune_t> trials;

	~batch()
	{
		if (s) HIPCHK(hipStreamSynchronize(s));
		for (auto &kv : graphs) {
			HIPCHK(hipGraphExecDestroy(kv.second.first));
			HIPCHK(hipGraphExecDestroy(kv.second.second));
		}
		if (s) HIPCHK(hipStreamDestroy(s));
	}
};

size_t lu_lds(int Nmax)
{
	return (size_t)Nmax * sizeof(double2) + (size_t)Nmax * sizeof(int) + 16;
}

template <int NB, int CB, typename TS>
void launch_lu_t(hipStream_t s, batch &B, int grid, int nj, size_t lds, TS *arena, const int *order)
{
	hipLaunchKernelGGL((k_b_lu<NB, CB, TS>), dim3(grid), dim3(256), lds, s, B.dd.p, B.dj.p, order, nj, B.dcnt.p, arena,
	                   (const double2 *)B.arena.p, B.rscr.p, (long)BMAXN, B.zdev.p, B.dinfo.p, B.dberr.p, B.dsweeps.p);
}

// fp32: the FP32 arena with refinement against the FP64 one; else the FP64 arena in place
void launch_lu(hipStream_t s, batch &B, const tune_t &t, int nj, size_t lds, bool fp32, const int *order)
{
	const int grid = std::max(1, std::min(nj, B.ncu * t.wpc));
	if (fp32) {
		float2 *a = B.arena32.p;
		if ((t.nb == 16) && (t.cb == 2)) launch_lu_t<16, 2>(s, B, grid, nj, lds, a, order);
		else if ((t.nb == 16) && (t.cb == 4)) launch_lu_t<16, 4>(s, B, grid, nj, lds, a, order);
		else if ((t.nb == 64) && (t.cb == 2)) launch_lu_t<64, 2>(s, B, grid, nj, lds, a, order);
		else if ((t.nb == 64) && (t.cb == 4)) launch_lu_t<64, 4>(s, B, grid, nj, lds, a, order);
		else if (t.cb == 4) launch_lu_t<32, 4>(s, B, grid, nj, lds, a, order);
		else launch_lu_t<32, 2>(s, B, grid, nj, lds, a, order);
	}
	else {
		double2 *a = B.arena.p;
		if ((t.nb == 16) && (t.cb == 2)) launch_lu_t<16, 2>(s, B, grid, nj, lds, a, order);
		else if ((t.nb == 16) && (t.cb == 4)) launch_lu_t<16, 4>(s, B, grid, nj, lds, a, order);
		else if ((t.nb == 64) && (t.cb == 2)) launch_lu_t<64, 2>(s, B, grid, nj, lds, a, order);
		else if ((t.nb == 64) && (t.cb == 4)) launch_lu_t<64, 4>(s, B, grid, nj, lds, a, order);
		else if (t.cb == 4) launch_lu_t<32, 4>(s, B, grid, nj, lds, a, order);
		else launch_lu_t<32, 2>(s, B, grid, nj, lds, a, order);
	}
}

void record_fill(hipStream_t s, batch &B, const chunk &c)
{
	const int nj = (int)c.J.size();
	hipLaunchKernelGGL(k_b_zero, dim3(4096), dim3(256), 0, s, B.arena.p, c.arena_n);
	if (c.tilesS > 0) hipLaunchKernelGGL(k_b_lp, dim3((unsigned)c.tilesS), dim3(256), 0, s, B.dd.p, B.dj.p, nj, B.arena.p);
	if (c.tilesG > 0) hipLaunchKernelGGL(k_b_g, dim3((unsigned)c.tilesG), dim3(256), 0, s, B.dd.p, B.dj.p, nj, B.gscr.p);
	if (c.nlist > 0) {
		const dim3 g((unsigned)((c.nlist + 255) / 256)), b(256);
		switch (B.km_class) {
		case pe::KM_WIRE: hipLaunchKernelGGL(k_b_list<pe::KM_WIRE>, g, b, 0, s, B.dd.p, B.dj.p, nj, c.nlist, B.arena.p, B.gscr.p); break;
		case pe::KM_RIBBON:
			hipLaunchKernelGGL(k_b_list<pe::KM_RIBBON>, g, b, 0, s, B.dd.p, B.dj.p, nj, c.nlist, B.arena.p, B.gscr.p);
			break;
		default: hipLaunchKernelGGL(k_b_list<15>, g, b, 0, s, B.dd.p, B.dj.p, nj, c.nlist, B.arena.p, B.gscr.p); break;
		}
	}
	if (c.tilesP > 0)
		hipLaunchKernelGGL(k_b_p, dim3((unsigned)c.tilesP), dim3(256), 0, s, B.dd.p, B.dj.p, nj, 4 * PI * EPS0,
		                   (const double2 *)B.gscr.p, B.arena.p);
}

void record_solve(hipStream_t s, batch &B, const chunk &c, const tune_t &t)
{
	const int nj = (int)c.J.size();
	if (c.nst > 0)
		hipLaunchKernelGGL(k_b_stamp, dim3((unsigned)((c.nst + 255) / 256)), dim3(256), 0, s, B.dj.p, B.dsj.p, B.dri.p,
		                   B.dci.p, B.dval.p, c.nst, B.arena.p);
	if (B.fp32) hipLaunchKernelGGL(k_b_demote, dim3(4096), dim3(256), 0, s, (const double2 *)B.arena.p, B.arena32.p, c.arena_n);
	hipLaunchKernelGGL(k_b_reset, dim3(1), dim3(64), 0, s, B.dcnt.p);
	launch_lu(s, B, t, nj, lu_lds(c.Nmax), B.fp32 != 0, B.dorder.p);
}

template <typename F>
hipGraphExec_t capture(hipStream_t s, F body)
{
	hipGraph_t g = nullptr;
	HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
	body(s);
	HIPCHK(hipStreamEndCapture(s, &g));
	hipGraphExec_t ex = nullptr;
	HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
	HIPCHK(hipGraphDestroy(g));
	return ex;
}

// job table of a list of (structure, frequency index) items
void build_chunk(const batch &B, const std::vector<std::pair<int, int>> &items, int order, chunk &c)
{
	c.J.resize(items.size());
	c.arena_n = c.tilesS = c.tilesP = c.tilesG = c.nlist = c.gn = c.zn = c.nst = 0;
	c.Nmax = 0;
	for (size_t k = 0; k < items.size(); k++) {
		const int s = items[k].first, fi = items[k].second;
		const bdesc &d = B.hd[s];
		const double f = freq_of(B.P[s], fi);
		bjob &j = c.J[k];
		j.s = s;
		j.N = d.N;
		j.f = f;
		j.omega = 2 * PI * f;
		j.kw = B.P[s]->retardation ? 2 * PI * f / C0 : 0.0;
		j.aoff = c.arena_n;
		j.zoff = c.zn;
		j.tS = c.tilesS;
		j.tP = c.tilesP;
		j.tG = c.tilesG;
		j.tL = c.nlist;
		j.goff = c.gn;
		if (!d.direct) c.nlist += d.nd_lp + ((d.ncell > 0) ? d.nd_ig : 0);
		c.arena_n += (long)d.N * d.N;
		c.zn += (long)d.np * d.np;
		c.tilesS += (long)d.ntS * (d.ntS + 1) / 2;
		const long ntc = (d.ncell + 15) / 16;
		c.tilesP += ntc * ntc;
		if (d.ncell > 0) {
			c.tilesG += (long)d.ntH * (d.ntH + 1) / 2;
			c.gn += (long)d.nh * d.nh;
		}
		c.Nmax = std::max(c.Nmax, d.N);
	}
	c.ord.resize(items.size());
	for (size_t k = 0; k < items.size(); k++) c.ord[k] = (int)k;
	if (order == 0) std::stable_sort(c.ord.begin(), c.ord.end(), [&](int a, int b) { return c.J[a].N > c.J[b].N; });
}

// stamps of every job on host threads into the pinned staging arrays (job order), then the
// uploads of the job table, the order and the stamps on the batch stream
void stage_chunk(batch &B, chunk &c)
{
	const int nj = (int)c.J.size();
	std::vector<stamps_t> st(nj);
	unsigned nth = std::thread::hardware_concurrency();
	nth = std::max(1u, std::min(nth, 16u));
	std::vector<std::thread> th;
	for (unsigned t = 0; t < nth; t++)
		th.emplace_back([&, t]() {
			for (int k = (int)t; k < nj; k += (int)nth) mna_stamps_charge(B.P[c.J[k].s], c.J[k].f, &st[k]);
		});
	for (auto &x : th) x.join();
	long tot = 0;
	for (int k = 0; k < nj; k++) tot += st[k].n;
	c.nst = tot;
	// the previous chunk's uploads must have been consumed before the staging is overwritten
	HIPCHK(hipStreamSynchronize(B.s));
	B.hsj.ensure(tot);
	B.hri.ensure(tot);
	B.hci.ensure(tot);
	B.hval.ensure(tot);
	long e = 0;
	for (int k = 0; k < nj; k++) {
		for (int q = 0; q < st[k].n; q++) {
			B.hsj.p[e] = k;
			B.hri.p[e] = st[k].ri[q];
			B.hci.p[e] = st[k].ci[q];
			B.hval.p[e] = make_double2(st[k].val[q].r, st[k].val[q].i);
			e++;
		}
		mna_stamps_free(&st[k]);
	}
	bool moved = false;
	moved |= B.dj.ensure(nj);
	moved |= B.dorder.ensure(nj);
	moved |= B.dinfo.ensure(nj);
	moved |= B.dsj.ensure(tot);
	moved |= B.dri.ensure(tot);
	moved |= B.dci.ensure(tot);
	moved |= B.dval.ensure(tot);
	moved |= B.arena.ensure(c.arena_n);
	moved |= B.zdev.ensure(c.zn);
	moved |= B.gscr.ensure(c.gn);
	moved |= B.dberr.ensure(nj);
	moved |= B.dsweeps.ensure(nj);
	if (B.fp32) {
		moved |= B.arena32.ensure(c.arena_n);
		moved |= B.rscr.ensure((long)BMAXN * 2 * B.ncu);
	}
	if (moved) {
		// captured graphs hold the old addresses
		for (auto &kv : B.graphs) {
			HIPCHK(hipGraphExecDestroy(kv.second.first));
			HIPCHK(hipGraphExecDestroy(kv.second.second));
		}
		B.graphs.clear();
	}
	HIPCHK(hipMemcpyAsync(B.dj.p, c.J.data(), (size_t)nj * sizeof(bjob), hipMemcpyHostToDevice, B.s));
	HIPCHK(hipMemcpyAsync(B.dorder.p, c.ord.data(), (size_t)nj * sizeof(int), hipMemcpyHostToDevice, B.s));
	if (tot > 0) {
		HIPCHK(hipMemcpyAsync(B.dsj.p, B.hsj.p, (size_t)tot * sizeof(int), hipMemcpyHostToDevice, B.s));
		HIPCHK(hipMemcpyAsync(B.dri.p, B.hri.p, (size_t)tot * sizeof(int), hipMemcpyHostToDevice, B.s));
		HIPCHK(hipMemcpyAsync(B.dci.p, B.hci.p, (size_t)tot * sizeof(int), hipMemcpyHostToDevice, B.s));
		HIPCHK(hipMemcpyAsync(B.dval.p, B.hval.p, (size_t)tot * sizeof(double2), hipMemcpyHostToDevice, B.s));
	}
}

// structures filled element by element: their blocks of every job, between the two graphs
void direct_blocks(batch &B, const chunk &c)
{
	bool any = false;
	for (const bdesc &d : B.hd) any = any || d.direct;
	if (!any) return;
	HIPCHK(hipStreamSynchronize(B.s));
	for (const bjob &j : c.J) {
		const bdesc &d = B.hd[j.s];
		if (!d.direct) continue;
		const void *lp = nullptr, *Pm = nullptr;
		pgpu_fill_direct_dev(B.G[j.s], j.kw, &lp, &Pm);
		const long tot = (long)d.nseg * d.nseg + (Pm ? (long)d.ncell * d.ncell : 0);
		hipLaunchKernelGGL(k_b_scatter, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, B.s, B.arena.p + j.aoff, d.N,
		                   d.offS, d.nseg, (const double2 *)lp, j.omega, d.offQ, Pm ? d.ncell : 0, (const double2 *)Pm);
		HIPCHK(hipStreamSynchronize(B.s));   // the mirror's buffers are reused by the next job
	}
}

std::pair<hipGraphExec_t, hipGraphExec_t> graphs_for(batch &B, const chunk &c, const tune_t &t)
{
	const gkey k{(int)c.J.size(), c.nst, c.arena_n, c.tilesS, c.tilesP, c.tilesG, c.nlist, t.nb, t.cb, t.wpc, B.fp32, c.Nmax};
	auto it = B.graphs.find(k);
	if (it != B.graphs.end()) return it->second;
	hipGraphExec_t gf = capture(B.s, [&](hipStream_t s) { record_fill(s, B, c); });
	hipGraphExec_t gs = capture(B.s, [&](hipStream_t s) { record_solve(s, B, c, t); });
	B.graphs[k] = {gf, gs};
	return {gf, gs};
}

// runs one staged chunk; returns the LU graph's time (ms) when timed
double run_chunk(batch &B, const chunk &c, const tune_t &t, double *fill_ms)
{
	auto gg = graphs_for(B, c, t);
	hipEvent_t e0, e1, e2;
	HIPCHK(hipEventCreate(&e0));
	HIPCHK(hipEventCreate(&e1));
	HIPCHK(hipEventCreate(&e2));
	HIPCHK(hipEventRecord(e0, B.s));
	HIPCHK(hipGraphLaunch(gg.first, B.s));
	direct_blocks(B, c);
	HIPCHK(hipEventRecord(e1, B.s));
	HIPCHK(hipGraphLaunch(gg.second, B.s));
	HIPCHK(hipEventRecord(e2, B.s));
	HIPCHK(hipEventSynchronize(e2));
	float a = 0, b = 0;
	HIPCHK(hipEventElapsedTime(&a, e0, e1));
	HIPCHK(hipEventElapsedTime(&b, e1, e2));
	HIPCHK(hipEventDestroy(e0));
	HIPCHK(hipEventDestroy(e1));
	HIPCHK(hipEventDestroy(e2));
	if (fill_ms) *fill_ms = a;
	return b;
}

int prepare(batch &B, peec_t **models, int nm, int ifreq0, int nfreq)
{
	int dev = 0;
	HIPCHK(hipGetDevice(&dev));
	HIPCHK(hipDeviceGetAttribute(&B.ncu, hipDeviceAttributeMultiprocessorCount, dev));
	HIPCHK(hipStreamCreateWithFlags(&B.s, hipStreamNonBlocking));
	{
		const int lds = (int)lu_lds(BMAXN);
		const void *fns[] = {(const void *)k_b_lu<16, 2, double2>, (const void *)k_b_lu<16, 4, double2>,
		                     (const void *)k_b_lu<32, 2, double2>, (const void *)k_b_lu<32, 4, double2>,
		                     (const void *)k_b_lu<64, 2, double2>, (const void *)k_b_lu<64, 4, double2>,
		                     (const void *)k_b_lu<16, 2, float2>,  (const void *)k_b_lu<16, 4, float2>,
		                     (const void *)k_b_lu<32, 2, float2>,  (const void *)k_b_lu<32, 4, float2>,
		                     (const void *)k_b_lu<64, 2, float2>,  (const void *)k_b_lu<64, 4, float2>};
		for (const void *fn : fns) HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds));
	}
	B.nm = nm;
	B.P.resize(nm);
	B.G.resize(nm);
	B.hd.resize(nm);
	B.views.resize(nm);
	for (int s = 0; s < nm; s++) {
		const peec_t *p = models[s];
		B.P[s] = p;
		if ((p->compress != 0) || (p->nport <= 0) || (p->nport > BMAXPORT) || (ifreq0 < 0) || (ifreq0 + nfreq > p->nfreq)) {
			printf("*** batched sweep: structure %d is not a dense deck with 1..%d ports and frequencies %d..%d\n", s,
			       BMAXPORT, ifreq0, ifreq0 + nfreq - 1);
			return 4;
		}
		const int nc = (p->capacitance && (p->ncell > 0)) ? p->ncell : 0;
		const int N = p->nunknown + nc;
		if (N > BMAXN) {
			printf("*** batched sweep: structure %d has %d unknowns (at most %d per batched system)\n", s, N, BMAXN);
			return 4;
		}
		B.G[s] = pgpu_get(p);
	}
	// k-factored records of every structure over its band, setups running concurrently
	std::vector<int> ok(nm, 0);
	for (int s = 0; s < nm; s++) {
		const peec_t *p = B.P[s];
		const double kr = p->retardation ? 2 * PI / C0 : 0.0;
		const double klo = kr * freq_of(p, ifreq0), khi = kr * freq_of(p, ifreq0 + nfreq - 1);
		ok[s] = (pgpu_kf_setup_launch(B.G[s], std::min(klo, khi), std::max(klo, khi)) == 0);
	}
	for (int s = 0; s < nm; s++)
		if (ok[s]) pgpu_kf_setup_finish(B.G[s]);
	// PEEC_BATCH_DIRECT=1 fills every structure element by element (the path of structures
	// whose records cannot be built)
	const char *fd = getenv("PEEC_BATCH_DIRECT");
	const bool force_direct = (fd != nullptr) && (atoi(fd) != 0);
	int kmu = 0;
	for (int s = 0; s < nm; s++) {
		const peec_t *p = B.P[s];
		pgpu_bview &v = B.views[s];
		pgpu_batch_view(B.G[s], &v);
		bdesc &d = B.hd[s];
		memset(&d, 0, sizeof(d));
		d.ncell = (p->capacitance && (p->ncell > 0)) ? p->ncell : 0;
		d.N = p->nunknown + d.ncell;
		d.offS = p->offS;
		d.nseg = p->nseg;
		d.offQ = p->nunknown;
		d.T = v.T;
		d.ntS = (p->nseg + 15) / 16;
		d.nh = v.nh;
		d.ntH = (v.nh + 15) / 16;
		d.nslS = kf_slots(p->nseg);
		d.nslH = (v.nh > 0) ? kf_slots(v.nh) : 0;
		d.rec_lp = v.rec_lp;
		d.rec_ig = v.rec_ig;
		d.csoff = v.csoff;
		d.csidx = v.csidx;
		d.carea = v.carea;
		d.np = p->nport;
		for (int k = 0; k < p->nport; k++) {
			d.i1[k] = p->nodemap[p->port[k].n1];
			d.i2[k] = p->nodemap[p->port[k].n2];
		}
		d.direct = force_direct || (!ok[s]) || (v.rec_lp == nullptr && p->nseg > 0) || (d.ncell > 0 && v.rec_ig == nullptr);
		d.nd_lp = v.nd_lp;
		d.nd_ig = v.nd_ig;
		d.list_lp = v.list_lp;
		d.list_ig = v.list_ig;
		d.fs = v.seg;
		d.fc = v.chg;
		kmu |= v.km_seg | ((d.ncell > 0) ? v.km_chg : 0);
	}
	B.km_class = ((kmu & ~pe::KM_WIRE) == 0) ? pe::KM_WIRE : ((kmu & ~pe::KM_RIBBON) == 0) ? pe::KM_RIBBON : 15;
	B.dd.ensure(nm);
	HIPCHK(hipMemcpy(B.dd.p, B.hd.data(), (size_t)nm * sizeof(bdesc), hipMemcpyHostToDevice));
	B.dcnt.ensure(1);
	return 0;
}

// HBM budget of one chunk's systems (bytes)
double arena_budget()
{
	const char *e = getenv("PEEC_BATCH_GB");
	const double gb = e ? atof(e) : 24.0;
	return ((gb > 0.01) ? gb : 24.0) * 1e9;
}

// chunks of whole frequency indices (all structures at each), within the HBM budget
std::vector<std::vector<std::pair<int, int>>> make_chunks(const batch &B, int ifreq0, int nfreq)
{
	double per_f = 0;
	for (const bdesc &d : B.hd) per_f += 16.0 * d.N * (double)d.N + ((d.ncell > 0) ? 16.0 * d.nh * (double)d.nh : 0.0);
	int fc = (int)std::max(1.0, std::floor(arena_budget() / std::max(per_f, 1.0)));
	fc = std::min(fc, nfreq);
	// even chunks so every chunk but possibly none differ in shape
	const int nch = (nfreq + fc - 1) / fc;
	fc = (nfreq + nch - 1) / nch;
	std::vector<std::vector<std::pair<int, int>>> out;
	for (int f0 = 0; f0 < nfreq; f0 += fc) {
		std::vector<std::pair<int, int>> items;
		for (int f = f0; f < std::min(nfreq, f0 + fc); f++)
			for (int s = 0; s < B.nm; s++) items.push_back({s, ifreq0 + f});
		out.push_back(items);
	}
	return out;
}

// timing autotuner on a sample of the first chunk: every candidate (panel width, workgroups per
// CU, job order) re-assembles the sample and factors it; the fastest LU wins
void autotune(batch &B, const std::vector<std::pair<int, int>> &first)
{
	std::vector<std::pair<int, int>> sample;
	const size_t stride = std::max<size_t>(1, first.size() / 512);
	for (size_t k = 0; k < first.size(); k += stride) sample.push_back(first[k]);
	const int nbs[3] = {16, 32, 64};
	const int cbs[2] = {2, 4};
	const int wpcs[2] = {1, 2};
	B.trials.clear();
	tune_t best;
	best.ms = 1e300;
	for (int order = 0; order < 2; order++) {
		chunk c;
		build_chunk(B, sample, order, c);
		stage_chunk(B, c);
		for (int nb : nbs)
			for (int cb : cbs)
				for (int wpc : wpcs) {
					tune_t t;
					t.nb = nb;
					t.cb = cb;
					t.wpc = wpc;
					t.order = order;
					t.ms = run_chunk(B, c, t, nullptr);
					B.trials.push_back(t);
					if (t.ms < best.ms) best = t;
				}
	}
	B.tune = best;
	B.tuned = true;
}

}  // namespace

// ------------------------------------------------------------------------------------------
extern "C" {

struct peec_batch {
	batch b;
	int ifreq0 = 0, nfreq = 0;
};

peec_batch_t *peec_batch_create(peec_t **models, int nmodel, int ifreq0, int nfreq)
{
	if ((models == nullptr) || (nmodel <= 0) || (nfreq <= 0)) return nullptr;
	peec_batch_t *h = new peec_batch();
	h->ifreq0 = ifreq0;
	h->nfreq = nfreq;
	const auto t0 = std::chrono::steady_clock::now();
	if (prepare(h->b, models, nmodel, ifreq0, nfreq) != 0) {
		delete h;
		return nullptr;
	}
	(void)t0;
	return h;
}

void peec_batch_free(peec_batch_t *h)
{
	delete h;
}

int peec_batch_tune(peec_batch_t *h, peec_batch_stats_t *st)
{
	if (h == nullptr) return 4;
	const auto t0 = std::chrono::steady_clock::now();
	auto chunks = make_chunks(h->b, h->ifreq0, h->nfreq);
	autotune(h->b, chunks[0]);
	if (st) {
		st->tune_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
		st->nb = h->b.tune.nb;
		st->tile_cols = 16 * h->b.tune.cb;
		st->wg_per_cu = h->b.tune.wpc;
		st->order = h->b.tune.order;
		st->tune_trials = (int)h->b.trials.size();
	}
	return 0;
}

int peec_batch_sweep(peec_batch_t *h, d_complex_t **zmat, peec_batch_stats_t *st)
{
	if (h == nullptr) return 4;
	batch &B = h->b;
	if (st) memset(st, 0, sizeof(*st));
	const auto t0 = std::chrono::steady_clock::now();
	auto chunks = make_chunks(B, h->ifreq0, h->nfreq);
	if (!B.tuned) autotune(B, chunks[0]);
	const tune_t tn = B.tune;
	int rc = 0;
	for (size_t ci = 0; ci < chunks.size() && rc == 0; ci++) {
		chunk c;
		build_chunk(B, chunks[ci], tn.order, c);
		const auto th = std::chrono::steady_clock::now();
		stage_chunk(B, c);
		if (st) st->stamp_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - th).count();
		double fms = 0;
		const double lms = run_chunk(B, c, tn, &fms);
		const int nj = (int)c.J.size();
		B.hz.ensure(c.zn);
		B.hinfo.ensure(nj);
		HIPCHK(hipMemcpyAsync(B.hinfo.p, B.dinfo.p, (size_t)nj * sizeof(int), hipMemcpyDeviceToHost, B.s));
		HIPCHK(hipStreamSynchronize(B.s));
		if (B.fp32) {
			// refinement statistics, and jobs whose refinement did not reach 1e-10: their FP64
			// systems (untouched by the FP32 factorisation) factored and solved in FP64
			std::vector<double> be(nj);
			std::vector<int> sw(nj);
			HIPCHK(hipMemcpyAsync(be.data(), B.dberr.p, (size_t)nj * sizeof(double), hipMemcpyDeviceToHost, B.s));
			HIPCHK(hipMemcpyAsync(sw.data(), B.dsweeps.p, (size_t)nj * sizeof(int), hipMemcpyDeviceToHost, B.s));
			HIPCHK(hipStreamSynchronize(B.s));
			std::vector<int> redo;
			for (int k = 0; k < nj; k++) {
				if (B.hinfo.p[k] == -2) redo.push_back(k);
				if (st && (B.hinfo.p[k] < 0)) {
					st->refine_sweeps += sw[k];
					if (B.hinfo.p[k] == -1 && be[k] > st->berr_max) st->berr_max = be[k];
				}
			}
			if (!redo.empty()) {
				const auto tr = std::chrono::steady_clock::now();
				HIPCHK(hipMemcpyAsync(B.dorder.p, redo.data(), redo.size() * sizeof(int), hipMemcpyHostToDevice, B.s));
				hipLaunchKernelGGL(k_b_reset, dim3(1), dim3(64), 0, B.s, B.dcnt.p);
				launch_lu(B.s, B, tn, (int)redo.size(), lu_lds(c.Nmax), false, B.dorder.p);
				HIPCHK(hipMemcpyAsync(B.hinfo.p, B.dinfo.p, (size_t)nj * sizeof(int), hipMemcpyDeviceToHost, B.s));
				HIPCHK(hipStreamSynchronize(B.s));
				if (st) {
					st->fp64_fallbacks += (int)redo.size();
					st->lu_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tr).count();
				}
			}
		}
		HIPCHK(hipMemcpyAsync(B.hz.p, B.zdev.p, (size_t)c.zn * sizeof(double2), hipMemcpyDeviceToHost, B.s));
		HIPCHK(hipStreamSynchronize(B.s));
		for (int k = 0; k < nj; k++) {
			const bjob &j = c.J[k];
			const int fi = chunks[ci][k].second - h->ifreq0;
			if (B.hinfo.p[k] >= 0) {
				printf("*** batched sweep: singular system (structure %d, f = %.5e Hz, row %d)\n", j.s, j.f,
				       B.hinfo.p[k]);
				rc = 2;
				continue;
			}
			const int np = B.hd[j.s].np;
			d_complex_t *z = zmat[j.s] + (size_t)fi * np * np;
			for (int e = 0; e < np * np; e++) z[e] = d_complex(B.hz.p[j.zoff + e].x, B.hz.p[j.zoff + e].y);
		}
		if (st) {
			st->fill_ms += fms;
			st->lu_ms += lms;
			st->chunks++;
			st->jobs += nj;
			st->direct_pairs += c.nlist;
			for (const bjob &j : c.J) {
				const double N = j.N;
				st->lu_flop += (8.0 / 3.0) * N * N * N + 8.0 * N * N * B.hd[j.s].np;
			}
		}
	}
	if (st) {
		st->total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
		st->nb = tn.nb;
		st->tile_cols = 16 * tn.cb;
		st->wg_per_cu = tn.wpc;
		st->order = tn.order;
		st->tune_trials = (int)B.trials.size();
		for (const bdesc &d : B.hd) st->direct_structures += d.direct;
	}
	return rc;
}

int peec_batch_lu_decomp(int nmat, const int *n, d_complex_t *a, int *piv, int nb)
{
	if ((nmat <= 0) || ((nb != 16) && (nb != 32) && (nb != 64))) return -2;
	std::vector<long> aoff(nmat), poff(nmat);
	long na = 0, np = 0;
	for (int j = 0; j < nmat; j++) {
		if ((n[j] <= 0) || (n[j] > BMAXN)) return -2;
		aoff[j] = na;
		poff[j] = np;
		na += (long)n[j] * n[j];
		np += n[j];
	}
	double2 *dA = nullptr;
	long *daoff = nullptr, *dpoff = nullptr;
	int *dn = nullptr, *dpiv = nullptr, *dinfo = nullptr;
	HIPCHK(hipMalloc((void **)&dA, (size_t)na * sizeof(double2)));
	HIPCHK(hipMalloc((void **)&daoff, (size_t)nmat * sizeof(long)));
	HIPCHK(hipMalloc((void **)&dpoff, (size_t)nmat * sizeof(long)));
	HIPCHK(hipMalloc((void **)&dn, (size_t)nmat * sizeof(int)));
	HIPCHK(hipMalloc((void **)&dpiv, (size_t)np * sizeof(int)));
	HIPCHK(hipMalloc((void **)&dinfo, (size_t)nmat * sizeof(int)));
	HIPCHK(hipMemcpy(dA, a, (size_t)na * sizeof(double2), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(daoff, aoff.data(), (size_t)nmat * sizeof(long), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(dpoff, poff.data(), (size_t)nmat * sizeof(long), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(dn, n, (size_t)nmat * sizeof(int), hipMemcpyHostToDevice));
	if (nb == 16) hipLaunchKernelGGL((k_b_lu_only<16>), dim3(nmat), dim3(256), 0, 0, dA, daoff, dn, dpiv, dpoff, dinfo);
	else if (nb == 64) hipLaunchKernelGGL((k_b_lu_only<64>), dim3(nmat), dim3(256), 0, 0, dA, daoff, dn, dpiv, dpoff, dinfo);
	else hipLaunchKernelGGL((k_b_lu_only<32>), dim3(nmat), dim3(256), 0, 0, dA, daoff, dn, dpiv, dpoff, dinfo);
	HIPCHK(hipGetLastError());
	std::vector<int> info(nmat);
	HIPCHK(hipMemcpy(a, dA, (size_t)na * sizeof(double2), hipMemcpyDeviceToHost));
	HIPCHK(hipMemcpy(piv, dpiv, (size_t)np * sizeof(int), hipMemcpyDeviceToHost));
	HIPCHK(hipMemcpy(info.data(), dinfo, (size_t)nmat * sizeof(int), hipMemcpyDeviceToHost));
	HIPCHK(hipFree(dA));
	HIPCHK(hipFree(daoff));
	HIPCHK(hipFree(dpoff));
	HIPCHK(hipFree(dn));
	HIPCHK(hipFree(dpiv));
	HIPCHK(hipFree(dinfo));
	for (int j = 0; j < nmat; j++)
		if (info[j] >= 0) return j;
	return -1;
}

int peec_batch_set_form(peec_batch_t *h, int form)
{
	if (h == nullptr) return 4;
	if ((form != PEEC_FORM_PLAIN) && (form != PEEC_FORM_PRECISION)) return 4;
	const int f32 = (form == PEEC_FORM_PRECISION) ? 1 : 0;
	if (f32 != h->b.fp32) {
		h->b.fp32 = f32;
		h->b.tuned = false;
	}
	return 0;
}

int peec_batch_trials(const peec_batch_t *h, int *nb, int *tile_cols, int *wpc, int *order, double *ms, int cap)
{
	if (h == nullptr) return 0;
	const int n = (int)h->b.trials.size();
	for (int k = 0; k < n && k < cap; k++) {
		nb[k] = h->b.trials[k].nb;
		tile_cols[k] = 16 * h->b.trials[k].cb;
		wpc[k] = h->b.trials[k].wpc;
		order[k] = h->b.trials[k].order;
		ms[k] = h->b.trials[k].ms;
	}
	return n;
}

}  // extern "C"
