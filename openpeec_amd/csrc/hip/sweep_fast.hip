// sweep_fast.hip — accelerated frequency sweep (PEEC_FORM_MIXED).
//
// Same physics and the same FP64 fills as the plain sweep; the solve changes:
//   * charge-unknown formulation (the reference's compressed-path system, mna.c:86-107, written
//     with dense kernel blocks): unknowns [nodes | L | V | branch currents | cell charges], P
//     enters as a block, so no C = P^-1 is formed per frequency.  The product A x is applied
//     block-wise as the reference's mna_apply does (mna.c:338-364): sparse stamps (CSR) plus
//     -jw Lp on the branch-current block plus P on the charge block, so each GMRES product
//     streams Lp and P once (1.4 GB on the 64 x 64 plate) instead of a dense n^2 matrix;
//   * one factorisation of A(f_ref) serves a run of frequencies as the right preconditioner of
//     FP64 flexible GMRES (the reference's acceleration = 1 scheme, solve.c:47-64 / iterative.c:
//     230-252).  It is factored through its Schur complement on the charge unknowns: C = P^-1
//     at f_ref, and the LU of S = A11 + jw E^T C E, the plain system, of order nunknown instead
//     of nunknown + ncell.  The factors of S are inverted once (recursive blocking on MFMA
//     GEMMs) and the two inverses and C stored in FP32, so applying the preconditioner is two
//     bandwidth-bound triangular matrix-vector products and two products with C instead of
//     chains of dependent substitution steps.  Flexible
//     GMRES keeps the preconditioned directions, so the preconditioner may be applied inexactly
//     (FP32 data, any summation order) without limiting the final residual;
//   * frequencies advance in lock-step batches of F systems sharing the preconditioner, so each
//     pass over the inverses serves F right-hand sides at once;
//   * GMRES runs to a relative residual of 1e-10 (ACCEL_TOL, solve.c:21), checked on the true
//     residual b - A x in FP64; a system that fails to converge within the restart budget
//     triggers a fresh factorisation at its own frequency (as accel_rhs does).
// Orthogonalisation is classical Gram-Schmidt with one re-orthogonalisation pass (two batched
// dot-product launches per iteration instead of j serial ones); all reductions run in a fixed
// order, so results are bitwise reproducible run to run.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "../peec_internal.h"
#include "../../../include/peec_mi355x.h"
#include "hip_check.hpp"

namespace {

constexpr int FMAX = 8;        // systems per batch
constexpr int MKRYLOV = 80;    // largest GMRES restart length (buffers are sized for it)
constexpr int MITS = 240;      // iterations of one solve before refactorising
constexpr int TB = 64;         // triangular-solve block
constexpr double TOL = 1e-10;  // relative residual
constexpr int REFACTOR_MIN_ITS = 24;   // batches below this many iterations never trigger a refactor

typedef std::complex<double> zc;

__device__ __forceinline__ double2 zmul(double2 a, double2 b)
{
	return make_double2((a.x * b.x) - (a.y * b.y), (a.x * b.y) + (a.y * b.x));
}

__device__ __forceinline__ double2 zinv(double2 z)
{
	const double n2 = (z.x * z.x) + (z.y * z.y);
	if (n2 <= 0) return make_double2(0, 0);
	return make_double2(z.x / n2, -z.y / n2);
}

// ---------------------------------------------------------------- assembly

// The preconditioner of the charge-unknown system A = [[A11, A12], [A21, P]] at f_ref is
// factored through its Schur complement: with A21 = -E (charge rows: -v of the cell's node) and
// A12 = jw E^T (node rows: jw q), S = A11 - A12 P^-1 A21 = A11 + jw E^T C E, C = P^-1, which is
// the plain MNA system of order nunknown.  Exactly A^-1 in exact arithmetic, at 0.65x the
// factorisation work of the full order on the full stack.

// S: -jw Lp on the branch block, jw C between the node rows of capacitive cells (rowcell[r] =
// the cell on node row r, or -1), 0 elsewhere; the sparse stamps follow (k_stamp_lead)
__global__ void k_assemble_schur(double2 *A, int n, int offS, int nseg, const double2 *lp, double omega,
                                 const int *rowcell, const double2 *C, int nc)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * n) return;
	const int r = (int)(e / n), c = (int)(e - (long)r * n);
	double2 v = make_double2(0, 0);
	if ((r >= offS) && (r < offS + nseg) && (c >= offS) && (c < offS + nseg)) {
		const double2 x = lp[(long)(r - offS) * nseg + (c - offS)];
		v = make_double2(omega * x.y, -omega * x.x);
	}
	else if (nc > 0) {
		const int cr = rowcell[r], cc = rowcell[c];
		if ((cr >= 0) && (cc >= 0)) {
			const double2 x = C[(long)cr * nc + cc];
			v = make_double2(-omega * x.y, omega * x.x);
		}
	}
	A[e] = v;
}

// the charge-unknown system's stamps that fall inside the leading n x n block (its couplings
// to the charge unknowns are the jw C block of S)
__global__ void k_stamp_lead(double2 *A, long lda, int n, const int *ri, const int *ci, const double2 *val, int nst)
{
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nst) return;
	const int r = ri[e], c = ci[e];
	if ((r >= n) || (c >= n)) return;
	double2 *a = &A[(long)r * lda + c];
	const double2 x = *a, v = val[e];
	*a = make_double2(v.x + x.x, v.y + x.y);
}

// zt[k][f] = r[perm[k]][f] with r = v1 - jw E^T w (w = C v2 per cell): the right-hand side of S,
// permuted for its factors
__global__ void k_permute_schur(const double2 *v, const double2 *w, const int *perm, const int *rowcell, double omega,
                                double2 *zt, int n, int F)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * F) return;
	const int k = (int)(e / F), f = (int)(e - (long)k * F);
	const int r = perm[k];
	double2 x = v[(long)r * F + f];
	const int cr = (rowcell != nullptr) ? rowcell[r] : -1;
	if (cr >= 0) {
		const double2 wv = w[(long)cr * F + f];
		x = make_double2(x.x + omega * wv.y, x.y - omega * wv.x);
	}
	zt[e] = x;
}

// t[i][f] = v2[i][f] + x1[cellrow[i]][f]: the charge rows' right-hand side once the node
// potentials x1 are known (cellrow[i] = the node row of cell i, or -1 on the reference node)
__global__ void k_charge_rhs(const double2 *v2, const double2 *x1, const int *cellrow, double2 *t, int nc, int F)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)nc * F) return;
	const int i = (int)(e / F), f = (int)(e - (long)i * F);
	double2 a = v2[e];
	const int r = cellrow[i];
	if (r >= 0) {
		const double2 x = x1[(long)r * F + f];
		a = make_double2(a.x + x.x, a.y + x.y);
	}
	t[e] = a;
}

__global__ void k_eye(double2 *B, int n)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * n) return;
	const int r = (int)(e / n), c = (int)(e - (long)r * n);
	B[e] = make_double2((r == c) ? 1.0 : 0.0, 0);
}

__global__ void k_demote(const double2 *src, float2 *dst, long count)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= count) return;
	const double2 v = src[e];
	dst[e] = make_float2((float)v.x, (float)v.y);
}

// dst = src^T in FP32 (both n x n, row-major), through 32 x 32 tiles in LDS so both the read
// and the write are coalesced
__device__ __forceinline__ double2 as_d2(const double2 &v) { return v; }
__device__ __forceinline__ double2 as_d2(const float2 &v) { return make_double2(v.x, v.y); }

template <typename TS>
__global__ __launch_bounds__(256) void k_transpose32(const TS *__restrict__ src, float2 *__restrict__ dst, int n)
{
	__shared__ double2 t[32][33];
	const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
	const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
	for (int k = 0; k < 32; k += 8) {
		const int r = r0 + ty + k, c = c0 + tx;
		if ((r < n) && (c < n)) t[ty + k][tx] = as_d2(src[(long)r * n + c]);
	}
	__syncthreads();
	for (int k = 0; k < 32; k += 8) {
		const int c = c0 + ty + k, r = r0 + tx;
		if ((r < n) && (c < n)) {
			const double2 v = t[tx][ty + k];
			dst[(long)c * n + r] = make_float2((float)v.x, (float)v.y);
		}
	}
}

// the lower triangle of a symmetric n x n matrix from its upper triangle (32 x 32 tiles through
// LDS; a workgroup per tile above the diagonal)
__global__ __launch_bounds__(256) void k_mirror_lower(double2 *M, int n)
{
	__shared__ double2 t[32][33];
	const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
	if (c0 < r0) return;
	const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
	for (int k = 0; k < 32; k += 8) {
		const int r = r0 + ty + k, c = c0 + tx;
		if ((r < n) && (c < n)) t[ty + k][tx] = M[(long)r * n + c];
	}
	__syncthreads();
	for (int k = 0; k < 32; k += 8) {
		const int c = c0 + ty + k, r = r0 + tx;
		if ((r < n) && (c < n) && (c > r)) M[(long)c * n + r] = t[tx][ty + k];
	}
}

void mirror_lower(hipStream_t s, double2 *M, int n)
{
	if (n <= 1) return;
	const unsigned nt = (unsigned)((n + 31) / 32);
	hipLaunchKernelGGL(k_mirror_lower, dim3(nt, nt), dim3(256), 0, s, M, n);
}

// ---------------------------------------------------------------- batched dense products

// One system of the batch: CSR stamps and its dense kernel blocks.
struct slot {
	const int *rp;          // n + 1
	const int *ci;
	const double2 *val;
	const double2 *lp;      // nseg x nseg (kernel Lp, scaled by -jw on the fly)
	const double2 *P;       // nc x nc
	double omega;
};
struct slots {
	slot s[FMAX];
};

// Y[:, f] = A_f X[:, f] for f < nact (vectors n x F row-major; columns f >= nact are zeroed),
// in two launches.
//
// k_mvdense: the dense blocks (Lp on the branch-current rows, P on the charge rows).  Both are
// symmetric (the fills evaluate one triangle and mirror it), so B[c][r] serves for B[r][c] and a
// lane can own a row while the wave reads columns: one load instruction takes 64 consecutive
// entries of a stored row (1 KB).  Items are (64-row block of either dense block, column segment
// of cs columns); each lane accumulates its row's sums over the segment for every active system
// (the F entries of X at a column are the same for the whole wave: scalar loads) and writes them
// to part[s][row][f].
//
// k_mvfin: per (row, f): the dense segments summed in increasing s, scaled by -jw on the Lp rows,
// plus the row's sparse stamps (CSR).
template <int F>
__global__ __launch_bounds__(256) void k_mvdense(slots S, int nact, int offS, int nseg, int offQ, int nc, int cs, int nrbS,
                                                 const double2 *__restrict__ X, double2 *__restrict__ part, int n)
{
	constexpr int CPS = (F >= 16) ? 1 : 16 / F;   // columns per step: 16 loads in flight per lane
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
	const int rb = blockIdx.x * 4 + wave, s = blockIdx.y;
	const bool lblk = rb < nrbS;
	const int nb = lblk ? nseg : nc, o = lblk ? offS : offQ;
	const int rlo = (lblk ? rb : rb - nrbS) * 64;
	const int cb = s * cs;
	if ((rlo >= nb) || (cb >= nb)) return;
	const int ce = min(nb, cb + cs);
	const int lr = rlo + lane;
	const bool act = lr < nb;
	const double2 *B[F];
#pragma unroll
	for (int f = 0; f < F; f++) B[f] = (f < nact) ? ((lblk ? S.s[f].lp : S.s[f].P) + (act ? lr : rlo)) : nullptr;
	double ar[F], ai[F];
#pragma unroll
	for (int f = 0; f < F; f++) ar[f] = ai[f] = 0;
	for (int c = cb; c < ce; c += CPS) {
		double2 m[CPS][F];
#pragma unroll
		for (int u = 0; u < CPS; u++)
#pragma unroll
			for (int f = 0; f < F; f++)
				m[u][f] = (act && (f < nact) && (c + u < ce)) ? B[f][(long)(c + u) * nb] : make_double2(0, 0);
#pragma unroll
		for (int u = 0; u < CPS; u++) {
			if (c + u < ce) {
				const double2 *xp = X + (long)(o + c + u) * F;
#pragma unroll
				for (int f = 0; f < F; f++) {
					if (f < nact) {
						const double2 x = xp[f];
						ar[f] = fma(m[u][f].x, x.x, ar[f]);
						ar[f] = fma(-m[u][f].y, x.y, ar[f]);
						ai[f] = fma(m[u][f].x, x.y, ai[f]);
						ai[f] = fma(m[u][f].y, x.x, ai[f]);
					}
				}
			}
		}
	}
	if (act) {
#pragma unroll
		for (int f = 0; f < F; f++) part[((long)s * n + o + lr) * F + f] = make_double2(ar[f], ai[f]);
	}
}

// k_mvsym: the same product reading only the upper triangle of each symmetric block (whole
// 64 x 64 diagonal tiles), half the bytes of k_mvdense.  A workgroup takes row block I of one
// block and a run of MV_JS tiles J >= I (J0 = I + MV_JS blockIdx.x); wave w takes rows
// 16 w .. 16 w + 15 of the row block and lane l column l of each tile, so one buffer load reads
// 1 KB of a stored row.  A tile above the diagonal serves both triangles:
//   * B[r][c] x[c] accumulates per lane and row over the run; the 16 row sums are reduced across
//     the wave once per run (partN[f][run][row]);
//   * B[r][c] x[r] accumulates per lane and column over the wave's 16 rows; the four waves' sums
//     are added in LDS, in wave order, at the end of the run (partT[f][I][column]).
// The next tile's loads are issued before the current tile's arithmetic.  Partial sums are stored
// system-major, so each store is a contiguous run of rows (the system-interleaved layout cost
// 12 % of the product's time in partial-line writes).  Reads past a block's last column wrap into
// the next stored row and meet a zero x; rows past the last one meet a zero x[r] and land in sums
// that are never written; the buffer descriptor's range ends at the end of the matrix, so nothing
// past it is read (such loads return zero).
constexpr int MV_JS = 8;     // tiles per run

__device__ __forceinline__ void cfma(double2 &acc, double2 b, double2 x)
{
	acc.x = fma(b.x, x.x, acc.x);
	acc.x = fma(-b.y, x.y, acc.x);
	acc.y = fma(b.x, x.y, acc.y);
	acc.y = fma(b.y, x.x, acc.y);
}

__device__ __forceinline__ double2 add2(double2 a, double2 b)
{
	return make_double2(a.x + b.x, a.y + b.y);
}

__device__ __forceinline__ double2 shx(double2 v, int m)
{
	return make_double2(__shfl_xor(v.x, m), __shfl_xor(v.y, m));
}

// v[0..2H) held by lane groups ST lanes apart: each step halves the values a lane keeps (by lane
// bit H ST), adding its partner's; the lane ends with index (lane / ST) % 2H in v[0]
template <int H, int ST>
__device__ __forceinline__ void halve_sums(double2 *v, int lane)
{
	if constexpr (H >= 1) {
		constexpr int m = H * ST;
		const bool up = lane & m;
#pragma unroll
		for (int k = 0; k < H; k++) {
			const double2 snd = up ? v[k] : v[k + H], kp = up ? v[k + H] : v[k];
			v[k] = add2(kp, shx(snd, m));
		}
		halve_sums<H / 2, ST>(v, lane);
	}
}

__global__ __launch_bounds__(256) void k_mvsym(slots S, int nact, unsigned am, int F, int offS, int nseg, int offQ, int nc,
                                               int NR, int NI, const double2 *__restrict__ X, double2 *__restrict__ partN,
                                               double2 *__restrict__ partT, int n)
{
	constexpr int RW = 16;
	__shared__ double2 tT[MV_JS][4][64];
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
	const int ntS = (nseg + 63) / 64;
	const int ib = blockIdx.y;
	const bool lblk = ib < ntS;
	const int nb = lblk ? nseg : nc, o = lblk ? offS : offQ;
	const int nt = (nb + 63) / 64;
	const int I = lblk ? ib : ib - ntS;
	if (I >= nt) return;
	const int sI = blockIdx.x;
	const int J0 = I + sI * MV_JS;
	if (J0 >= nt) return;                       // uniform over the workgroup
	const int J1 = min(nt, J0 + MV_JS);
	const int rlo = I * 64, nr = min(64, nb - rlo);
	const int r0 = RW * wave;
	const long mat_bytes = (long)nb * nb * 16;
	for (int f = 0; f < nact; f++) {
		if (!((am >> f) & 1u)) continue;   // a system outside the mask: its sums are not read
		const double2 *Bf = nullptr;
#pragma unroll
		for (int h = 0; h < FMAX; h++)
			if (h == f) Bf = lblk ? S.s[h].lp : S.s[h].P;
		double2 xr[RW];                          // x at the wave's rows (uniform)
#pragma unroll
		for (int i = 0; i < RW; i++) {
			const int r = r0 + i;
			xr[i] = (r < nr) ? X[(long)(o + rlo + r) * F + f] : make_double2(0, 0);
		}
		double2 acc[RW];
#pragma unroll
		for (int i = 0; i < RW; i++) acc[i] = make_double2(0, 0);
		double2 bA[RW], bB[RW], xjA, xjB;
		auto load = [&](int J, double2 *b, double2 &xj) {
			const int c = J * 64 + lane;
			xj = (c < nb) ? X[(long)(o + c) * F + f] : make_double2(0, 0);
			const long tofs = ((long)(rlo + r0) * nb + J * 64) * 16;
			const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
				(void *)((const char *)Bf + tofs), (short)0, (int)max(0L, min(mat_bytes - tofs, 0x7fffffffL)), 0x00020000);
#pragma unroll
			for (int u = 0; u < RW; u++) {
				const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, lane * 16, u * nb * 16, 0);
				b[u] = *(const double2 *)&v;
			}
		};
		auto tile = [&](int J, const double2 *b, double2 xj) {
			const bool diag = (J == I);
			double2 t0 = make_double2(0, 0), t1 = make_double2(0, 0);
#pragma unroll
			for (int u = 0; u < RW; u++) {
				cfma(acc[u], b[u], xj);
				if (!diag) cfma((u & 1) ? t1 : t0, b[u], xr[u]);
			}
			tT[J - J0][wave][lane] = add2(t0, t1);
		};
		load(J0, bA, xjA);
#pragma unroll 1
		for (int J = J0; J < J1; J += 2) {
			if (J + 1 < J1) load(J + 1, bB, xjB);
			tile(J, bA, xjA);
			if (J + 1 < J1) {
				if (J + 2 < J1) load(J + 2, bA, xjA);
				tile(J + 1, bB, xjB);
			}
		}
		// mirror sums of the run's tiles: the four waves' partials added in wave order
		__syncthreads();
		for (int j = wave; j < J1 - J0; j += 4) {
			if (J0 + j == I) continue;
			const double2 v = add2(add2(add2(tT[j][0][lane], tT[j][1][lane]), tT[j][2][lane]), tT[j][3][lane]);
			const int c = (J0 + j) * 64 + lane;
			if (c < nb) partT[((long)f * NI + I) * n + o + c] = v;
		}
		__syncthreads();
		// the 16 row sums over the wave: halved over lane bits 5..2, then the four lanes of a row
		halve_sums<RW / 2, 4>(acc, lane);
		acc[0] = add2(acc[0], shx(acc[0], 2));
		acc[0] = add2(acc[0], shx(acc[0], 1));
		const int row = r0 + (lane >> 2);
		if (((lane & 3) == 0) && (row < nr)) partN[((long)f * NR + sI) * n + o + rlo + row] = acc[0];
	}
}

// per (system, row) of the symmetric product: the row's run sums (partN, increasing run), then
// the mirror contributions of the tiles above it in its column (partT, increasing I), the -jw
// scale on the Lp rows, and the row's sparse stamps
__global__ __launch_bounds__(256) void k_mvfin_sym(slots S, int n, int F, int nact, unsigned am, int offS, int nseg, int offQ,
                                                   int nc, int NR, int NI, const double2 *__restrict__ partN,
                                                   const double2 *__restrict__ partT, const double2 *__restrict__ X,
                                                   double2 *__restrict__ Y)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * F) return;
	const int f = (int)(e / n), row = (int)(e - (long)f * n);
	double2 y = make_double2(0, 0);
	if ((f < nact) && ((am >> f) & 1u)) {
		const slot &sl = S.s[f];
		const bool l = (row >= offS) && (row < offS + nseg);
		const bool q = (row >= offQ) && (row < offQ + nc);
		if (l || q) {
			const int nb = l ? nseg : nc, o = l ? offS : offQ;
			const int nt = (nb + 63) / 64, Ib = (row - o) / 64;
			const int nsg = (nt - Ib + MV_JS - 1) / MV_JS;
			// the partial sums loaded eight at a time, then added one by one in the same order
			double tr = 0, ti = 0;
			auto sum = [&](const double2 *base, int cnt) {
				for (int q0 = 0; q0 < cnt; q0 += 8) {
					double2 p[8];
#pragma unroll
					for (int q = 0; q < 8; q++) p[q] = (q0 + q < cnt) ? base[(long)(q0 + q) * n] : make_double2(0, 0);
#pragma unroll
					for (int q = 0; q < 8; q++)
						if (q0 + q < cnt) {
							tr += p[q].x;
							ti += p[q].y;
						}
				}
			};
			sum(partN + (long)f * NR * n + row, nsg);
			sum(partT + (long)f * NI * n + row, Ib);
			y = l ? make_double2(sl.omega * ti, -sl.omega * tr) : make_double2(tr, ti);   // Lp rows: -jw * sum
		}
		for (int k = sl.rp[row]; k < sl.rp[row + 1]; k++) {
			const double2 p = zmul(sl.val[k], X[(long)sl.ci[k] * F + f]);
			y.x += p.x;
			y.y += p.y;
		}
	}
	Y[(long)row * F + f] = y;
}

__global__ __launch_bounds__(256) void k_mvfin(slots S, int n, int F, int nact, int offS, int nseg, int offQ, int nc, int cs,
                                               const double2 *__restrict__ part, const double2 *__restrict__ X,
                                               double2 *__restrict__ Y)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * F) return;
	const int row = (int)(e / F), f = (int)(e - (long)row * F);
	double2 y = make_double2(0, 0);
	if (f < nact) {
		const slot &sl = S.s[f];
		const bool l = (row >= offS) && (row < offS + nseg);
		const bool q = (row >= offQ) && (row < offQ + nc);
		if (l || q) {
			const int nsg = ((l ? nseg : nc) + cs - 1) / cs;
			double tr = 0, ti = 0;
			for (int sg = 0; sg < nsg; sg++) {
				const double2 p = part[((long)sg * n + row) * F + f];
				tr += p.x;
				ti += p.y;
			}
			y = l ? make_double2(sl.omega * ti, -sl.omega * tr) : make_double2(tr, ti);   // Lp rows: -jw * sum
		}
		for (int k = sl.rp[row]; k < sl.rp[row + 1]; k++) {
			const double2 p = zmul(sl.val[k], X[(long)sl.ci[k] * F + f]);
			y.x += p.x;
			y.y += p.y;
		}
	}
	Y[e] = y;
}

// Y = B - Y (residual after a product), all columns
__global__ void k_resid(const double2 *B, double2 *Y, long cnt)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= cnt) return;
	const double2 b = B[e], y = Y[e];
	Y[e] = make_double2(b.x - y.x, b.y - y.y);
}

// partial[blk][k][f] = sum over the block's rows of conj(V_k[i][f]) W[i][f], k < nk
constexpr int DOT_ROWS = 256;
__global__ __launch_bounds__(256) void k_dots(const double2 *V, long vstride, int nk, const double2 *W, int n, int F,
                                              double2 *partial)
{
	__shared__ double2 red[256];
	const int blk = blockIdx.x;
	const int k = blockIdx.y;
	const int r0 = blk * DOT_ROWS;
	const int t = threadIdx.x;
	const int f = t % F, rr = t / F, rstep = 256 / F;
	double sr = 0, si = 0;
	if (rr < rstep) {
		const double2 *vk = V + (long)k * vstride;
		for (int r = r0 + rr; r < min(n, r0 + DOT_ROWS); r += rstep) {
			const double2 v = vk[(long)r * F + f], w = W[(long)r * F + f];
			sr += (v.x * w.x) + (v.y * w.y);
			si += (v.x * w.y) - (v.y * w.x);
		}
	}
	red[t] = make_double2(sr, si);
	__syncthreads();
	if (t < F) {
		double ar = 0, ai = 0;
		for (int q = t; q < rstep * F; q += F) {
			ar += red[q].x;
			ai += red[q].y;
		}
		partial[((long)blk * gridDim.y + k) * F + t] = make_double2(ar, ai);
	}
}

// h[k][f] = sum over blocks (fixed order)
__global__ void k_dots_final(const double2 *partial, int nblk, int nk, int F, double2 *h)
{
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nk * F) return;
	// the block partials loaded eight at a time, then added one by one in block order
	double sr = 0, si = 0;
	const long st = (long)nk * F;
	for (int b0 = 0; b0 < nblk; b0 += 8) {
		double2 p[8];
#pragma unroll
		for (int q = 0; q < 8; q++) p[q] = (b0 + q < nblk) ? partial[(long)(b0 + q) * st + e] : make_double2(0, 0);
#pragma unroll
		for (int q = 0; q < 8; q++)
			if (b0 + q < nblk) {
				sr += p[q].x;
				si += p[q].y;
			}
	}
	h[e] = make_double2(sr, si);
}

// W[:, f] -= sum_k h[k][f] V_k[:, f]
__global__ void k_orth(const double2 *V, long vstride, int nk, const double2 *h, double2 *W, int n, int F)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * F) return;
	const int f = (int)(e % F);
	double2 w = W[e];
	for (int k = 0; k < nk; k++) {
		const double2 p = zmul(h[k * F + f], V[(long)k * vstride + e]);
		w.x -= p.x;
		w.y -= p.y;
	}
	W[e] = w;
}

// Vout[:, f] = scale[f] * Win[:, f]
__global__ void k_scale(const double2 *Win, double2 *Vout, const double *scale, int n, int F)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * F) return;
	const double s = scale[e % F];
	const double2 w = Win[e];
	Vout[e] = make_double2(s * w.x, s * w.y);
}

// Vout[:, f] = Win[:, f] / sqrt(nn[f].x) (0 where the norm is 0): the next Arnoldi vector from
// the device-side norm, the same arithmetic as 1 / sqrt on the host
__global__ void k_scale_inv_norm(const double2 *Win, double2 *Vout, const double2 *nn, int n, int F)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * F) return;
	const double hn = sqrt(nn[e % F].x);
	const double s = (hn > 0) ? 1.0 / hn : 0.0;
	const double2 w = Win[e];
	Vout[e] = make_double2(s * w.x, s * w.y);
}

// U[:, f] = sum_k y[k][f] V_k[:, f]
__global__ void k_combine(const double2 *V, long vstride, int nk, const double2 *y, double2 *U, int n, int F)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * F) return;
	const int f = (int)(e % F);
	double2 u = make_double2(0, 0);
	for (int k = 0; k < nk; k++) {
		const double2 p = zmul(y[k * F + f], V[(long)k * vstride + e]);
		u.x += p.x;
		u.y += p.y;
	}
	U[e] = u;
}

__global__ void k_add(double2 *X, const double2 *D, long cnt)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= cnt) return;
	X[e] = make_double2(X[e].x + D[e].x, X[e].y + D[e].y);
}

// initial guesses of a batch from the previous batch's solutions of the same port (n x F each):
// X[:, b] = sum_q w[b][q] P[:, col[q]] for b < nb (Lagrange extrapolation in frequency through
// up to three previous points), zero for the unused columns
struct xcoef {
	double w[FMAX][3];
	int col[3];
	int nq;
};
__global__ void k_extrap(const double2 *P, double2 *X, int n, int F, int nb, xcoef tc)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * F) return;
	const int i = (int)(e / F), b = (int)(e - (long)i * F);
	double2 x = make_double2(0, 0);
	if (b < nb) {
		for (int q = 0; q < tc.nq; q++) {
			const double2 a = P[(long)i * F + tc.col[q]];
			x.x += tc.w[b][q] * a.x;
			x.y += tc.w[b][q] * a.y;
		}
	}
	X[e] = x;
}

__global__ void k_zero(double2 *X, long cnt)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e < cnt) X[e] = make_double2(0, 0);
}

__global__ void k_zero_u32(unsigned *x, int cnt)
{
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e < cnt) x[e] = 0;
}

// dst[k][f] = src[perm[k]][f]
__global__ void k_permute(const double2 *src, double2 *dst, const int *perm, int n, int F)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * F) return;
	const int k = (int)(e / F), f = (int)(e - (long)k * F);
	dst[e] = src[(long)perm[k] * F + f];
}

// ---------------------------------------------------------------- preconditioner
// The preconditioner is applied through explicit inverses of the two factors, Linv (unit lower)
// and Uinv (upper), stored in FP32: z = Uinv (Linv (P v)).  Two triangular matrix-vector
// products with F right-hand sides stream each inverse once (half of n^2 FP32 entries each) and
// have no sequential dependence, unlike substitution, whose 64-row steps each wait for the
// previous one.  The inverses are built once per factorisation by recursive blocking (below).

// Y = T X for a triangular FP32 T stored by columns (Tc[c * n + r] = T[r][c]; entries outside
// the triangle are stored zeros) and F FP64 columns X (n x F, row-major).  The triangle is cut
// into items (64-row block, column segment s of cs columns) so every wave streams about the same
// share of T whatever its rows; a wave takes one item and lane l owns row 64 rb + l, so one
// load instruction reads 64 consecutive entries of a column (512 B).  A step covers 16 columns
// (the next step's entries loaded before the current step's products); the step's X entries are
// the same for every lane (scalar loads); FP64 fused multiply-adds.  No cross-lane reduction:
// each item writes its row sums to part[s][row][f] and k_tgsum adds a row's segments in
// increasing s (deterministic).
constexpr int TG_RB = 64;

template <bool LOWER, int F, bool FULL = false>
__global__ __launch_bounds__(256) void k_tgemv(const float2 *__restrict__ Tc, int n, int cs,
                                               const double2 *__restrict__ X, double2 *__restrict__ part)
{
	// the wave index through readfirstlane: item bounds, and so the X addresses, are then known
	// to be uniform and X is read with scalar loads
	const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
	const int rlo = (blockIdx.x * 4 + wave) * TG_RB, s = blockIdx.y;
	if (rlo >= n) return;
	const int rhi = min(n, rlo + TG_RB);
	int cb, ce;
	if (FULL) {
		// a full matrix (the Schur preconditioner's C, symmetric, so its rows are its columns)
		cb = s * cs;
		ce = min(n, cb + cs);
	}
	else if (LOWER) {
		cb = s * cs;
		ce = min(rhi, cb + cs);
	}
	else {
		cb = max(s * cs, rlo);
		ce = min(n, (s + 1) * cs);
	}
	if (cb >= ce) return;   // item outside the triangle
	const int row = rlo + lane;
	const bool act = row < n;
	const float2 *tr = Tc + (act ? row : rlo);
	double ar[F], ai[F];
#pragma unroll
	for (int f = 0; f < F; f++) ar[f] = ai[f] = 0;
	auto load = [&](int c, float2 (&d)[16]) {
#pragma unroll
		for (int u = 0; u < 16; u++) d[u] = act ? tr[(long)(c + u) * n] : make_float2(0, 0);
	};
	auto macc = [&](int c, const float2 (&d)[16], int cnt) {
#pragma unroll
		for (int j = 0; j < 16; j++) {
			if (j < cnt) {
				const double mre = d[j].x, mim = d[j].y;
				const double2 *xp = X + (long)(c + j) * F;
#pragma unroll
				for (int f = 0; f < F; f++) {
					const double2 x = xp[f];
					ar[f] = fma(mre, x.x, ar[f]);
					ar[f] = fma(-mim, x.y, ar[f]);
					ai[f] = fma(mre, x.y, ai[f]);
					ai[f] = fma(mim, x.x, ai[f]);
				}
			}
		}
	};
	const int nstep = (ce - cb) / 16;
	float2 m[16], mn[16];
	if (nstep > 0) load(cb, m);
	for (int t = 0; t < nstep; t++) {
		const int c = cb + 16 * t;
		if (t + 1 < nstep) load(c + 16, mn);
		macc(c, m, 16);
#pragma unroll
		for (int j = 0; j < 16; j++) m[j] = mn[j];
	}
	const int ct = cb + 16 * nstep;
	if (ct < ce) {
		const int cnt = ce - ct;
#pragma unroll
		for (int u = 0; u < 16; u++) m[u] = (act && (u < cnt)) ? tr[(long)(ct + u) * n] : make_float2(0, 0);
		macc(ct, m, cnt);
	}
	if (act) {
#pragma unroll
		for (int f = 0; f < F; f++) part[((long)s * n + row) * F + f] = make_double2(ar[f], ai[f]);
	}
}

// Y[row][f] = sum over the row's segments of part[s][row][f] (the items k_tgemv ran for the
// row's 64-row block), in increasing s
template <bool LOWER, bool FULL = false>
__global__ void k_tgsum(const double2 *__restrict__ part, int n, int F, int cs, double2 *__restrict__ Y)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * F) return;
	const int row = (int)(e / F);
	const int rlo = (row / TG_RB) * TG_RB, rhi = min(n, rlo + TG_RB);
	const int s0 = (LOWER || FULL) ? 0 : rlo / cs;
	const int s1 = FULL ? (n - 1) / cs : LOWER ? (rhi - 1) / cs : (n - 1) / cs;
	const long stride = (long)n * F;
	double2 a = part[(long)s0 * stride + e];
	// the segments loaded eight at a time, then added one by one in increasing s
	for (int q0 = s0 + 1; q0 <= s1; q0 += 8) {
		double2 p[8];
#pragma unroll
		for (int q = 0; q < 8; q++) p[q] = (q0 + q <= s1) ? part[(long)(q0 + q) * stride + e] : make_double2(0, 0);
#pragma unroll
		for (int q = 0; q < 8; q++)
			if (q0 + q <= s1) {
				a.x += p[q].x;
				a.y += p[q].y;
			}
	}
	Y[e] = a;
}

// column segment of the triangular products (PEEC_TG_CS, a multiple of 64 in [64, 8192]; 512 by
// default)
int tg_cs()
{
	static int cs = -1;
	if (cs < 0) {
		const char *e = getenv("PEEC_TG_CS");
		const int v = e ? atoi(e) : 512;
		cs = ((v >= 64) && (v <= 8192) && (v % 64 == 0)) ? v : 512;
	}
	return cs;
}

// column segment of the dense blocks of A x (PEEC_MV_CS, a multiple of 64 in [64, 8192]; 256 by
// default: about 5000 items on the 64 x 64 plate, enough to fill every SIMD)
int mv_cs()
{
	static int cs = -1;
	if (cs < 0) {
		const char *e = getenv("PEEC_MV_CS");
		const int v = e ? atoi(e) : 256;
		cs = ((v >= 64) && (v <= 8192) && (v % 64 == 0)) ? v : 256;
	}
	return cs;
}

// dense blocks of A x read as symmetric, upper triangle only (PEEC_MV_SYM=0: both triangles,
// k_mvdense)
bool mv_sym()
{
	const char *e = getenv("PEEC_MV_SYM");   // read per product (a few per GMRES iteration)
	return !(e && (e[0] == '0'));
}

// doubles of partial sums the products need for order n and F columns
size_t tg_part_count(int n, int F)
{
	return (size_t)((n + tg_cs() - 1) / tg_cs()) * n * F;
}

template <bool LOWER, int F, bool FULL>
void tgemv_f(hipStream_t s, const float2 *Mt, int n, const double2 *X, double2 *Y, double2 *part)
{
	const int cs = tg_cs();
	const dim3 grid((unsigned)((n + 4 * TG_RB - 1) / (4 * TG_RB)), (unsigned)((n + cs - 1) / cs));
	hipLaunchKernelGGL((k_tgemv<LOWER, F, FULL>), grid, dim3(256), 0, s, Mt, n, cs, X, part);
	hipLaunchKernelGGL((k_tgsum<LOWER, FULL>), dim3((unsigned)(((long)n * F + 255) / 256)), dim3(256), 0, s,
	                   (const double2 *)part, n, F, cs, Y);
}

// Y = T X for a triangle (LOWER or upper) or, with FULL, a full matrix stored by columns
template <bool LOWER, bool FULL = false>
void tgemv(hipStream_t s, const float2 *Mt, int n, const double2 *X, double2 *Y, int F, double2 *part)
{
	switch (F) {
	case 1: tgemv_f<LOWER, 1, FULL>(s, Mt, n, X, Y, part); break;
	case 2: tgemv_f<LOWER, 2, FULL>(s, Mt, n, X, Y, part); break;
	case 3: tgemv_f<LOWER, 3, FULL>(s, Mt, n, X, Y, part); break;
	case 4: tgemv_f<LOWER, 4, FULL>(s, Mt, n, X, Y, part); break;
	case 5: tgemv_f<LOWER, 5, FULL>(s, Mt, n, X, Y, part); break;
	case 6: tgemv_f<LOWER, 6, FULL>(s, Mt, n, X, Y, part); break;
	case 7: tgemv_f<LOWER, 7, FULL>(s, Mt, n, X, Y, part); break;
	default: tgemv_f<LOWER, 8, FULL>(s, Mt, n, X, Y, part); break;
	}
}

template <typename T>
T *dal(size_t n)
{
	T *p = nullptr;
	HIPCHK(hipMalloc((void **)&p, (n ? n : 1) * sizeof(T)));
	return p;
}

unsigned blocks(long n, int t = 256)
{
	return (unsigned)((n + t - 1) / t);
}

// ---------------------------------------------------------------- host-side context

struct slotbuf {
	double2 *lp = nullptr, *P = nullptr;
	int *rp = nullptr, *ci = nullptr, *ri = nullptr;
	double2 *val = nullptr;
	int cap = 0, nnz = 0;
	double omega = 0;
	bool upper = false;   // lp and P hold their upper triangles and diagonal 64 x 64 blocks only
};

struct fast {
	int n = 0, nc = 0, F = 0, nseg = 0;
	int ns = 0;               // order of the factored Schur complement (the plain system)
	slotbuf sl[FMAX];
	double2 *luw = nullptr;   // FP64 factorisation workspace
	double2 *inv = nullptr;   // FP64 inverse of one factor (built, then demoted)
	double2 *scr = nullptr;   // GEMM scratch of the recursive inversion
	double2 *scr2 = nullptr;  // the same for the second triangle (inverted concurrently)
	hipStream_t s2 = nullptr; // second stream of the inversion (default priority)
	hipEvent_t ev_f = nullptr, ev_u = nullptr;
	float2 *ML = nullptr, *MU = nullptr;   // FP32 Linv, Uinv, stored by columns
	double2 *dinvL = nullptr, *dinvU = nullptr;
	int *piv = nullptr, *perm = nullptr;
	double fref = -1;
	bool stale = false;       // refactor before the next batch
	int its_ref = -1;         // lock-step iterations of the first batch after a factorisation
	int lock_ref = -1;        // the same summed over ports (cost rule)
	double lu_ms = 0;         // wall time of the last factorisation and inversion
	double excess_ms = 0;     // GMRES time beyond the best batch's rate since the last factorisation
	std::vector<std::pair<int, double>> hist;   // per batch since the factorisation: lock-step
	                                            // iterations, ms per iteration
	int served = 0;           // points solved with the current factorisation
	int span = 0;             // points the previous factorisation served (0: none yet)
	std::vector<double2 *> xw;   // per port: the previous batch's solutions (n x F)
	int xw_nb = 0;            // points in the previous batch (0: none)
	double xw_f[FMAX] = {};   // their frequencies
	slotbuf ahead;            // system at a point ahead of the current batch (stamps; its Lp and P
	                          // live in the inverse workspace until the factorisation)
	double2 *V = nullptr;     // (MKRYLOV + 1) x n x F Arnoldi basis
	double2 *Z = nullptr;     // MKRYLOV x n x F preconditioned directions
	double2 *W = nullptr, *Xv = nullptr, *Bv = nullptr, *R = nullptr, *T1 = nullptr;
	double2 *tgp = nullptr;   // segment partial sums of the triangular products
	double2 *mvp = nullptr;   // segment partial sums of the dense blocks of A x
	double2 *mvt = nullptr;   // mirror partial sums of the symmetric product (per row block and column)
	double2 *partial = nullptr, *h = nullptr, *yd = nullptr;
	double2 *h2 = nullptr, *nn = nullptr;   // second Gram-Schmidt pass and norms, device side
	double2 *hp = nullptr;    // pinned host copy of h, h2, nn (one transfer per Arnoldi step)
	double *scale = nullptr;
	// the Schur preconditioner's charge part: C = P(f_ref)^-1 (FP64, then FP32 by columns),
	// the node row of each cell and the cell of each node row, two n_c x F work vectors
	double2 *Pw = nullptr, *Cd = nullptr, *cwork = nullptr;
	float2 *Cc = nullptr;
	int *ppiv = nullptr, *pperm = nullptr, *rowcell = nullptr, *cellrow = nullptr;
	double2 *Wc = nullptr, *Wt = nullptr;
	double omega_ref = 0;
};

// both FP32 inverses from the FP64 factors in c.luw, stored by columns
void build_inverses(fast &c, hipStream_t s)
{
	const int n = c.ns;
	for (int pass = 0; pass < 2; pass++) {
		const bool lower = (pass == 0);
		dlu_tri_inverse(s, n, (const d_complex_t *)c.luw, n, lower ? 1 : 0, (d_complex_t *)c.inv, (d_complex_t *)c.scr,
		                (d_complex_t *)(lower ? c.dinvL : c.dinvU));
		const unsigned nt = (unsigned)((n + 31) / 32);
		hipLaunchKernelGGL(k_transpose32<double2>, dim3(nt, nt), dim3(256), 0, s, (const double2 *)c.inv,
		                   lower ? c.ML : c.MU, n);
	}
}

// z = M^-1 v, v and z n x F (n = ns + nc): block elimination through the Schur complement,
//   w = C v2,  x1 = S^-1 (v1 - jw E^T w) = Uinv (Linv (P (...))),  x2 = C (v2 + E x1)
// (without capacitive cells, nc = 0: z = Uinv (Linv (P v)))
void precond(fast &c, hipStream_t s, const double2 *v, double2 *z, int F)
{
	const int ns = c.ns, nc = c.nc;
	const double2 *v2 = v + (long)ns * F;
	double2 *x2 = z + (long)ns * F;
	if (nc > 0) tgemv<false, true>(s, (const float2 *)c.Cc, nc, v2, c.Wc, F, c.tgp);
	hipLaunchKernelGGL(k_permute_schur, dim3(blocks((long)ns * F)), dim3(256), 0, s, v, (const double2 *)c.Wc, c.perm,
	                   (nc > 0) ? (const int *)c.rowcell : nullptr, c.omega_ref, z, ns, F);
	tgemv<true>(s, (const float2 *)c.ML, ns, (const double2 *)z, c.T1, F, c.tgp);
	tgemv<false>(s, (const float2 *)c.MU, ns, (const double2 *)c.T1, z, F, c.tgp);
	if (nc > 0) {
		hipLaunchKernelGGL(k_charge_rhs, dim3(blocks((long)nc * F)), dim3(256), 0, s, v2, (const double2 *)z, c.cellrow, c.Wt,
		                   nc, F);
		tgemv<false, true>(s, (const float2 *)c.Cc, nc, (const double2 *)c.Wt, x2, F, c.tgp);
	}
}

int g_offS = 0, g_offQ = 0;

// kernel timing of the GMRES products (HIP events on the sweep's stream), read back when the
// stream is next synchronised
struct mvtimer {
	bool on = false;
	std::vector<hipEvent_t> ev;
	int used = 0;
	double bytes = 0;
	void reset()
	{
		used = 0;
		bytes = 0;
	}
} g_mvt;

// y = A_f x for the first nact systems of the batch (columns f >= nact are set to zero).  With
// the symmetric product, systems outside the mask am are skipped and their columns set to zero
// as well (their matrices are not read); the two-triangle product ignores the mask.
void matvec(fast &c, hipStream_t s, const double2 *x, double2 *y, int F, int nact, unsigned am = ~0u)
{
	if (g_mvt.on) {
		if ((int)g_mvt.ev.size() < g_mvt.used + 2) {
			// timing-only events: no system-scope fence (its cache writeback would be charged to
			// the kernel between them)
			hipEvent_t e0, e1;
			HIPCHK(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
			HIPCHK(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
			g_mvt.ev.push_back(e0);
			g_mvt.ev.push_back(e1);
		}
		HIPCHK(hipEventRecord(g_mvt.ev[g_mvt.used], s));
	}
	slots S;
	for (int f = 0; f < FMAX; f++) {
		const slotbuf &b = c.sl[(f < nact) ? f : 0];
		S.s[f] = slot{b.rp, b.ci, b.val, b.lp, b.P, b.omega};
	}
	if (mv_sym()) {
		const int ntS = (c.nseg + 63) / 64, ntQ = (c.nc + 63) / 64, ntm = std::max(ntS, ntQ);
		const int NR = (ntm + MV_JS - 1) / MV_JS;   // runs per row block at most (partN rows per system)
		if ((ntS + ntQ) > 0)
			hipLaunchKernelGGL(k_mvsym, dim3((unsigned)NR, (unsigned)(ntS + ntQ)), dim3(256), 0, s, S, nact, am, F, g_offS,
			                   c.nseg, g_offQ, c.nc, NR, ntm, x, c.mvp, c.mvt, c.n);
		if (g_mvt.on) {
			// the upper triangle's tiles of the nact systems (16 B per entry), x once, the run and
			// mirror sums written
			HIPCHK(hipEventRecord(g_mvt.ev[g_mvt.used + 1], s));
			g_mvt.used += 2;
			double ent = 0, sums = 0;
			for (int blk = 0; blk < 2; blk++) {
				const int nb = blk ? c.nc : c.nseg, nt = (nb + 63) / 64;
				for (int I = 0; I < nt; I++) {
					const int rows = std::min(64, nb - 64 * I);
					ent += (double)rows * (nb - 64 * I);
					sums += (double)((nt - I + MV_JS - 1) / MV_JS) * rows + (double)(nb - 64 * (I + 1) > 0 ? nb - 64 * (I + 1) : 0);
				}
			}
			int nin = 0;
			for (int f = 0; f < nact; f++) nin += (am >> f) & 1u;
			g_mvt.bytes += 16.0 * nin * ent + 16.0 * c.n * F + 16.0 * nin * sums;
		}
		hipLaunchKernelGGL(k_mvfin_sym, dim3(blocks((long)c.n * F)), dim3(256), 0, s, S, c.n, F, nact, am, g_offS, c.nseg,
		                   g_offQ, c.nc, NR, ntm, (const double2 *)c.mvp, (const double2 *)c.mvt, x, y);
		return;
	}
	const int cs = mv_cs();
	const int nrbS = (c.nseg + 63) / 64, nrbQ = (c.nc + 63) / 64;
	const dim3 grid((unsigned)((nrbS + nrbQ + 3) / 4), (unsigned)((std::max(c.nseg, c.nc) + cs - 1) / cs));
	if ((nrbS + nrbQ) > 0) {
		switch (F) {
		case 1: hipLaunchKernelGGL(k_mvdense<1>, grid, dim3(256), 0, s, S, nact, g_offS, c.nseg, g_offQ, c.nc, cs, nrbS, x, c.mvp, c.n); break;
		case 2: hipLaunchKernelGGL(k_mvdense<2>, grid, dim3(256), 0, s, S, nact, g_offS, c.nseg, g_offQ, c.nc, cs, nrbS, x, c.mvp, c.n); break;
		case 3: hipLaunchKernelGGL(k_mvdense<3>, grid, dim3(256), 0, s, S, nact, g_offS, c.nseg, g_offQ, c.nc, cs, nrbS, x, c.mvp, c.n); break;
		case 4: hipLaunchKernelGGL(k_mvdense<4>, grid, dim3(256), 0, s, S, nact, g_offS, c.nseg, g_offQ, c.nc, cs, nrbS, x, c.mvp, c.n); break;
		case 5: hipLaunchKernelGGL(k_mvdense<5>, grid, dim3(256), 0, s, S, nact, g_offS, c.nseg, g_offQ, c.nc, cs, nrbS, x, c.mvp, c.n); break;
		case 6: hipLaunchKernelGGL(k_mvdense<6>, grid, dim3(256), 0, s, S, nact, g_offS, c.nseg, g_offQ, c.nc, cs, nrbS, x, c.mvp, c.n); break;
		case 7: hipLaunchKernelGGL(k_mvdense<7>, grid, dim3(256), 0, s, S, nact, g_offS, c.nseg, g_offQ, c.nc, cs, nrbS, x, c.mvp, c.n); break;
		default: hipLaunchKernelGGL(k_mvdense<8>, grid, dim3(256), 0, s, S, nact, g_offS, c.nseg, g_offQ, c.nc, cs, nrbS, x, c.mvp, c.n); break;
		}
	}
	if (g_mvt.on) {
		// k_mvdense alone: the dense blocks of the nact systems (16 B per entry), x once, and the
		// segment sums written
		HIPCHK(hipEventRecord(g_mvt.ev[g_mvt.used + 1], s));
		g_mvt.used += 2;
		const double nsum = (double)((c.nseg + cs - 1) / cs) * c.nseg + (double)((c.nc + cs - 1) / cs) * c.nc;
		g_mvt.bytes += 16.0 * nact * ((double)c.nseg * c.nseg + (double)c.nc * c.nc) + 16.0 * c.n * F + 16.0 * nsum * F;
	}
	hipLaunchKernelGGL(k_mvfin, dim3(blocks((long)c.n * F)), dim3(256), 0, s, S, c.n, F, nact, g_offS, c.nseg, g_offQ, c.nc, cs,
	                   (const double2 *)c.mvp, x, y);
}

// sum the recorded product timings into st (stream already synchronised by the caller)
void mv_collect(pgpu_stats_t *st)
{
	for (int k = 0; k + 1 < g_mvt.used; k += 2) {
		float ms = 0;
		HIPCHK(hipEventElapsedTime(&ms, g_mvt.ev[k], g_mvt.ev[k + 1]));
		st->mv_ms += ms;
		st->mv_launches++;
	}
	st->mv_bytes += g_mvt.bytes;
	g_mvt.reset();
}

// h[k][f] = V_k^H W for k < nk, copied to host
void dots(fast &c, hipStream_t s, const double2 *V, long vstride, int nk, const double2 *W, int F, std::vector<double2> &hh)
{
	const int nb = (c.n + DOT_ROWS - 1) / DOT_ROWS;
	hipLaunchKernelGGL(k_dots, dim3(nb, nk), dim3(256), 0, s, V, vstride, nk, W, c.n, F, c.partial);
	hipLaunchKernelGGL(k_dots_final, dim3(blocks((long)nk * F)), dim3(256), 0, s, c.partial, nb, nk, F, c.h);
	hh.resize((size_t)nk * F);
	HIPCHK(hipMemcpyAsync(hh.data(), c.h, (size_t)nk * F * sizeof(double2), hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
}

// h[k][f] = V_k^H W for k < nk into the device buffer hout (no transfer)
void dots_dev(fast &c, hipStream_t s, const double2 *V, long vstride, int nk, const double2 *W, int F, double2 *hout)
{
	const int nb = (c.n + DOT_ROWS - 1) / DOT_ROWS;
	hipLaunchKernelGGL(k_dots, dim3(nb, nk), dim3(256), 0, s, V, vstride, nk, W, c.n, F, c.partial);
	hipLaunchKernelGGL(k_dots_final, dim3(blocks((long)nk * F)), dim3(256), 0, s, c.partial, nb, nk, F, hout);
}

// batched restarted flexible GMRES on F systems A_f x_f = b_f (b in c.Bv, x in c.Xv); returns
// per-system iteration counts, -1 for a system that did not converge
// warm: c.Xv already holds the initial guesses, else they are zero
// GMRES restart length (PEEC_GMRES_M, 10 .. MKRYLOV; 40 by default); the cycles before a
// refactorisation make up MITS iterations
int gmres_m()
{
	static int m = -1;
	if (m < 0) {
		const char *e = getenv("PEEC_GMRES_M");
		const int v = e ? atoi(e) : 40;
		m = ((v >= 10) && (v <= MKRYLOV)) ? v : 40;
	}
	return m;
}

std::vector<int> gmres_batch(fast &c, hipStream_t s, int F, int nactive, bool warm)
{
	const int mk = gmres_m(), ncyc = std::max(1, MITS / mk);
	const int n = c.n;
	const long nf = (long)n * F;
	const long vstride = nf;
	std::vector<int> its(F, 0);
	std::vector<char> done(F, 0);
	for (int f = nactive; f < F; f++) done[f] = 1;   // unused slots of a short batch
	std::vector<double> bnorm(F);
	std::vector<double2> hh;
	dots(c, s, c.Bv, 0, 1, c.Bv, F, hh);
	for (int f = 0; f < F; f++) bnorm[f] = std::sqrt(hh[f].x);
	if (!warm) hipLaunchKernelGGL(k_zero, dim3(blocks(nf)), dim3(256), 0, s, c.Xv, nf);
	// systems whose products are still needed: not done (residual), and not converged in this
	// cycle (Arnoldi steps); the others' columns come out zero
	auto mask = [&](const std::vector<int> *kd) {
		unsigned am = 0;
		for (int f = 0; f < F; f++)
			if (!done[f] && (!kd || ((*kd)[f] < 0))) am |= 1u << f;
		return am;
	};
	for (int cyc = 0; cyc < ncyc; cyc++) {
		// r = b - A x
		matvec(c, s, c.Xv, c.R, F, nactive, mask(nullptr));
		hipLaunchKernelGGL(k_resid, dim3(blocks(nf)), dim3(256), 0, s, c.Bv, c.R, nf);
		dots(c, s, c.R, 0, 1, c.R, F, hh);
		std::vector<double> beta(F);
		bool all = true;
		for (int f = 0; f < F; f++) {
			beta[f] = std::sqrt(hh[f].x);
			if (!done[f] && ((beta[f] <= TOL * bnorm[f]) || (bnorm[f] == 0))) done[f] = 1;
			all = all && done[f];
		}
		if (all) break;
		std::vector<double> sc(F);
		for (int f = 0; f < F; f++) sc[f] = (beta[f] > 0) ? 1.0 / beta[f] : 0.0;
		HIPCHK(hipMemcpyAsync(c.scale, sc.data(), F * sizeof(double), hipMemcpyHostToDevice, s));
		hipLaunchKernelGGL(k_scale, dim3(blocks(nf)), dim3(256), 0, s, c.R, c.V, c.scale, n, F);
		// per-system Hessenberg, Givens, rhs
		std::vector<std::vector<zc>> H(F, std::vector<zc>((size_t)(MKRYLOV + 1) * MKRYLOV));
		std::vector<std::vector<zc>> cs(F, std::vector<zc>(MKRYLOV)), sn(F, std::vector<zc>(MKRYLOV));
		std::vector<std::vector<zc>> g(F, std::vector<zc>(MKRYLOV + 1));
		std::vector<int> kdone(F, -1);
		for (int f = 0; f < F; f++) g[f][0] = beta[f];
		int j;
		for (j = 0; j < mk; j++) {
			double2 *vj = c.V + (long)j * vstride;
			double2 *zj = c.Z + (long)j * vstride;
			precond(c, s, vj, zj, F);
			matvec(c, s, zj, c.W, F, nactive, mask(&kdone));
			// classical Gram-Schmidt, two passes, and the new vector's norm, all on the device;
			// the coefficients reach the host in one transfer for the Hessenberg column
			dots_dev(c, s, c.V, vstride, j + 1, c.W, F, c.h);
			hipLaunchKernelGGL(k_orth, dim3(blocks(nf)), dim3(256), 0, s, c.V, vstride, j + 1, c.h, c.W, n, F);
			dots_dev(c, s, c.V, vstride, j + 1, c.W, F, c.h2);
			hipLaunchKernelGGL(k_orth, dim3(blocks(nf)), dim3(256), 0, s, c.V, vstride, j + 1, c.h2, c.W, n, F);
			dots_dev(c, s, c.W, 0, 1, c.W, F, c.nn);
			hipLaunchKernelGGL(k_scale_inv_norm, dim3(blocks(nf)), dim3(256), 0, s, c.W, c.V + (long)(j + 1) * vstride, c.nn,
			                   n, F);
			const size_t nh = (size_t)(j + 1) * F;
			const double2 *h1 = c.hp, *h2 = c.hp + nh, *nn = c.hp + 2 * nh;
			HIPCHK(hipMemcpyAsync(c.hp, c.h, nh * sizeof(double2), hipMemcpyDeviceToHost, s));
			HIPCHK(hipMemcpyAsync(c.hp + nh, c.h2, nh * sizeof(double2), hipMemcpyDeviceToHost, s));
			HIPCHK(hipMemcpyAsync(c.hp + 2 * nh, c.nn, (size_t)F * sizeof(double2), hipMemcpyDeviceToHost, s));
			HIPCHK(hipStreamSynchronize(s));
			bool all_conv = true;
			for (int f = 0; f < F; f++) {
				const double hn = std::sqrt(nn[f].x);
				if (done[f] || (kdone[f] >= 0)) continue;
				zc *Hc = &H[f][(size_t)j * (MKRYLOV + 1)];
				for (int i = 0; i <= j; i++) Hc[i] = zc(h1[(size_t)i * F + f].x + h2[(size_t)i * F + f].x, h1[(size_t)i * F + f].y + h2[(size_t)i * F + f].y);
				Hc[j + 1] = hn;
				for (int i = 0; i < j; i++) {
					const zc a = Hc[i], b = Hc[i + 1];
					Hc[i] = std::conj(cs[f][i]) * a + std::conj(sn[f][i]) * b;
					Hc[i + 1] = cs[f][i] * b - sn[f][i] * a;
				}
				const double rho = std::sqrt(std::norm(Hc[j]) + std::norm(Hc[j + 1]));
				if (rho <= 0) {
					kdone[f] = j;
					continue;
				}
				cs[f][j] = Hc[j] / rho;
				sn[f][j] = Hc[j + 1] / rho;
				Hc[j] = rho;
				Hc[j + 1] = 0;
				g[f][j + 1] = -sn[f][j] * g[f][j];
				g[f][j] = std::conj(cs[f][j]) * g[f][j];
				its[f]++;
				if (std::abs(g[f][j + 1]) <= TOL * bnorm[f]) kdone[f] = j + 1;
			}
			for (int f = 0; f < F; f++) all_conv = all_conv && (done[f] || (kdone[f] >= 0));
			if (all_conv) {
				j++;
				break;
			}
		}
		// y = H^-1 g per system; x += Z y
		const int kmax = std::min(j, mk);
		std::vector<double2> yv((size_t)kmax * F, make_double2(0, 0));
		for (int f = 0; f < F; f++) {
			if (done[f]) continue;
			const int k = (kdone[f] >= 0) ? kdone[f] : kmax;
			std::vector<zc> y(k);
			for (int i = k - 1; i >= 0; i--) {
				zc acc = g[f][i];
				for (int l = i + 1; l < k; l++) acc -= H[f][(size_t)l * (MKRYLOV + 1) + i] * y[l];
				y[i] = acc / H[f][(size_t)i * (MKRYLOV + 1) + i];
			}
			for (int i = 0; i < k; i++) yv[(size_t)i * F + f] = make_double2(y[i].real(), y[i].imag());
		}
		if (kmax > 0) {
			HIPCHK(hipMemcpyAsync(c.yd, yv.data(), yv.size() * sizeof(double2), hipMemcpyHostToDevice, s));
			hipLaunchKernelGGL(k_combine, dim3(blocks(nf)), dim3(256), 0, s, c.Z, vstride, kmax, c.yd, c.W, n, F);
			hipLaunchKernelGGL(k_add, dim3(blocks(nf)), dim3(256), 0, s, c.Xv, c.W, nf);
		}
	}
	// final true-residual check
	matvec(c, s, c.Xv, c.R, F, nactive);
	hipLaunchKernelGGL(k_resid, dim3(blocks(nf)), dim3(256), 0, s, c.Bv, c.R, nf);
	dots(c, s, c.R, 0, 1, c.R, F, hh);
	for (int f = 0; f < F; f++)
		if (std::sqrt(hh[f].x) > TOL * bnorm[f] * 1.0000001) its[f] = -1;
	return its;
}

fast g_fast;

void free_fast(fast &c)
{
	for (int f = 0; f < FMAX; f++) {
		slotbuf &b = c.sl[f];
		void *bb[] = {b.lp, b.P, b.rp, b.ci, b.ri, b.val};
		for (void *p : bb)
			if (p) HIPCHK(hipFree(p));
	}
	for (double2 *p : c.xw)
		if (p) HIPCHK(hipFree(p));
	{
		void *bb[] = {c.ahead.rp, c.ahead.ci, c.ahead.ri, c.ahead.val};
		for (void *p : bb)
			if (p) HIPCHK(hipFree(p));
	}
	void *bufs[] = {c.luw, c.inv, c.scr, c.ML, c.MU, c.dinvL, c.dinvU, c.piv, c.perm, c.V, c.Z, c.W, c.Xv, c.Bv, c.R, c.T1, c.tgp, c.mvp, c.mvt, c.partial, c.h, c.yd, c.scale};
	for (void *b : bufs)
		if (b) HIPCHK(hipFree(b));
	if (c.scr2) HIPCHK(hipFree(c.scr2));
	if (c.s2) HIPCHK(hipStreamDestroy(c.s2));
	if (c.ev_f) HIPCHK(hipEventDestroy(c.ev_f));
	if (c.ev_u) HIPCHK(hipEventDestroy(c.ev_u));
	if (c.h2) HIPCHK(hipFree(c.h2));
	if (c.nn) HIPCHK(hipFree(c.nn));
	if (c.hp) HIPCHK(hipHostFree(c.hp));
	{
		void *bb[] = {c.Pw, c.Cd, c.cwork, c.Cc, c.ppiv, c.pperm, c.rowcell, c.cellrow, c.Wc, c.Wt};
		for (void *p : bb)
			if (p) HIPCHK(hipFree(p));
	}
	c = fast();
}

// the cached preconditioner and its buffers live for the process: released at exit
void fast_release_at_exit()
{
	free_fast(g_fast);
}

// stamps of one slot -> device triplets sorted by row (stable) + CSR row pointer
void upload_stamps(slotbuf &b, const stamps_t &sp, int n, hipStream_t s)
{
	std::vector<int> cnt(n + 1, 0);
	for (int e = 0; e < sp.n; e++) cnt[sp.ri[e] + 1]++;
	for (int i = 0; i < n; i++) cnt[i + 1] += cnt[i];
	std::vector<int> pos(cnt.begin(), cnt.end() - 1), ri(sp.n), ci(sp.n);
	std::vector<double2> val(sp.n);
	for (int e = 0; e < sp.n; e++) {
		const int k = pos[sp.ri[e]]++;
		ri[k] = sp.ri[e];
		ci[k] = sp.ci[e];
		val[k] = make_double2(sp.val[e].r, sp.val[e].i);
	}
	if (sp.n > b.cap) {
		if (b.ci) {
			HIPCHK(hipFree(b.ci));
			HIPCHK(hipFree(b.ri));
			HIPCHK(hipFree(b.val));
		}
		b.cap = sp.n;
		b.ci = dal<int>(sp.n);
		b.ri = dal<int>(sp.n);
		b.val = dal<double2>(sp.n);
	}
	b.nnz = sp.n;
	HIPCHK(hipMemcpyAsync(b.rp, cnt.data(), (size_t)(n + 1) * sizeof(int), hipMemcpyHostToDevice, s));
	if (sp.n > 0) {
		HIPCHK(hipMemcpyAsync(b.ri, ri.data(), (size_t)sp.n * sizeof(int), hipMemcpyHostToDevice, s));
		HIPCHK(hipMemcpyAsync(b.ci, ci.data(), (size_t)sp.n * sizeof(int), hipMemcpyHostToDevice, s));
		HIPCHK(hipMemcpyAsync(b.val, val.data(), (size_t)sp.n * sizeof(double2), hipMemcpyHostToDevice, s));
	}
	HIPCHK(hipStreamSynchronize(s));
}

}  // namespace

// helpers of the plain runtime (gpu_runtime.hip)
extern "C" {
void *pgpu_stream(pgpu_t *g);
const double2 *pgpu_lp_dev(pgpu_t *g);
const double2 *pgpu_p_dev(pgpu_t *g);
void mna_stamps_charge(const peec_t *p, double f, stamps_t *st);
}

// PEEC_TRACE=1 prints the iteration counts of every batch to stderr
static const bool g_trace = getenv("PEEC_TRACE") && getenv("PEEC_TRACE")[0] == '1';

// initial guesses extrapolated from the previous batch through its last PEEC_WARM points (1..3;
// 3, quadratic, by default; 0 starts every batch from zero)
static int warm_start()
{
	static int w = -1;
	if (w < 0) {
		const char *e = getenv("PEEC_WARM");
		w = e ? std::min(3, std::max(0, atoi(e))) : 3;
	}
	return w;
}

// the cost rule counts the batch just solved once more, as the next batch's expected excess
// (PEEC_REFACTOR_LOOKAHEAD=0: only the excess already spent)
static bool refactor_lookahead()
{
	static int on = -1;
	if (on < 0) {
		const char *e = getenv("PEEC_REFACTOR_LOOKAHEAD");
		on = !(e && (e[0] == '0'));
	}
	return on != 0;
}

// refactor rule: PEEC_REFACTOR_ALPHA > 0 prices extra iterations against the factorisation time
// (1 by default, the break-even point: on the full stack 3 factorisations and 15.3 s against 4
// and 16.5 s at 0.75, and 18.2 s at 1.5, measured on one GPU); 0 selects the iteration-ratio
// rule (more than twice the first batch's)
static double refactor_alpha()
{
	static double a = -1;
	if (a < 0) {
		const char *e = getenv("PEEC_REFACTOR_ALPHA");
		a = e ? std::max(0.0, atof(e)) : 1.0;
	}
	return a;
}

// PEEC_MIXED_LU64=1 factors and inverts the preconditioner in FP64 (default: FP32 storage, FP64
// panels, FP32 MFMA updates and inversion GEMMs)
static bool lu64()
{
	const char *e = getenv("PEEC_MIXED_LU64");
	return e && (e[0] == '1');
}

// PEEC_KF=0 in the environment switches the k-factored fill off (element-by-element fills)
static bool pgpu_kf_enabled()
{
	const char *e = getenv("PEEC_KF");
	return !(e && e[0] == '0');
}

// Sweep nfreq points from ifreq0 with the accelerated form.  zmat as pgpu_sweep.
extern "C" int pgpu_sweep_fast(pgpu_t *g, const peec_t *p, int ifreq0, int nfreq, d_complex_t *zmat, pgpu_stats_t *st,
                               int batch)
{
	fast &c = g_fast;
	hipStream_t s = (hipStream_t)pgpu_stream(g);
	const int np = p->nport;
	const int usecap = p->capacitance && (p->ncell > 0);
	const int nc = usecap ? p->ncell : 0;
	const int n = p->nunknown + nc;
	const int offQ = p->nunknown;
	const int ns = p->nunknown;   // order of the factored Schur complement
	const int F = std::max(1, std::min(batch, FMAX));
	if ((c.n != n) || (c.F < F) || (c.nseg != p->nseg) || (c.nc != nc)) {
		free_fast(c);
		c.n = n;
		c.nc = nc;
		c.ns = ns;
		c.F = F;
		c.nseg = p->nseg;
		for (int f = 0; f < F; f++) {
			c.sl[f].lp = dal<double2>((size_t)p->nseg * p->nseg);
			c.sl[f].P = dal<double2>((size_t)nc * nc);
			// zeros once: the batch fill leaves most of the lower triangles as they were, and the
			// symmetric product's last partial tile column multiplies a few of them by zero
			hipLaunchKernelGGL(k_zero, dim3(blocks((long)p->nseg * p->nseg)), dim3(256), 0, s, c.sl[f].lp, (long)p->nseg * p->nseg);
			if (nc > 0) hipLaunchKernelGGL(k_zero, dim3(blocks((long)nc * nc)), dim3(256), 0, s, c.sl[f].P, (long)nc * nc);
			c.sl[f].rp = dal<int>(n + 1);
			hipLaunchKernelGGL(k_zero_u32, dim3(blocks(n + 1, 64)), dim3(64), 0, s, (unsigned *)c.sl[f].rp, n + 1);
		}
		c.ahead.rp = dal<int>(n + 1);
		const int nblk = (ns + TB - 1) / TB;
		c.luw = dal<double2>((size_t)ns * ns);
		// the inverse workspace also holds a look-ahead point's Lp and P until its factorisation
		c.inv = dal<double2>(std::max((size_t)ns * ns, (size_t)p->nseg * p->nseg + (size_t)nc * nc));
		c.scr = dal<double2>((size_t)((ns + 1) / 2 + TB) * ((ns + 1) / 2 + TB));
		c.scr2 = dal<double2>((size_t)((ns + 1) / 2 + TB) * ((ns + 1) / 2 + TB));
		pgpu_at_exit(fast_release_at_exit);
		HIPCHK(hipStreamCreate(&c.s2));
		HIPCHK(hipEventCreateWithFlags(&c.ev_f, hipEventDisableTiming));
		HIPCHK(hipEventCreateWithFlags(&c.ev_u, hipEventDisableTiming));
		c.ML = dal<float2>((size_t)ns * ns);
		c.MU = dal<float2>((size_t)ns * ns);
		c.dinvL = dal<double2>((size_t)nblk * TB * TB);
		c.dinvU = dal<double2>((size_t)nblk * TB * TB);
		c.piv = dal<int>(ns);
		c.perm = dal<int>(ns);
		c.rowcell = dal<int>(ns);
		if (nc > 0) {
			c.Pw = dal<double2>((size_t)nc * nc);
			c.Cd = dal<double2>((size_t)nc * nc);
			c.cwork = dal<double2>((size_t)nc * nc);
			c.Cc = dal<float2>((size_t)nc * nc);
			c.ppiv = dal<int>(nc);
			c.pperm = dal<int>(nc);
			c.cellrow = dal<int>(nc);
			c.Wc = dal<double2>((size_t)nc * FMAX);
			c.Wt = dal<double2>((size_t)nc * FMAX);
		}
		const long nf = (long)n * FMAX;
		c.V = dal<double2>((size_t)(MKRYLOV + 1) * nf);
		c.Z = dal<double2>((size_t)MKRYLOV * nf);
		c.W = dal<double2>(nf);
		c.Xv = dal<double2>(nf);
		c.Bv = dal<double2>(nf);
		c.R = dal<double2>(nf);
		c.T1 = dal<double2>(nf);
		c.tgp = dal<double2>(tg_part_count(n, FMAX));
		{
			const int ntm = (std::max(p->nseg, nc) + 63) / 64;
			const size_t segs = std::max((size_t)((std::max(p->nseg, nc) + mv_cs() - 1) / mv_cs()),
			                             (size_t)((ntm + MV_JS - 1) / MV_JS));
			c.mvp = dal<double2>(segs * n * FMAX);
			c.mvt = dal<double2>((size_t)std::max(1, ntm) * n * FMAX);
		}
		c.partial = dal<double2>((size_t)((n + DOT_ROWS - 1) / DOT_ROWS) * (MKRYLOV + 1) * FMAX);
		c.h = dal<double2>((size_t)(MKRYLOV + 1) * FMAX);
		c.h2 = dal<double2>((size_t)(MKRYLOV + 1) * FMAX);
		c.nn = dal<double2>(FMAX);
		HIPCHK(hipHostMalloc((void **)&c.hp, (size_t)(2 * (MKRYLOV + 1) + 1) * FMAX * sizeof(double2), hipHostMallocDefault));
		c.yd = dal<double2>((size_t)(MKRYLOV + 1) * FMAX);
		c.scale = dal<double>(FMAX);
	}
	g_offS = p->offS;
	g_offQ = offQ;
	// the cell of each node row and the node row of each cell (the incidence E of the charge
	// rows), for the Schur preconditioner; the deck's, so set at every call
	{
		std::vector<int> rc(ns, -1), cr(nc, -1);
		for (int i = 0; i < nc; i++) {
			const int mi = p->nodemap[p->cellid[i]];
			if ((mi >= 0) && (mi < ns)) {
				cr[i] = mi;
				rc[mi] = i;
			}
		}
		HIPCHK(hipMemcpyAsync(c.rowcell, rc.data(), (size_t)ns * sizeof(int), hipMemcpyHostToDevice, s));
		if (nc > 0) HIPCHK(hipMemcpyAsync(c.cellrow, cr.data(), (size_t)nc * sizeof(int), hipMemcpyHostToDevice, s));
		HIPCHK(hipStreamSynchronize(s));
	}
	std::vector<int> i1(np), i2(np);
	for (int j = 0; j < np; j++) {
		i1[j] = p->nodemap[p->port[j].n1];
		i2[j] = p->nodemap[p->port[j].n2];
	}
	std::vector<double2> hx((size_t)n * F);
	int rc = 0;
	// k-factored fill over the band of the deck's sweep (and of this call)
	bool kf = false;
	if (p->retardation && pgpu_kf_enabled()) {
		double lo = 1e300, hi = 0;
		const int ends[4] = {0, (p->nfreq > 0) ? p->nfreq - 1 : 0, ifreq0, ifreq0 + nfreq - 1};
		for (int e : ends) {
			const double fe = freq_at(p, e);
			lo = std::min(lo, fe);
			hi = std::max(hi, fe);
		}
		const auto ta = std::chrono::steady_clock::now();
		kf = (pgpu_kf_setup(g, 2 * PI * lo / C0, 2 * PI * hi / C0) == 0);
		const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta).count();
		if (st) {
			st->kf_setup_ms += ms;
			st->fill_ms += ms;
			st->kf_direct_pairs = kf ? pgpu_kf_direct_pairs(g) : -1;
		}
	}
	c.xw_nb = 0;   // initial guesses only from batches of this call
	for (int f0 = 0; (f0 < nfreq) && !rc; f0 += F) {
		const int nb = std::min(F, nfreq - f0);
		const auto t0 = std::chrono::steady_clock::now();
		// the system of sweep point k (relative to this call) into slot buffers sb (its dense
		// blocks too when blocks is set)
		auto fill_system = [&](slotbuf &sb, int k, bool blocks) {
			const double freq = freq_at(p, ifreq0 + k);
			const double kw = p->retardation ? (2 * PI * freq / C0) : 0;
			sb.omega = 2 * PI * freq;
			if (!blocks) {
				// dense blocks already filled (batch fill)
			}
			else if (kf) {
				sb.upper = false;
				pgpu_kf_fill(g, kw, (d_complex_t *)sb.lp, nc ? (d_complex_t *)sb.P : nullptr);
			}
			else {
				sb.upper = false;
				pgpu_fill_lp(g, kw);
				if (usecap) pgpu_fill_pot(g, kw);
				HIPCHK(hipMemcpyAsync(sb.lp, pgpu_lp_dev(g), (size_t)p->nseg * p->nseg * sizeof(double2), hipMemcpyDeviceToDevice, s));
				if (nc) HIPCHK(hipMemcpyAsync(sb.P, pgpu_p_dev(g), (size_t)nc * nc * sizeof(double2), hipMemcpyDeviceToDevice, s));
			}
			stamps_t sp;
			mna_stamps_charge(p, freq, &sp);
			upload_stamps(sb, sp, n, s);
			mna_stamps_free(&sp);
		};
		// fill each system of the batch into its slot; with records, the batch's Lp and P in one
		// pass over them
		if (kf) {
			double kws[FMAX];
			d_complex_t *lps[FMAX], *Ps[FMAX];
			for (int b = 0; b < nb; b++) {
				const double freq = freq_at(p, ifreq0 + f0 + b);
				kws[b] = p->retardation ? (2 * PI * freq / C0) : 0;
				lps[b] = (d_complex_t *)c.sl[b].lp;
				Ps[b] = (d_complex_t *)c.sl[b].P;
			}
			// the symmetric product reads the upper triangles: the batch fill leaves the rest of
			// the lower ones to the factorisation, which mirrors them when it takes a system
			const bool up = mv_sym();
			pgpu_kf_fill_multi(g, nb, kws, lps, nc ? Ps : nullptr, up ? 1 : 0);
			for (int b = 0; b < nb; b++) c.sl[b].upper = up;
		}
		for (int b = 0; b < nb; b++) fill_system(c.sl[b], f0 + b, !kf);
		const auto t1 = std::chrono::steady_clock::now();
		if (st) st->fill_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
		// (re)factor the preconditioner on the system in sb, that of sweep point k
		auto refactor_on = [&](const slotbuf &sb, int k) -> int {
			const auto ta = std::chrono::steady_clock::now();
			const int n = ns;   // the factored order from here on: the Schur complement's
			if (sb.upper) {
				mirror_lower(s, sb.lp, p->nseg);
				if (nc > 0) mirror_lower(s, sb.P, nc);
			}
			if (nc > 0) {
				// C = P(f_ref)^-1 in FP64 (the plain form's P LU and unit solves), then FP32 by columns
				HIPCHK(hipMemcpyAsync(c.Pw, sb.P, (size_t)nc * nc * sizeof(double2), hipMemcpyDeviceToDevice, s));
				if (dlu_factor(s, nc, (d_complex_t *)c.Pw, nc, c.ppiv, c.pperm, nullptr) >= 0) return 2;
				hipLaunchKernelGGL(k_eye, dim3(blocks((long)nc * nc)), dim3(256), 0, s, c.Cd, nc);
				dlu_solve_perm(s, nc, (const d_complex_t *)c.Pw, nc, c.pperm, (d_complex_t *)c.Cd, nc, nc,
				               (d_complex_t *)c.cwork);
				const unsigned ntc = (unsigned)((nc + 31) / 32);
				hipLaunchKernelGGL(k_transpose32<double2>, dim3(ntc, ntc), dim3(256), 0, s, (const double2 *)c.Cd, c.Cc, nc);
			}
			c.omega_ref = sb.omega;
			hipLaunchKernelGGL(k_assemble_schur, dim3(blocks((long)n * n)), dim3(256), 0, s, c.luw, n, p->offS, p->nseg,
			                   (const double2 *)sb.lp, sb.omega, (const int *)c.rowcell, (const double2 *)c.Cd, nc);
			if (sb.nnz > 0)
				hipLaunchKernelGGL(k_stamp_lead, dim3(blocks(sb.nnz)), dim3(256), 0, s, c.luw, (long)n, n, sb.ri, sb.ci,
				                   (const double2 *)sb.val, sb.nnz);
			if (lu64()) {
				const int info = dlu_factor(s, n, (d_complex_t *)c.luw, n, c.piv, c.perm, st);
				if (info >= 0) return 2;
				build_inverses(c, s);
			}
			else {
				// FP32 factors (FP64 panels, FP32 MFMA updates) inverted on FP32 MFMA: the
				// preconditioner is stored in FP32 either way; the FP64 inverse workspace
				// holds the FP32 copy
				float2 *a32 = (float2 *)c.inv;
				hipLaunchKernelGGL(k_demote, dim3(blocks((long)n * n)), dim3(256), 0, s, (const double2 *)c.luw, a32, (long)n * n);
				int info = slu_factor(s, n, a32, n, c.piv, c.perm, st);
				if (info >= 0) {
					// a pivot that rounds to zero in FP32: the FP64 factorisation
					const int info64 = dlu_factor(s, n, (d_complex_t *)c.luw, n, c.piv, c.perm, st);
					if (info64 >= 0) return 2;
					build_inverses(c, s);
				}
				else {
					// each inverse built row-major in free workspace, then stored by columns; the two
					// are independent and run on two streams (the recursion's small levels leave
					// most of the GPU idle): Linv from the free half of the inverse workspace on s,
					// Uinv from the FP64 assembly buffer (no longer needed) on the second stream
					float2 *tmpL = a32 + (size_t)n * n, *tmpU = (float2 *)c.luw;
					const unsigned nt = (unsigned)((n + 31) / 32);
					HIPCHK(hipEventRecord(c.ev_f, s));
					HIPCHK(hipStreamWaitEvent(c.s2, c.ev_f, 0));
					slu_tri_inverse(c.s2, n, a32, n, 0, tmpU, c.scr2, c.dinvU);
					hipLaunchKernelGGL(k_transpose32<float2>, dim3(nt, nt), dim3(256), 0, c.s2, (const float2 *)tmpU, c.MU, n);
					HIPCHK(hipEventRecord(c.ev_u, c.s2));
					slu_tri_inverse(s, n, a32, n, 1, tmpL, c.scr, c.dinvL);
					hipLaunchKernelGGL(k_transpose32<float2>, dim3(nt, nt), dim3(256), 0, s, (const float2 *)tmpL, c.ML, n);
					HIPCHK(hipStreamWaitEvent(s, c.ev_u, 0));
				}
			}
			HIPCHK(hipStreamSynchronize(s));
			c.fref = freq_at(p, ifreq0 + k);
			c.lu_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta).count();
			if (st) {
				st->n_lu++;
				st->lu_ms += c.lu_ms;
			}
			c.served = 0;
			c.its_ref = -1;
			c.lock_ref = -1;
			c.excess_ms = 0;
			return 0;
		};
		auto refactor = [&](int b) -> int { return refactor_on(c.sl[b], f0 + b); };
		// a fresh factorisation when there is none, or when the refactor rule marked the current
		// one stale.  Iteration counts grow with the distance from the factorisation's
		// frequency on either side, so the factorisation goes to the middle of the run of points
		// it is expected to serve: half the previous factorisation's run ahead of this batch
		// (the middle of this batch for the first one, or under the iteration-ratio rule)
		if ((c.fref < 0) || c.stale) {
			if (c.served > 0) c.span = c.served;
			// (any point of the deck's sweep: a sweep called in runs of points still looks ahead)
			const int last = std::max(nfreq - 1, p->nfreq - 1 - ifreq0);
			const int k = ((c.span > 0) && (refactor_alpha() > 0)) ? std::min(last, f0 + c.span / 2) : f0 + nb / 2;
			if (k < f0 + nb) {
				rc = refactor(k - f0);
			}
			else {
				// the system of point k: Lp and P in the inverse workspace (free until the
				// factorisation demotes into it), stamps in their own buffers
				c.ahead.lp = c.inv;
				c.ahead.P = c.inv + (size_t)p->nseg * p->nseg;
				fill_system(c.ahead, k, true);
				rc = refactor_on(c.ahead, k);
			}
			if (rc) break;
			c.stale = false;
		}
		int its_batch = 0, its_lock = 0;
		double batch_ms = 0;
		// every port: batched GMRES over the nb systems
		for (int j = 0; (j < np) && !rc; j++) {
			std::vector<double2> hb((size_t)n * F, make_double2(0, 0));
			for (int b = 0; b < F; b++) {
				if (i1[j] >= 0) hb[(size_t)i1[j] * F + b].x += 1;
				if (i2[j] >= 0) hb[(size_t)i2[j] * F + b].x -= 1;
			}
			HIPCHK(hipMemcpyAsync(c.Bv, hb.data(), hb.size() * sizeof(double2), hipMemcpyHostToDevice, s));
			const auto tg = std::chrono::steady_clock::now();
			g_mvt.on = (st != nullptr);
			g_mvt.reset();
			// initial guesses: the previous batch's solutions of this port extrapolated to this
			// batch's frequencies (the solution varies smoothly between sweep points)
			bool warm = false;
			if ((warm_start() > 0) && (c.xw_nb > 0) && (j < (int)c.xw.size())) {
				xcoef tc{};
				tc.nq = std::min(warm_start(), c.xw_nb);
				for (int q = 0; q < tc.nq; q++) tc.col[q] = c.xw_nb - 1 - q;
				for (int q = 1; q < tc.nq; q++)
					if (c.xw_f[tc.col[q]] == c.xw_f[tc.col[q - 1]]) tc.nq = 1;   // repeated frequency: constant
				for (int b = 0; b < nb; b++) {
					const double fb = freq_at(p, ifreq0 + f0 + b);
					for (int q = 0; q < tc.nq; q++) {
						double w = 1;
						for (int r = 0; r < tc.nq; r++)
							if (r != q) w *= (fb - c.xw_f[tc.col[r]]) / (c.xw_f[tc.col[q]] - c.xw_f[tc.col[r]]);
						tc.w[b][q] = w;
					}
				}
				hipLaunchKernelGGL(k_extrap, dim3(blocks((long)n * F)), dim3(256), 0, s, (const double2 *)c.xw[j], c.Xv, n, F,
				                   nb, tc);
				warm = true;
			}
			std::vector<int> its = gmres_batch(c, s, F, nb, warm);
			if (st) mv_collect(st);
			int lock = 0;
			for (int b = 0; b < nb; b++) {
				if (its[b] < 0) {
					// stale preconditioner: refactor at this frequency and solve the batch again
					if ((rc = refactor(b))) break;
					its = gmres_batch(c, s, F, nb, true);
					if (st) mv_collect(st);
					if (its[b] < 0) {
						rc = 4;
						break;
					}
				}
				if (st) st->gmres_its += its[b];
				its_batch = std::max(its_batch, its[b]);
				lock = std::max(lock, its[b]);
			}
			its_lock += lock;
			batch_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tg).count();
			if (g_trace) {
				fprintf(stderr, "[sweep_fast] points %d..%d port %d its", ifreq0 + f0, ifreq0 + f0 + nb - 1, j);
				for (int b = 0; b < nb; b++) fprintf(stderr, " %d", its[b]);
				fprintf(stderr, " (ref %.4g Hz, excess %.0f ms of %.0f)\n", c.fref, c.excess_ms, c.lu_ms);
			}
			if (st) st->solve_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tg).count();
			if ((int)c.xw.size() <= j) c.xw.resize(j + 1, nullptr);
			if (c.xw[j] == nullptr) c.xw[j] = dal<double2>((size_t)n * FMAX);
			HIPCHK(hipMemcpyAsync(c.xw[j], c.Xv, (size_t)n * F * sizeof(double2), hipMemcpyDeviceToDevice, s));
			HIPCHK(hipMemcpyAsync(hx.data(), c.Xv, hx.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
			HIPCHK(hipStreamSynchronize(s));
			for (int b = 0; b < nb; b++)
				for (int i = 0; i < np; i++) {
					const double2 v1 = (i1[i] < 0) ? make_double2(0, 0) : hx[(size_t)i1[i] * F + b];
					const double2 v2 = (i2[i] < 0) ? make_double2(0, 0) : hx[(size_t)i2[i] * F + b];
					zmat[((size_t)(f0 + b) * np + i) * np + j] = d_complex(v1.x - v2.x, v1.y - v2.y);
				}
		}
		c.xw_nb = nb;
		for (int b = 0; b < nb; b++) c.xw_f[b] = freq_at(p, ifreq0 + f0 + b);
		c.served += nb;
		if (c.its_ref < 0) {
			c.its_ref = its_batch;
			c.lock_ref = its_lock;
			c.hist.assign(1, std::make_pair(its_lock, batch_ms / std::max(1, its_lock)));
		}
		else if (refactor_alpha() <= 0) {
			// iteration-ratio rule
			if (its_batch > std::max(REFACTOR_MIN_ITS, 2 * c.its_ref)) c.stale = true;
		}
		else if (its_lock > 0) {
			// cost rule: the lock-step iterations beyond the fewest a batch needed since the
			// factorisation (what a fresh factorisation would give), priced at this batch's time
			// per iteration, accumulate until they reach alpha times the factorisation's own
			// time (the ski-rental rule)
			c.lock_ref = std::min(c.lock_ref, its_lock);
			c.hist.push_back(std::make_pair(its_lock, batch_ms / its_lock));
			c.excess_ms = 0;
			for (const auto &h : c.hist) c.excess_ms += (double)(h.first - c.lock_ref) * h.second;
			// with look-ahead, the next batch is assumed to cost at least this one's excess, so
			// the factorisation comes a batch earlier where the iteration counts climb fast
			const double next = refactor_lookahead() ? (double)(its_lock - c.lock_ref) * (batch_ms / its_lock) : 0;
			if (c.excess_ms + next >= refactor_alpha() * c.lu_ms) c.stale = true;
		}
		if (st) st->total_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
	}
	return rc;
}

// forget the preconditioner (next fast sweep refactors)
extern "C" void pgpu_fast_reset(void)
{
	g_fast.fref = -1;
	g_fast.stale = false;
	g_fast.its_ref = -1;
	g_fast.lock_ref = -1;
	g_fast.excess_ms = 0;
	g_fast.served = 0;
	g_fast.span = 0;
	g_fast.xw_nb = 0;
}
