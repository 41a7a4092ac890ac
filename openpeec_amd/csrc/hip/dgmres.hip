// dgmres.hip — one right-preconditioned GMRES(60) cycle on a dense device matrix, preconditioned
// by LU factors of a nearby matrix: the acceleration = 1 path of the reference (gmres_solve,
// iterative.c:83-196 and 230-252; accel_rhs, solve.c:47-64).
//
// Same contract as the reference's cycle: start from x0 = 0 (r = b), build the Krylov basis of
// A M^-1, stop when the least-squares residual |g_{j+1}| reaches tol |b| (return j + 1) or after
// 60 steps (return -1); the solution x = M^-1 V y is formed only on convergence.  The Arnoldi
// step orthogonalises by classical Gram-Schmidt with one re-orthogonalisation (two multi-vector
// passes instead of the reference's j + 1 serial dot / axpy pairs; equal in exact arithmetic, and
// at least as stable); every reduction sums in a fixed order, so results are bitwise reproducible.
// Products: y = A x one wave per row (HBM bound, n^2 16 B per product); M^-1 z through the
// permuted triangular solves of lu.hip.
#include <hip/hip_runtime.h>
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>
#include "../peec_internal.h"
#include "hip_check.hpp"

namespace {

constexpr int GM = 60;        // GMRES_MAXIT (iterative.c:28)
constexpr int DB = 64;        // row blocks of the multi-vector dot products

// y = A x: one wave per row, lanes across the row, the 64 lane sums combined by a fixed butterfly
__global__ __launch_bounds__(256) void k_dgemv(const double2 *__restrict__ A, long lda, int n, const double2 *__restrict__ x,
                                               double2 *__restrict__ y)
{
	const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
	if (row >= n) return;   // the whole wave leaves together
	const double2 *a = A + (long)row * lda;
	double sr = 0, si = 0;
	for (int j = lane; j < n; j += 64) {
		const double2 v = a[j], w = x[j];
		sr += (v.x * w.x) - (v.y * w.y);
		si += (v.x * w.y) + (v.y * w.x);
	}
	for (int off = 32; off > 0; off >>= 1) {
		sr += __shfl_xor(sr, off);
		si += __shfl_xor(si, off);
	}
	if (lane == 0) y[row] = make_double2(sr, si);
}

// part[i][b] = sum over row block b of conj(V_i) . w   (i < nv); grid (DB, nv)
__global__ __launch_bounds__(256) void k_ddots(const double2 *__restrict__ V, long ldv, int n, const double2 *__restrict__ w,
                                               double2 *__restrict__ part)
{
	__shared__ double sr_[256], si_[256];
	const int i = blockIdx.y, b = blockIdx.x, t = threadIdx.x;
	const int rows = (n + DB - 1) / DB, r0 = b * rows, r1 = min(n, r0 + rows);
	const double2 *v = V + (long)i * ldv;
	double sr = 0, si = 0;
	for (int r = r0 + t; r < r1; r += 256) {
		const double2 a = v[r], c = w[r];
		sr += (a.x * c.x) + (a.y * c.y);
		si += (a.x * c.y) - (a.y * c.x);
	}
	sr_[t] = sr;
	si_[t] = si;
	__syncthreads();
	for (int h = 128; h > 0; h >>= 1) {
		if (t < h) {
			sr_[t] += sr_[t + h];
			si_[t] += si_[t + h];
		}
		__syncthreads();
	}
	if (t == 0) part[(long)i * DB + b] = make_double2(sr_[0], si_[0]);
}

// h[i] = sum_b part[i][b] in block order
__global__ void k_dfin(const double2 *part, int nv, double2 *h)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= nv) return;
	double sr = 0, si = 0;
	for (int b = 0; b < DB; b++) {
		sr += part[(long)i * DB + b].x;
		si += part[(long)i * DB + b].y;
	}
	h[i] = make_double2(sr, si);
}

// w -= sum_i h[i] V_i, or w = sum_i h[i] V_i when overwrite is set
__global__ void k_dcomb(double2 *w, const double2 *__restrict__ V, long ldv, int n, int nv, const double2 *__restrict__ h,
                        int overwrite)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= n) return;
	double sr = 0, si = 0;
	for (int i = 0; i < nv; i++) {
		const double2 c = h[i], v = V[(long)i * ldv + r];
		sr += (c.x * v.x) - (c.y * v.y);
		si += (c.x * v.y) + (c.y * v.x);
	}
	if (overwrite) w[r] = make_double2(sr, si);
	else w[r] = make_double2(w[r].x - sr, w[r].y - si);
}

__global__ void k_dscale(const double2 *x, double2 *y, double s, int n)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r < n) y[r] = make_double2(s * x[r].x, s * x[r].y);
}

unsigned nb256(long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct dgmres_ws {
	int n = 0;
	double2 *V = nullptr;      // (GM + 1) x n
	double2 *w = nullptr, *z = nullptr, *work = nullptr;
	double2 *part = nullptr, *h = nullptr;
};

extern "C" {

dgmres_ws *dgmres_ws_create(int n)
{
	dgmres_ws *ws = new dgmres_ws();
	ws->n = n;
	HIPCHK(hipMalloc((void **)&ws->V, (size_t)(GM + 1) * n * sizeof(double2)));
	HIPCHK(hipMalloc((void **)&ws->w, (size_t)n * sizeof(double2)));
	HIPCHK(hipMalloc((void **)&ws->z, (size_t)n * sizeof(double2)));
	HIPCHK(hipMalloc((void **)&ws->work, (size_t)n * sizeof(double2)));
	HIPCHK(hipMalloc((void **)&ws->part, (size_t)(GM + 1) * DB * sizeof(double2)));
	HIPCHK(hipMalloc((void **)&ws->h, (size_t)(GM + 1) * sizeof(double2)));
	return ws;
}

void dgmres_ws_free(dgmres_ws *ws)
{
	if (ws == nullptr) return;
	void *b[] = {ws->V, ws->w, ws->z, ws->work, ws->part, ws->h};
	for (void *x : b)
		if (x) HIPCHK(hipFree(x));
	delete ws;
}

// One cycle on A (n x n, lda) with M = the LU factors LU (lda) and row permutation perm (as
// dlu_factor records it).  b: device right-hand side, replaced by the solution when the cycle
// converges (left untouched otherwise).  Returns iterations (>= 1) or -1.
int dgmres_cycle(void *stream, dgmres_ws *ws, int n, const d_complex_t *A_, int lda, const d_complex_t *LU_, const int *perm,
                 d_complex_t *b_, double tol)
{
	typedef std::complex<double> zc;
	hipStream_t s = (hipStream_t)stream;
	const double2 *A = (const double2 *)A_;
	double2 *b = (double2 *)b_;
	double2 *V = ws->V;
	const long ldv = n;
	auto dots = [&](const double2 *w, int nv, std::vector<zc> &out) {
		hipLaunchKernelGGL(k_ddots, dim3(DB, nv), dim3(256), 0, s, V, ldv, n, w, ws->part);
		hipLaunchKernelGGL(k_dfin, dim3((nv + 63) / 64), dim3(64), 0, s, (const double2 *)ws->part, nv, ws->h);
		std::vector<double2> hh(nv);
		HIPCHK(hipMemcpyAsync(hh.data(), ws->h, (size_t)nv * sizeof(double2), hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		out.resize(nv);
		for (int i = 0; i < nv; i++) out[i] = zc(hh[i].x, hh[i].y);
	};
	auto norm_of = [&](const double2 *w) -> double {
		hipLaunchKernelGGL(k_ddots, dim3(DB, 1), dim3(256), 0, s, w, ldv, n, w, ws->part);
		hipLaunchKernelGGL(k_dfin, dim3(1), dim3(64), 0, s, (const double2 *)ws->part, 1, ws->h);
		double2 hh;
		HIPCHK(hipMemcpyAsync(&hh, ws->h, sizeof(double2), hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		return std::sqrt(hh.x);
	};
	// coefficients to the device; the host buffer lives for the whole cycle and every upload is
	// followed by a synchronising read-back before the next one
	std::vector<double2> hup(GM + 1);
	auto put_h = [&](const std::vector<zc> &c) {
		for (size_t i = 0; i < c.size(); i++) hup[i] = make_double2(c[i].real(), c[i].imag());
		HIPCHK(hipMemcpyAsync(ws->h, hup.data(), c.size() * sizeof(double2), hipMemcpyHostToDevice, s));
	};
	const double beta = norm_of(b);
	if (beta <= 0) {
		// zero right-hand side: the solution is 0 (iterative.c:117-122)
		hipLaunchKernelGGL(k_dscale, dim3(nb256(n)), dim3(256), 0, s, (const double2 *)b, b, 0.0, n);
		HIPCHK(hipStreamSynchronize(s));
		return 1;
	}
	const double tolabs = tol * beta;
	hipLaunchKernelGGL(k_dscale, dim3(nb256(n)), dim3(256), 0, s, (const double2 *)b, V, 1.0 / beta, n);
	std::vector<zc> H((size_t)(GM + 1) * GM, zc(0, 0)), cs(GM), sn(GM), g(GM + 1, zc(0, 0)), c1, c2;
	g[0] = beta;
	int iters = -1;
	for (int j = 0; j < GM; j++) {
		// w = A M^-1 v_j
		HIPCHK(hipMemcpyAsync(ws->z, V + (long)j * ldv, (size_t)n * sizeof(double2), hipMemcpyDeviceToDevice, s));
		dlu_solve_perm(s, n, LU_, lda, perm, (d_complex_t *)ws->z, 1, 1, (d_complex_t *)ws->work);
		hipLaunchKernelGGL(k_dgemv, dim3((n + 3) / 4), dim3(256), 0, s, A, (long)lda, n, (const double2 *)ws->z, ws->w);
		// two Gram-Schmidt passes against v_0 .. v_j
		dots(ws->w, j + 1, c1);
		put_h(c1);
		hipLaunchKernelGGL(k_dcomb, dim3(nb256(n)), dim3(256), 0, s, ws->w, (const double2 *)V, ldv, n, j + 1,
		                   (const double2 *)ws->h, 0);
		dots(ws->w, j + 1, c2);
		put_h(c2);
		hipLaunchKernelGGL(k_dcomb, dim3(nb256(n)), dim3(256), 0, s, ws->w, (const double2 *)V, ldv, n, j + 1,
		                   (const double2 *)ws->h, 0);
		zc *Hc = &H[(size_t)j * (GM + 1)];
		for (int i = 0; i <= j; i++) Hc[i] = c1[i] + c2[i];
		const double hn = norm_of(ws->w);
		Hc[j + 1] = hn;
		if (hn > 0)
			hipLaunchKernelGGL(k_dscale, dim3(nb256(n)), dim3(256), 0, s, (const double2 *)ws->w, V + (long)(j + 1) * ldv,
			                   1.0 / hn, n);
		// earlier rotations on the new column, then the rotation that zeroes H[j+1][j]
		for (int i = 0; i < j; i++) {
			const zc h1 = Hc[i], h2 = Hc[i + 1];
			Hc[i] = std::conj(cs[i]) * h1 + std::conj(sn[i]) * h2;
			Hc[i + 1] = cs[i] * h2 - sn[i] * h1;
		}
		const zc h1 = Hc[j], h2 = Hc[j + 1];
		const double rho = std::sqrt(std::norm(h1) + std::norm(h2));
		if (rho <= 0) break;   // breakdown: already exact
		cs[j] = h1 / rho;
		sn[j] = h2 / rho;
		Hc[j] = rho;
		Hc[j + 1] = 0;
		g[j + 1] = -sn[j] * g[j];
		g[j] = std::conj(cs[j]) * g[j];
		if (std::abs(g[j + 1]) <= tolabs) {
			iters = j + 1;
			break;
		}
	}
	if (iters > 0) {
		// y = H^-1 g by back substitution; b = M^-1 (V y)
		const int k = iters;
		std::vector<zc> y(k);
		for (int i = k - 1; i >= 0; i--) {
			zc acc = g[i];
			for (int l = i + 1; l < k; l++) acc -= H[(size_t)l * (GM + 1) + i] * y[l];
			y[i] = acc / H[(size_t)i * (GM + 1) + i];
		}
		put_h(y);
		hipLaunchKernelGGL(k_dcomb, dim3(nb256(n)), dim3(256), 0, s, b, (const double2 *)V, ldv, n, k, (const double2 *)ws->h, 1);
		dlu_solve_perm(s, n, LU_, lda, perm, b_, 1, 1, (d_complex_t *)ws->work);
	}
	HIPCHK(hipStreamSynchronize(s));
	return iters;
}

// gmres_solve (iterative.c:230-252) with host arrays: a is the current matrix, alu / piv the LU
// factors of a nearby one as lu_decomp leaves them.  b is replaced by the solution when the
// cycle converges to tol |b| and left untouched otherwise.  Returns iterations or -1.
int gmres_solve(int n, const d_complex_t *a, const d_complex_t *alu, const int *piv, d_complex_t *b, double tol)
{
	if (n <= 0) return -1;
	if (!pgpu_device_ok()) pgpu_fatal("no HIP device");
	// lu_decomp's interchange record as the row permutation of the factors
	std::vector<int> perm(n);
	for (int i = 0; i < n; i++) perm[i] = i;
	for (int k = 0; k < n; k++)
		if (piv[k] != k) std::swap(perm[k], perm[piv[k]]);
	const size_t bytes = (size_t)n * n * sizeof(d_complex_t);
	d_complex_t *dA = nullptr, *dLU = nullptr, *db = nullptr;
	int *dperm = nullptr;
	HIPCHK(hipMalloc((void **)&dA, bytes));
	HIPCHK(hipMalloc((void **)&dLU, bytes));
	HIPCHK(hipMalloc((void **)&db, (size_t)n * sizeof(d_complex_t)));
	HIPCHK(hipMalloc((void **)&dperm, (size_t)n * sizeof(int)));
	HIPCHK(hipMemcpy(dA, a, bytes, hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(dLU, alu, bytes, hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(db, b, (size_t)n * sizeof(d_complex_t), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(dperm, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
	dgmres_ws *ws = dgmres_ws_create(n);
	const int it = dgmres_cycle(nullptr, ws, n, dA, n, dLU, dperm, db, tol);
	if (it > 0) HIPCHK(hipMemcpy(b, db, (size_t)n * sizeof(d_complex_t), hipMemcpyDeviceToHost));
	dgmres_ws_free(ws);
	void *bb[] = {dA, dLU, db, dperm};
	for (void *x : bb) HIPCHK(hipFree(x));
	return it;
}

}  // extern "C"
