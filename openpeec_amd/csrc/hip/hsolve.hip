// hsolve.hip — the compressed (compression = 1) solve on the device: matrix-free MNA product,
// leaf-block + circuit-Schur preconditioner, restarted GMRES, and the per-frequency sweep.
//
// Reference: mna_sparse / mna_apply (mna.c:317-364), pc_build / pc_apply (precond.c:139-433),
// gmres_hmat (iterative.c:285-354), the compressed branch of solve (solve.c:246-300).
//
//   A = | S11  B_1  B_2 |   n1 = offS circuit unknowns (nodes, inductor and source currents)
//       | C_1  D_1   0  |   D_b = scale_b x kernel block b (Lp: -jw, P: 1) + its sparse diagonal
//       | C_2   0   D_2 |
// The preconditioner keeps of each D_b only the diagonal blocks of its cluster-tree leaves
// (m <= 32), inverts them (one workgroup each), forms the Schur complement
// S = S11 - sum B_b D_b^-1 C_b densely (each row of S owned by one workgroup, contributions
// in the reference's leaf / cell / coupling order), factors S on the device and stores the
// explicit inverses of its factors, so M^-1 is a sequence of sparse and dense products
// without substitution chains.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>
#include "../peec_internal.h"
#include "../../../include/peec_mi355x.h"
#include "fill_kernels.hpp"
#include "hip_check.hpp"
#include "hmat_dev.hpp"
#include "mlp.hpp"

namespace {

constexpr int PC_MAXM = 64;        // largest leaf (HM_LEAF = 32)
constexpr int PC_MAXT = 4 * PC_MAXM;
constexpr int GM_M = 60;           // GMRES restart length (GMRES_MAXIT, iterative.c:29)
constexpr int GM_CYC = 10;         // restarts (GMRES_CYCLES)

template <typename T>
T *dal(size_t n)
{
	T *p = nullptr;
	HIPCHK(hipMalloc((void **)&p, (n ? n : 1) * sizeof(T)));
	return p;
}

template <typename T>
T *up(hipStream_t s, const std::vector<T> &v)
{
	T *p = dal<T>(v.size());
	if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
	return p;
}

unsigned nb(long n, int t = 256)
{
	return (unsigned)((n + t - 1) / t);
}

__device__ __forceinline__ double2 cm(double2 a, double2 b)
{
	return make_double2((a.x * b.x) - (a.y * b.y), (a.x * b.y) + (a.y * b.x));
}

// ---------------------------------------------------------------- sparse rows (CSR, entry order kept)

// y[r] = sum_e val[e] x[col[e]] over the row's entries, in order; add != 0 adds to y instead
__global__ void k_csr(int n, const int *rp, const int *col, const double2 *val, const double2 *x, double2 *y, int add,
                      double sign)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= n) return;
	double2 acc = add ? y[r] : make_double2(0, 0);
	for (int e = rp[r]; e < rp[r + 1]; e++) {
		const double2 p = cm(val[e], x[col[e]]);
		acc.x = acc.x + (sign * p.x);
		acc.y = acc.y + (sign * p.y);
	}
	y[r] = acc;
}

// y[off + k] += scale w[k]
__global__ void k_addscaled(double2 *y, const double2 *w, int n, double2 sc)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n) return;
	const double2 p = cm(sc, w[k]);
	y[k] = make_double2(y[k].x + p.x, y[k].y + p.y);
}

__global__ void k_copy(const double2 *a, double2 *b, int n)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n) b[k] = a[k];
}

__global__ void k_zero(double2 *a, long n)
{
	const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n) a[k] = make_double2(0, 0);
}

// ---------------------------------------------------------------- leaf blocks

struct leafd {
	int m;            // cells in the leaf
	int idx;          // offset of its cell list (block-local numbers) in the index pool
	long doff;        // its m x m block in the D pool
	long xoff;        // its m x t Schur factor X = D^-1 C_r in the X pool
	int t, toff;      // touched circuit unknowns and their offset in the tset pool
};

// D = scale * (leaf diagonal block of the kernel matrix)
__global__ void k_leaf_init(const leafd *L, const long *src, const double2 *Dk, double2 sc, double2 *D)
{
	const leafd l = L[blockIdx.x];
	const long s0 = src[blockIdx.x];
	for (int e = threadIdx.x; e < l.m * l.m; e += blockDim.x) D[l.doff + e] = cm(sc, Dk[s0 + e]);
}

// same-leaf sparse entries added in their order (one thread per leaf)
__global__ void k_leaf_add(int nleaf, const int *aoff, const int *apos, const double2 *aval, const leafd *L, double2 *D)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= nleaf) return;
	for (int e = aoff[r]; e < aoff[r + 1]; e++) {
		double2 &d = D[L[r].doff + apos[e]];
		d = make_double2(d.x + aval[e].x, d.y + aval[e].y);
	}
}

// in place D <- D^-1 by Gauss-Jordan with row pivoting on |re| + |im| (first maximum), one
// workgroup of 2 * PC_MAXM threads per leaf; thread c owns column c of [D | I].  A pivot below
// 1e-30 marks the leaf singular (*bad = leaf + 1).
__global__ __launch_bounds__(2 * PC_MAXM) void k_leaf_inv(const leafd *L, double2 *D, int *bad)
{
	__shared__ double2 A[PC_MAXM][2 * PC_MAXM + 1];
	__shared__ double2 colk[PC_MAXM];
	__shared__ int sp;
	const leafd l = L[blockIdx.x];
	const int m = l.m, c = threadIdx.x;
	if (c < 2 * m)
		for (int r = 0; r < m; r++)
			A[r][c] = (c < m) ? D[l.doff + (long)r * m + c] : make_double2((c - m == r) ? 1.0 : 0.0, 0);
	__syncthreads();
	for (int k = 0; k < m; k++) {
		if (c == 0) {
			int p = -1;
			double best = -1;
			for (int r = k; r < m; r++) {
				const double v = fabs(A[r][k].x) + fabs(A[r][k].y);
				if (v > best) {
					best = v;
					p = r;
				}
			}
			sp = (best < 1e-30) ? -1 : p;
			if (sp < 0) atomicCAS(bad, 0, (int)blockIdx.x + 1);
		}
		__syncthreads();
		const int p = sp;
		if (p < 0) return;
		if ((c < 2 * m) && (p != k)) {
			const double2 t = A[k][c];
			A[k][c] = A[p][c];
			A[p][c] = t;
		}
		__syncthreads();
		const double2 piv = A[k][k];
		const double n2 = (piv.x * piv.x) + (piv.y * piv.y);
		const double2 pinv = make_double2(piv.x / n2, -piv.y / n2);
		__syncthreads();
		if (c < 2 * m) A[k][c] = cm(A[k][c], pinv);
		if (c < m) colk[c] = A[c][k];
		__syncthreads();
		if (c < 2 * m)
			for (int r = 0; r < m; r++) {
				if (r == k) continue;
				const double2 f = colk[r], a = A[k][c];
				const double2 q = cm(f, a);
				A[r][c] = make_double2(A[r][c].x - q.x, A[r][c].y - q.y);
			}
		__syncthreads();
	}
	if ((c >= m) && (c < 2 * m))
		for (int r = 0; r < m; r++) D[l.doff + (long)r * m + (c - m)] = A[r][c];
}

// X_r = D_r^-1 T_r (m x t), T_r the leaf's coupling columns to the circuit unknowns it touches
__global__ void k_leaf_x(const leafd *L, const double2 *Dinv, const double2 *T, double2 *X)
{
	const leafd l = L[blockIdx.x];
	for (int e = threadIdx.x; e < l.m * l.t; e += blockDim.x) {
		const int a = e / l.t, cc = e - a * l.t;
		double2 acc = make_double2(0, 0);
		for (int b = 0; b < l.m; b++) {
			const double2 p = cm(Dinv[l.doff + (long)a * l.m + b], T[l.xoff + (long)b * l.t + cc]);
			acc.x += p.x;
			acc.y += p.y;
		}
		X[l.xoff + e] = acc;
	}
}

// S row i -= sum over its (leaf, cell, coupling) entries, in order, of bval * X_leaf[cell][:]
__global__ void k_schur_rows(int n1, const int *roff, const int *rleaf, const int *rcell, const double2 *rval,
                             const leafd *L, const int *tset, const double2 *X, double2 *S)
{
	const int i = blockIdx.x;
	if (i >= n1) return;
	for (int e = roff[i]; e < roff[i + 1]; e++) {
		const leafd l = L[rleaf[e]];
		const int a = rcell[e];
		const double2 bv = rval[e];
		for (int cc = threadIdx.x; cc < l.t; cc += blockDim.x) {
			const double2 p = cm(bv, X[l.xoff + (long)a * l.t + cc]);
			double2 &s = S[(long)i * n1 + tset[l.toff + cc]];
			s = make_double2(s.x - p.x, s.y - p.y);
		}
		__syncthreads();
	}
}

// z[cells of leaf] = D^-1 u[cells of leaf], one thread per cell row (block-local numbering); the
// leaf's entries of u are gathered into LDS first (leaves of up to LEAF_LDS cells), so a row is
// a chain of independent loads of D^-1; the sum runs over b in increasing order either way
constexpr int LEAF_LDS = 128;
__global__ void k_leaf_apply(int nleaf, const leafd *L, const int *lidx, const double2 *Dinv, const double2 *u, double2 *z)
{
	__shared__ double2 us[LEAF_LDS];
	const int r = blockIdx.x;
	if (r >= nleaf) return;
	const leafd l = L[r];
	const bool staged = l.m <= LEAF_LDS;
	if (staged) {
		for (int b = threadIdx.x; b < l.m; b += blockDim.x) us[b] = u[lidx[l.idx + b]];
		__syncthreads();
	}
	for (int a = threadIdx.x; a < l.m; a += blockDim.x) {
		const double2 *d = Dinv + l.doff + (long)a * l.m;
		double2 acc = make_double2(0, 0);
#pragma unroll 8
		for (int b = 0; b < l.m; b++) {
			const double2 p = cm(d[b], staged ? us[b] : u[lidx[l.idx + b]]);
			acc.x += p.x;
			acc.y += p.y;
		}
		z[lidx[l.idx + a]] = acc;
	}
}

// S[key[e]] = val[e] (distinct keys)
__global__ void k_scatter(const long *key, const double2 *val, int n, double2 *S)
{
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e < n) S[key[e]] = val[e];
}

// y = P x (y[k] = x[perm[k]])
__global__ void k_perm(const int *perm, const double2 *x, double2 *y, int n)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n) y[k] = x[perm[k]];
}

// y = T x for a dense triangular FP64 matrix (lower: columns <= row), one wave per row
template <bool LOWER>
__global__ __launch_bounds__(256) void k_trmv(const double2 *T, int n, const double2 *x, double2 *y)
{
	const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
	if (row >= n) return;
	const int c0 = LOWER ? 0 : row, c1 = LOWER ? row + 1 : n;
	double sr = 0, si = 0;
	for (int c = c0 + lane; c < c1; c += 64) {
		const double2 p = cm(T[(long)row * n + c], x[c]);
		sr += p.x;
		si += p.y;
	}
	for (int off = 32; off > 0; off >>= 1) {
		sr += __shfl_xor(sr, off);
		si += __shfl_xor(si, off);
	}
	if (lane == 0) y[row] = make_double2(sr, si);
}

// ---------------------------------------------------------------- GMRES helpers (fixed-order)

constexpr int RD = 256;

// partial[b] = sum over block b of conj(a) . c
__global__ __launch_bounds__(RD) void k_dotp(const double2 *a, const double2 *c, int n, double2 *partial)
{
	__shared__ double2 red[RD];
	const int t = threadIdx.x;
	double sr = 0, si = 0;
	for (int i = blockIdx.x * RD + t; i < n; i += gridDim.x * RD) {
		const double2 u = a[i], v = c[i];
		sr += (u.x * v.x) + (u.y * v.y);
		si += (u.x * v.y) - (u.y * v.x);
	}
	red[t] = make_double2(sr, si);
	__syncthreads();
	for (int s = RD / 2; s > 0; s >>= 1) {
		if (t < s) red[t] = make_double2(red[t].x + red[t + s].x, red[t].y + red[t + s].y);
		__syncthreads();
	}
	if (t == 0) partial[blockIdx.x] = red[0];
}

// conj(V_i) . w for the nv vectors V_i = V + i n: block (b, i) sums rows b, b + MD_BLK, ... of
// vector i into partial[i MD_BLK + b] (fixed order)
constexpr int MD_BLK = 64;
__global__ __launch_bounds__(RD) void k_mdot(const double2 *V, long n, int nv, const double2 *w, double2 *partial)
{
	__shared__ double2 red[RD];
	const int t = threadIdx.x, i = blockIdx.y;
	const double2 *a = V + (long)i * n;
	double sr = 0, si = 0;
	for (long r = (long)blockIdx.x * RD + t; r < n; r += (long)MD_BLK * RD) {
		const double2 u = a[r], v = w[r];
		sr += (u.x * v.x) + (u.y * v.y);
		si += (u.x * v.y) - (u.y * v.x);
	}
	red[t] = make_double2(sr, si);
	__syncthreads();
	for (int s2 = RD / 2; s2 > 0; s2 >>= 1) {
		if (t < s2) red[t] = make_double2(red[t].x + red[t + s2].x, red[t].y + red[t + s2].y);
		__syncthreads();
	}
	if (t == 0) partial[(long)i * MD_BLK + blockIdx.x] = red[0];
}

// w -= sum_i h_i V_i
__global__ void k_maxpy(double2 *w, const double2 *V, long n, int nv, const double2 *h)
{
	const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= n) return;
	double2 acc = w[r];
	for (int i = 0; i < nv; i++) {
		const double2 v = V[(long)i * n + r], c = h[i];
		acc.x -= (c.x * v.x) - (c.y * v.y);
		acc.y -= (c.x * v.y) + (c.y * v.x);
	}
	w[r] = acc;
}

// coefficients of one Gram-Schmidt pass from the multi-vector dot partials, summed over the
// partial blocks in increasing order: h[i] = sum_b part[i][b]; acc[i] = h[i] on the first pass,
// acc[i] + h[i] on the second
__global__ void k_mdot_fin(const double2 *part, int nv, int nblk, double2 *h, double2 *acc, int pass)
{
	for (int i = threadIdx.x; i < nv; i += blockDim.x) {
		double sr = 0, si = 0;
		for (int b = 0; b < nblk; b++) {
			sr += part[(size_t)i * nblk + b].x;
			si += part[(size_t)i * nblk + b].y;
		}
		h[i] = make_double2(sr, si);
		acc[i] = (pass == 0) ? make_double2(sr, si) : make_double2(acc[i].x + sr, acc[i].y + si);
	}
}

// w -= h v  and  x += y v  style updates
__global__ void k_axpy(double2 *w, const double2 *v, double2 h, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const double2 p = cm(h, v[i]);
	w[i] = make_double2(w[i].x + p.x, w[i].y + p.y);
}

__global__ void k_scal(const double2 *w, double2 *v, double s, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) v[i] = make_double2(s * w[i].x, s * w[i].y);
}

__global__ void k_sub(const double2 *b, double2 *r, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) r[i] = make_double2(b[i].x - r[i].x, b[i].y - r[i].y);
}

constexpr int DOT_BLK = 128;

// PEEC_TRACE=1 prints the true residual of every GMRES restart to stderr
const bool g_htrace = getenv("PEEC_TRACE") && getenv("PEEC_TRACE")[0] == '1';

}  // namespace

// ------------------------------------------------------------------------------------------
// device operator (sparse part + scaled compressed blocks) and preconditioner

struct hop {
	int n = 0;
	hipStream_t s = nullptr;
	int *rp = nullptr, *col = nullptr;
	double2 *val = nullptr;
	int nblk = 0;
	int off[HBLK_MAX] = {0, 0}, bn[HBLK_MAX] = {0, 0};
	const hmat_t *h[HBLK_MAX] = {nullptr, nullptr};
	double2 scale[HBLK_MAX];
	double2 *work = nullptr;
	double2 *partial = nullptr;
	std::vector<double2> hp;
	double2 *mpart = nullptr, *mh = nullptr;   // multi-vector dots: partials, coefficients
	double2 *hacc = nullptr;                    // both passes' coefficients summed
	std::vector<double2> hmp, hmh;
};

struct pcblk {
	int off = 0, n = 0, nleaf = 0;
	leafd *dL = nullptr;
	int *dlidx = nullptr;
	double2 *dD = nullptr;        // leaf inverses
	// B: circuit rows x block columns (CSR over circuit rows); C: block rows x circuit columns
	int *brp = nullptr, *bcol = nullptr;
	double2 *bval = nullptr;
	int *crp = nullptr, *ccol = nullptr;
	double2 *cval = nullptr;
};

struct precond_t {
	int n = 0, n1 = 0, nblk = 0;
	hipStream_t s = nullptr;
	pcblk blk[HBLK_MAX];
	double2 *Linv = nullptr, *Uinv = nullptr;   // Schur complement factors, explicit inverses
	int *perm = nullptr;
	double2 *w1 = nullptr, *w2 = nullptr, *wb = nullptr, *zb = nullptr;
	double2 *d_b = nullptr;
	double mb = 0;
};

namespace {

// CSR of the triplets (rows < nrow), entries of a row in their original order
void csr_build(int nrow, const std::vector<int> &ri, const std::vector<int> &ci, const std::vector<double2> &v,
               std::vector<int> &rp, std::vector<int> &col, std::vector<double2> &val)
{
	rp.assign(nrow + 1, 0);
	for (int r : ri) rp[r + 1]++;
	for (int i = 0; i < nrow; i++) rp[i + 1] += rp[i];
	std::vector<int> pos(rp.begin(), rp.end() - 1);
	col.resize(ri.size());
	val.resize(ri.size());
	for (size_t e = 0; e < ri.size(); e++) {
		const int k = pos[ri[e]]++;
		col[k] = ci[e];
		val[k] = v[e];
	}
}

}  // namespace

// y = A x on the device (x, y device vectors of length n)
void hop_apply(hop &o, const double2 *x, double2 *y)
{
	hipLaunchKernelGGL(k_csr, dim3(nb(o.n)), dim3(256), 0, o.s, o.n, (const int *)o.rp, (const int *)o.col,
	                   (const double2 *)o.val, x, y, 0, 1.0);
	for (int b = 0; b < o.nblk; b++) {
		if ((o.h[b] == nullptr) || (o.bn[b] <= 0)) continue;
		hm_matvec_dev(o.h[b], x + o.off[b], o.work);
		hipLaunchKernelGGL(k_addscaled, dim3(nb(o.bn[b])), dim3(256), 0, o.s, y + o.off[b], (const double2 *)o.work, o.bn[b],
		                   o.scale[b]);
	}
}

void hop_init(hop &o, const peec_t *p, const mna_sparse_t *sp, const hblock_t *blk, int nblk, hipStream_t s)
{
	o.n = p->nunknown;
	o.s = s;
	std::vector<int> ri(sp->ri, sp->ri + sp->nnz), ci(sp->ci, sp->ci + sp->nnz), rp, col;
	std::vector<double2> v(sp->nnz), val;
	for (int e = 0; e < sp->nnz; e++) v[e] = make_double2(sp->val[e].r, sp->val[e].i);
	csr_build(o.n, ri, ci, v, rp, col, val);
	o.rp = up(s, rp);
	o.col = up(s, col);
	o.val = up(s, val);
	o.nblk = nblk;
	int maxn = 1;
	for (int b = 0; b < nblk; b++) {
		o.off[b] = blk[b].off;
		o.bn[b] = blk[b].n;
		o.h[b] = blk[b].h;
		o.scale[b] = make_double2(blk[b].scale.r, blk[b].scale.i);
		maxn = std::max(maxn, blk[b].n);
	}
	o.work = dal<double2>(maxn);
	o.partial = dal<double2>(DOT_BLK);
	o.hp.resize(DOT_BLK);
	o.mpart = dal<double2>((size_t)(GM_M + 1) * MD_BLK);
	o.mh = dal<double2>(GM_M + 1);
	o.hacc = dal<double2>(GM_M + 1);
	o.hmp.resize((size_t)(GM_M + 1) * MD_BLK);
	o.hmh.resize(GM_M + 1);
}

void hop_free(hop &o)
{
	void *b[] = {o.rp, o.col, o.val, o.work, o.partial, o.mpart, o.mh, o.hacc};
	for (void *x : b)
		if (x) HIPCHK(hipFree(x));
	o = hop();
}

// fixed-order conj(a) . c on the device, summed on the host in block order
std::complex<double> hdot(hop &o, const double2 *a, const double2 *c, int n)
{
	hipLaunchKernelGGL(k_dotp, dim3(DOT_BLK), dim3(RD), 0, o.s, a, c, n, o.partial);
	HIPCHK(hipMemcpyAsync(o.hp.data(), o.partial, DOT_BLK * sizeof(double2), hipMemcpyDeviceToHost, o.s));
	HIPCHK(hipStreamSynchronize(o.s));
	double sr = 0, si = 0;
	for (int b = 0; b < DOT_BLK; b++) {
		sr += o.hp[b].x;
		si += o.hp[b].y;
	}
	return {sr, si};
}

// classical Gram-Schmidt with one reorthogonalisation of w against V_0 .. V_j: per pass one
// multi-vector dot kernel, the coefficients summed on the device and one multi-vector update;
// then |w|^2, and one copy to the host for both: Hc[i] += the coefficients of both passes, and
// the return value is |w|^2 (the partials summed in the order hdot sums them)
double orth_cgs2(hop &o, const double2 *V, int j, double2 *w, std::complex<double> *Hc)
{
	const int n = o.n, nv = j + 1;
	for (int pass = 0; pass < 2; pass++) {
		hipLaunchKernelGGL(k_mdot, dim3(MD_BLK, nv), dim3(RD), 0, o.s, V, (long)n, nv, (const double2 *)w, o.mpart);
		hipLaunchKernelGGL(k_mdot_fin, dim3(1), dim3(64), 0, o.s, (const double2 *)o.mpart, nv, (int)MD_BLK, o.mh, o.hacc, pass);
		hipLaunchKernelGGL(k_maxpy, dim3(nb(n)), dim3(256), 0, o.s, w, V, (long)n, nv, (const double2 *)o.mh);
	}
	hipLaunchKernelGGL(k_dotp, dim3(DOT_BLK), dim3(RD), 0, o.s, (const double2 *)w, (const double2 *)w, n, o.partial);
	HIPCHK(hipMemcpyAsync(o.hmh.data(), o.hacc, (size_t)nv * sizeof(double2), hipMemcpyDeviceToHost, o.s));
	HIPCHK(hipMemcpyAsync(o.hp.data(), o.partial, DOT_BLK * sizeof(double2), hipMemcpyDeviceToHost, o.s));
	HIPCHK(hipStreamSynchronize(o.s));
	for (int i = 0; i < nv; i++) Hc[i] += std::complex<double>(o.hmh[i].x, o.hmh[i].y);
	double sr = 0;
	for (int b = 0; b < DOT_BLK; b++) sr += o.hp[b].x;
	return sr;
}

// z = M^-1 v on the device (v, z length n; v is not modified)
void pc_apply_dev(precond_t *pc, const double2 *v, double2 *z)
{
	hipStream_t s = pc->s;
	const int n1 = pc->n1;
	// w1 = u1 - sum_b B_b D_b^-1 u_b
	hipLaunchKernelGGL(k_copy, dim3(nb(n1)), dim3(256), 0, s, v, pc->w1, n1);
	for (int k = 0; k < pc->nblk; k++) {
		const pcblk &b = pc->blk[k];
		hipLaunchKernelGGL(k_leaf_apply, dim3(b.nleaf), dim3(64), 0, s, b.nleaf, (const leafd *)b.dL, (const int *)b.dlidx,
		                   (const double2 *)b.dD, v + b.off, pc->zb);
		hipLaunchKernelGGL(k_csr, dim3(nb(n1)), dim3(256), 0, s, n1, (const int *)b.brp, (const int *)b.bcol,
		                   (const double2 *)b.bval, (const double2 *)pc->zb, pc->w1, 1, -1.0);
	}
	// w1 = S^-1 w1 = Uinv (Linv (P w1))
	if (n1 > 0) {
		hipLaunchKernelGGL(k_perm, dim3(nb(n1)), dim3(256), 0, s, (const int *)pc->perm, (const double2 *)pc->w1, pc->w2, n1);
		hipLaunchKernelGGL(k_trmv<true>, dim3((n1 + 3) / 4), dim3(256), 0, s, (const double2 *)pc->Linv, n1,
		                   (const double2 *)pc->w2, pc->w1);
		hipLaunchKernelGGL(k_trmv<false>, dim3((n1 + 3) / 4), dim3(256), 0, s, (const double2 *)pc->Uinv, n1,
		                   (const double2 *)pc->w1, pc->w2);
		hipLaunchKernelGGL(k_copy, dim3(nb(n1)), dim3(256), 0, s, (const double2 *)pc->w2, z, n1);
	}
	// z_b = D_b^-1 (u_b - C_b z1)
	for (int k = 0; k < pc->nblk; k++) {
		const pcblk &b = pc->blk[k];
		hipLaunchKernelGGL(k_copy, dim3(nb(b.n)), dim3(256), 0, s, v + b.off, pc->wb, b.n);
		hipLaunchKernelGGL(k_csr, dim3(nb(b.n)), dim3(256), 0, s, b.n, (const int *)b.crp, (const int *)b.ccol,
		                   (const double2 *)b.cval, (const double2 *)z, pc->wb, 1, -1.0);
		hipLaunchKernelGGL(k_leaf_apply, dim3(b.nleaf), dim3(64), 0, s, b.nleaf, (const leafd *)b.dL, (const int *)b.dlidx,
		                   (const double2 *)b.dD, (const double2 *)pc->wb, z + b.off);
	}
}

// Flexible restarted GMRES on the device: x = A^-1 b until a cycle's least-squares residual
// reaches tol ||b|| (the true residual b - A x is recomputed at every restart, as gmres_hmat
// does).  Returns total iterations, or -1.
int gmres_dev(hop &o, precond_t *pc, const double2 *b, double2 *x, double tol)
{
	typedef std::complex<double> zc;
	const int n = o.n;
	hipStream_t s = o.s;
	double2 *V = dal<double2>((size_t)(GM_M + 1) * n);
	double2 *Z = dal<double2>((size_t)GM_M * n);
	double2 *w = dal<double2>(n);
	double2 *r = dal<double2>(n);
	const double bnorm = std::sqrt(hdot(o, b, b, n).real());
	hipLaunchKernelGGL(k_zero, dim3(nb(n)), dim3(256), 0, s, x, (long)n);
	int total = 0, done = 0;   // iterations as the reference counts them (a failed cycle counts GM_M)
	if (bnorm <= 0) done = 1;
	const double tolabs = tol * bnorm;
	for (int cyc = 0; (cyc < GM_CYC) && !done; cyc++) {
		hop_apply(o, x, r);
		hipLaunchKernelGGL(k_sub, dim3(nb(n)), dim3(256), 0, s, b, r, n);
		const double beta = std::sqrt(hdot(o, r, r, n).real());
		if (g_htrace) fprintf(stderr, "[gmres_hmat] cycle %d: true residual %.3e of |b| after %d iterations\n", cyc,
		                      beta / bnorm, total);
		if (beta <= tolabs) {
			done = 1;
			break;
		}
		hipLaunchKernelGGL(k_scal, dim3(nb(n)), dim3(256), 0, s, (const double2 *)r, V, 1.0 / beta, n);
		std::vector<zc> H((size_t)(GM_M + 1) * GM_M), cs(GM_M), sn(GM_M), g(GM_M + 1);
		g[0] = beta;
		int j, kdone = -1;
		for (j = 0; j < GM_M; j++) {
			const double2 *vj = V + (long)j * n;
			double2 *zj = Z + (long)j * n;
			if (pc) pc_apply_dev(pc, vj, zj);
			else hipLaunchKernelGGL(k_copy, dim3(nb(n)), dim3(256), 0, s, vj, zj, n);
			hop_apply(o, zj, w);
			zc *Hc = &H[(size_t)j * (GM_M + 1)];
			const double hn = std::sqrt(orth_cgs2(o, V, j, w, Hc));
			Hc[j + 1] = hn;
			if (hn > 0) hipLaunchKernelGGL(k_scal, dim3(nb(n)), dim3(256), 0, s, (const double2 *)w, V + (long)(j + 1) * n, 1.0 / hn, n);
			for (int i = 0; i < j; i++) {
				const zc a = Hc[i], c = Hc[i + 1];
				Hc[i] = std::conj(cs[i]) * a + std::conj(sn[i]) * c;
				Hc[i + 1] = cs[i] * c - sn[i] * a;
			}
			const double rho = std::sqrt(std::norm(Hc[j]) + std::norm(Hc[j + 1]));
			if (rho <= 0) break;
			cs[j] = Hc[j] / rho;
			sn[j] = Hc[j + 1] / rho;
			Hc[j] = rho;
			Hc[j + 1] = 0;
			g[j + 1] = -sn[j] * g[j];
			g[j] = std::conj(cs[j]) * g[j];
			if (std::abs(g[j + 1]) <= tolabs) {
				kdone = j + 1;
				break;
			}
		}
		total += (kdone > 0) ? kdone : GM_M;
		const int k = (kdone > 0) ? kdone : std::min(j, GM_M);
		std::vector<zc> y(k);
		for (int i = k - 1; i >= 0; i--) {
			zc acc = g[i];
			for (int l = i + 1; l < k; l++) acc -= H[(size_t)l * (GM_M + 1) + i] * y[l];
			y[i] = acc / H[(size_t)i * (GM_M + 1) + i];
		}
		for (int l = 0; l < k; l++)
			hipLaunchKernelGGL(k_axpy, dim3(nb(n)), dim3(256), 0, s, x, Z + (long)l * n, make_double2(y[l].real(), y[l].imag()), n);
		// as the reference (iterative.c:336-342): a cycle whose least-squares residual reached the
		// tolerance ends the solve; the true residual is checked at the start of every restart
		if (kdone > 0) done = 1;
	}
	HIPCHK(hipStreamSynchronize(s));
	HIPCHK(hipFree(V));
	HIPCHK(hipFree(Z));
	HIPCHK(hipFree(w));
	HIPCHK(hipFree(r));
	return done ? std::max(1, total) : -1;
}

// ------------------------------------------------------------------------------------------
// preconditioner build (precond.c:139-399)

precond_t *pc_build_dev(const peec_t *p, const mna_sparse_t *sp, const hblock_t *blk, int nblk, hipStream_t s)
{
	const int n1 = p->offS, n = p->nunknown;
	precond_t *pc = new precond_t();
	pc->n = n;
	pc->n1 = n1;
	pc->nblk = nblk;
	pc->s = s;
	std::vector<int> ublk(n, -1), upos(n, -1);
	int maxn = 1;
	// leaves of every block, their cells, the D pool layout
	std::vector<std::vector<leafd>> L(nblk);
	std::vector<std::vector<int>> lidx(nblk), sleaf(nblk), spos(nblk);
	std::vector<std::vector<long>> lsrc(nblk);
	for (int b = 0; b < nblk; b++) {
		const hmat_t *h = blk[b].h;
		pcblk &pb = pc->blk[b];
		pb.off = blk[b].off;
		pb.n = blk[b].n;
		maxn = std::max(maxn, pb.n);
		for (int i = 0; i < pb.n; i++) {
			ublk[pb.off + i] = b;
			upos[pb.off + i] = i;
		}
		const hm_tree_t &t = h->t;
		pb.nleaf = t.nleaf;
		sleaf[b].assign(pb.n, -1);
		spos[b].assign(pb.n, -1);
		long doff = 0;
		for (int r = 0; r < t.nleaf; r++) {
			const int a = t.leafid[r];
			const int m = t.hi[a] - t.lo[a];
			if (m > PC_MAXM) {
				pc_free(pc);
				return nullptr;
			}
			long src = -1;
			for (int e = t.rnoff[r]; e < t.rnoff[r + 1]; e++) {
				const int q = t.rnidx[e];
				if ((t.nr[q] == a) && (t.nc2[q] == a)) src = h->near[q].off;
			}
			if (src < 0) {
				pc_free(pc);
				return nullptr;
			}
			leafd l{m, (int)lidx[b].size(), doff, 0, 0, 0};
			for (int q = 0; q < m; q++) {
				const int cell = t.perm[t.lo[a] + q];
				sleaf[b][cell] = r;
				spos[b][cell] = q;
				lidx[b].push_back(cell);
			}
			L[b].push_back(l);
			lsrc[b].push_back(src);
			doff += (long)m * m;
		}
		pb.dD = dal<double2>(doff);
	}
	// split the sparse part: S11 (dense init), B_b, C_b, same-leaf additions to D_b
	std::vector<double2> S11;
	std::vector<int> s11i, s11j;
	std::vector<std::vector<int>> bri(nblk), bci(nblk), cri(nblk), cci(nblk);
	std::vector<std::vector<double2>> bv(nblk), cv(nblk);
	std::vector<std::vector<std::vector<std::pair<int, double2>>>> ladd(nblk);
	for (int b = 0; b < nblk; b++) ladd[b].resize(pc->blk[b].nleaf);
	for (int e = 0; e < sp->nnz; e++) {
		const int i = sp->ri[e], j = sp->ci[e];
		const double2 v = make_double2(sp->val[e].r, sp->val[e].i);
		if ((i < n1) && (j < n1)) {
			s11i.push_back(i);
			s11j.push_back(j);
			S11.push_back(v);
		}
		else if (i < n1) {
			const int b = ublk[j];
			if (b < 0) continue;
			bri[b].push_back(i);
			bci[b].push_back(upos[j]);
			bv[b].push_back(v);
		}
		else if (j < n1) {
			const int b = ublk[i];
			if (b < 0) continue;
			cri[b].push_back(upos[i]);
			cci[b].push_back(j);
			cv[b].push_back(v);
		}
		else if ((ublk[i] >= 0) && (ublk[i] == ublk[j])) {
			const int b = ublk[i], k = upos[i], l = upos[j];
			if (sleaf[b][k] == sleaf[b][l]) {
				const int r = sleaf[b][k];
				ladd[b][r].push_back({spos[b][k] * L[b][r].m + spos[b][l], v});
			}
		}
	}
	int *dbad = dal<int>(1);
	HIPCHK(hipMemsetAsync(dbad, 0, sizeof(int), s));
	// leaf blocks: init, same-leaf additions, inverses
	std::vector<leafd *> dL(nblk);
	for (int b = 0; b < nblk; b++) {
		pcblk &pb = pc->blk[b];
		pb.dL = up(s, L[b]);
		pb.dlidx = up(s, lidx[b]);
		long *dsrc = up(s, lsrc[b]);
		hipLaunchKernelGGL(k_leaf_init, dim3(pb.nleaf), dim3(256), 0, s, (const leafd *)pb.dL, (const long *)dsrc,
		                   (const double2 *)blk[b].h->d_D, make_double2(blk[b].scale.r, blk[b].scale.i), pb.dD);
		std::vector<int> aoff(pb.nleaf + 1, 0), apos;
		std::vector<double2> aval;
		for (int r = 0; r < pb.nleaf; r++) {
			for (auto &pr : ladd[b][r]) {
				apos.push_back(pr.first);
				aval.push_back(pr.second);
			}
			aoff[r + 1] = (int)apos.size();
		}
		int *daoff = up(s, aoff), *dapos = up(s, apos);
		double2 *daval = up(s, aval);
		hipLaunchKernelGGL(k_leaf_add, dim3(nb(pb.nleaf)), dim3(256), 0, s, pb.nleaf, (const int *)daoff, (const int *)dapos,
		                   (const double2 *)daval, (const leafd *)pb.dL, pb.dD);
		hipLaunchKernelGGL(k_leaf_inv, dim3(pb.nleaf), dim3(2 * PC_MAXM), 0, s, (const leafd *)pb.dL, pb.dD, dbad);
		HIPCHK(hipStreamSynchronize(s));
		HIPCHK(hipFree(dsrc));
		HIPCHK(hipFree(daoff));
		HIPCHK(hipFree(dapos));
		HIPCHK(hipFree(daval));
		// B (CSR by circuit row) and C (CSR by block row), entry order kept
		std::vector<int> rp, col;
		std::vector<double2> val;
		csr_build(n1, bri[b], bci[b], bv[b], rp, col, val);
		pb.brp = up(s, rp);
		pb.bcol = up(s, col);
		pb.bval = up(s, val);
		csr_build(pb.n, cri[b], cci[b], cv[b], rp, col, val);
		pb.crp = up(s, rp);
		pb.ccol = up(s, col);
		pb.cval = up(s, val);
	}
	int bad = 0;
	HIPCHK(hipMemcpyAsync(&bad, dbad, sizeof(int), hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	HIPCHK(hipFree(dbad));
	if (bad) {
		pc_free(pc);
		return nullptr;
	}
	// Schur complement: S11 then, per circuit row, the leaf contributions in the reference's order
	double2 *S = dal<double2>((size_t)std::max(1, n1) * std::max(1, n1));
	hipLaunchKernelGGL(k_zero, dim3(nb((long)n1 * n1)), dim3(256), 0, s, S, (long)n1 * n1);
	{
		// duplicates of S11 entries are summed in their order on the host, then scattered
		std::vector<long> key(S11.size());
		for (size_t e = 0; e < S11.size(); e++) key[e] = (long)s11i[e] * n1 + s11j[e];
		std::vector<size_t> ord(S11.size());
		for (size_t e = 0; e < ord.size(); e++) ord[e] = e;
		std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return key[a] < key[b]; });
		std::vector<long> ukey;
		std::vector<double2> uval;
		for (size_t q = 0; q < ord.size(); q++) {
			const size_t e = ord[q];
			if (ukey.empty() || (ukey.back() != key[e])) {
				ukey.push_back(key[e]);
				uval.push_back(make_double2(0 + S11[e].x, 0 + S11[e].y));
			}
			else {
				uval.back() = make_double2(uval.back().x + S11[e].x, uval.back().y + S11[e].y);
			}
		}
		long *dk = up(s, ukey);
		double2 *dv = up(s, uval);
		if (!ukey.empty())
			hipLaunchKernelGGL(k_scatter, dim3(nb((long)ukey.size())), dim3(256), 0, s, (const long *)dk, (const double2 *)dv,
			                   (int)ukey.size(), S);
		HIPCHK(hipStreamSynchronize(s));
		HIPCHK(hipFree(dk));
		HIPCHK(hipFree(dv));
	}
	for (int b = 0; b < nblk; b++) {
		pcblk &pb = pc->blk[b];
		// per leaf: touched circuit unknowns (scan order), T = C_r (m x t)
		std::vector<std::vector<int>> cofrow(pb.n);
		for (size_t e = 0; e < cri[b].size(); e++) cofrow[cri[b][e]].push_back((int)e);
		std::vector<std::vector<int>> bofcol(pb.n);
		for (size_t e = 0; e < bri[b].size(); e++) bofcol[bci[b][e]].push_back((int)e);
		std::vector<int> tmap(std::max(1, n1), -1), tset;
		std::vector<double2> T;
		long xoff = 0;
		std::vector<std::vector<std::pair<int, std::pair<int, double2>>>> rows(n1);   // row -> (leaf, (cell, bval))
		for (int r = 0; r < pb.nleaf; r++) {
			leafd &l = L[b][r];
			const int *idx = &lidx[b][l.idx];
			const int t0 = (int)tset.size();
			int t = 0;
			for (int a = 0; a < l.m; a++)
				for (int e : cofrow[idx[a]]) {
					const int j = cci[b][e];
					if (tmap[j] < 0) {
						if (t >= PC_MAXT) {
							pc_free(pc);
							HIPCHK(hipFree(S));
							return nullptr;
						}
						tmap[j] = t;
						tset.push_back(j);
						t++;
					}
				}
			l.t = t;
			l.toff = t0;
			l.xoff = xoff;
			T.resize((size_t)(xoff + (long)l.m * t), make_double2(0, 0));
			for (int a = 0; a < l.m; a++)
				for (int e : cofrow[idx[a]]) {
					double2 &x = T[xoff + (long)a * t + tmap[cci[b][e]]];
					x = make_double2(x.x + cv[b][e].x, x.y + cv[b][e].y);
				}
			if (t > 0)
				for (int a = 0; a < l.m; a++)
					for (int e : bofcol[idx[a]]) rows[bri[b][e]].push_back({r, {a, bv[b][e]}});
			for (int c = 0; c < t; c++) tmap[tset[t0 + c]] = -1;
			xoff += (long)l.m * t;
		}
		HIPCHK(hipMemcpyAsync(pb.dL, L[b].data(), L[b].size() * sizeof(leafd), hipMemcpyHostToDevice, s));
		double2 *dT = up(s, T), *dX = dal<double2>(std::max(1L, xoff));
		int *dtset = up(s, tset);
		hipLaunchKernelGGL(k_leaf_x, dim3(pb.nleaf), dim3(256), 0, s, (const leafd *)pb.dL, (const double2 *)pb.dD,
		                   (const double2 *)dT, dX);
		std::vector<int> roff(n1 + 1, 0), rleaf, rcell;
		std::vector<double2> rval;
		for (int i = 0; i < n1; i++) {
			for (auto &q : rows[i]) {
				rleaf.push_back(q.first);
				rcell.push_back(q.second.first);
				rval.push_back(q.second.second);
			}
			roff[i + 1] = (int)rleaf.size();
		}
		int *droff = up(s, roff), *drleaf = up(s, rleaf), *drcell = up(s, rcell);
		double2 *drval = up(s, rval);
		if (n1 > 0)
			hipLaunchKernelGGL(k_schur_rows, dim3(n1), dim3(64), 0, s, n1, (const int *)droff, (const int *)drleaf,
			                   (const int *)drcell, (const double2 *)drval, (const leafd *)pb.dL, (const int *)dtset,
			                   (const double2 *)dX, S);
		HIPCHK(hipStreamSynchronize(s));
		void *tmp[] = {dT, dX, dtset, droff, drleaf, drcell, drval};
		for (void *x : tmp) HIPCHK(hipFree(x));
	}
	// S = LU; explicit inverses of its factors
	if (n1 > 0) {
		int *piv = dal<int>(n1);
		pc->perm = dal<int>(n1);
		const int info = dlu_factor(s, n1, (d_complex_t *)S, n1, piv, pc->perm, nullptr);
		HIPCHK(hipFree(piv));
		if (info >= 0) {
			HIPCHK(hipFree(S));
			pc_free(pc);
			return nullptr;
		}
		pc->Linv = dal<double2>((size_t)n1 * n1);
		pc->Uinv = dal<double2>((size_t)n1 * n1);
		const int nblk64 = (n1 + 63) / 64;
		double2 *scr = dal<double2>((size_t)((n1 + 1) / 2 + 64) * ((n1 + 1) / 2 + 64));
		double2 *dinv = dal<double2>((size_t)nblk64 * 64 * 64);
		dlu_tri_inverse(s, n1, (const d_complex_t *)S, n1, 1, (d_complex_t *)pc->Linv, (d_complex_t *)scr, (d_complex_t *)dinv);
		dlu_tri_inverse(s, n1, (const d_complex_t *)S, n1, 0, (d_complex_t *)pc->Uinv, (d_complex_t *)scr, (d_complex_t *)dinv);
		HIPCHK(hipStreamSynchronize(s));
		HIPCHK(hipFree(scr));
		HIPCHK(hipFree(dinv));
	}
	HIPCHK(hipFree(S));
	pc->w1 = dal<double2>(std::max(1, n1));
	pc->w2 = dal<double2>(std::max(1, n1));
	pc->wb = dal<double2>(maxn);
	pc->zb = dal<double2>(maxn);
	pc->d_b = dal<double2>(n);
	long nz = (long)n1 * n1 * 2;
	for (int b = 0; b < nblk; b++) nz += L[b].empty() ? 0 : (L[b].back().doff + (long)L[b].back().m * L[b].back().m);
	pc->mb = (double)nz * sizeof(d_complex_t) / (1024.0 * 1024.0);
	return pc;
}

// ------------------------------------------------------------------------------------------
// the reference's C interface

// ------------------------------------------------------------------------------------------
// learned near-field correction of the preconditioner
//
// P^-1 = M^-1 - theta M^-1 N M^-1, with M the preconditioner of precond.c (leaf blocks and the
// circuit Schur complement, possibly built at another frequency of the sweep), N the near-field
// blocks of the compressed kernel matrices off their block diagonal (scaled as in the system:
// -jw for Lp, 1 for P) at the current frequency, and theta a complex weight predicted by the
// network (mlp.hip) from features of the system.  theta = 0 is precond.c's method.  GMRES's
// true residual |b - A x| <= tol |b| decides convergence, so theta changes iteration counts
// but not the answer; a solve that does not converge is redone with a fresh M and theta = 0.

namespace {

}  // namespace

// k-factored records of the band [f0, f1] for the compressed builds (entries by one Horner each);
// PEEC_KF=0 keeps the element code
extern "C" void pgpu_kf_band(pgpu_t *g, const peec_t *p, double f0, double f1)
{
	const char *e = getenv("PEEC_KF");
	if ((e && e[0] == '0') || !p->retardation) return;
	const double kr = 2 * PI / C0;
	pgpu_kf_setup(g, kr * std::min(f0, f1), kr * std::max(f0, f1));
}

namespace {

// the network's parameters set by the caller (empty: theta = 0)
std::vector<double> g_lpc_model;

// y = N x over the kernel blocks (rows outside them 0)
void near_apply(hop &o, const double2 *x, double2 *y)
{
	hipLaunchKernelGGL(k_zero, dim3(nb(o.n)), dim3(256), 0, o.s, y, (long)o.n);
	for (int b = 0; b < o.nblk; b++) {
		if ((o.h[b] == nullptr) || (o.bn[b] <= 0)) continue;
		hm_near_off_dev(o.h[b], x + o.off[b], o.work);
		hipLaunchKernelGGL(k_addscaled, dim3(nb(o.bn[b])), dim3(256), 0, o.s, y + o.off[b], (const double2 *)o.work, o.bn[b],
		                   o.scale[b]);
	}
}

struct lpc_ws {
	double2 *t = nullptr, *w = nullptr;
	int n = 0;
	void ensure(int nn)
	{
		if (nn <= n) return;
		if (t) HIPCHK(hipFree(t));
		if (w) HIPCHK(hipFree(w));
		t = dal<double2>(nn);
		w = dal<double2>(nn);
		n = nn;
	}
	~lpc_ws()
	{
		if (t) HIPCHK(hipFree(t));
		if (w) HIPCHK(hipFree(w));
	}
};

// z = (M^-1 - theta M^-1 N M^-1) v
void lpc_apply(hop &o, precond_t *pc, std::complex<double> theta, lpc_ws &ws, const double2 *v, double2 *z)
{
	pc_apply_dev(pc, v, z);
	if (theta == std::complex<double>(0, 0)) return;
	near_apply(o, z, ws.t);
	pc_apply_dev(pc, ws.t, ws.w);
	hipLaunchKernelGGL(k_axpy, dim3(nb(o.n)), dim3(256), 0, o.s, z, (const double2 *)ws.w,
	                   make_double2(-theta.real(), -theta.imag()), o.n);
}

// Right-preconditioned restarted GMRES(GM_M) with a fixed preconditioner, ended by the true
// residual |b - A x| <= tol |b| checked at every restart (at most maxcyc cycles).  hist (may be
// null) receives copies of the first cycle's Arnoldi vectors (capacity GM_M), *nhist their
// number.  Returns the iterations, or -1 without convergence.
int gmres_fn(hop &o, const std::function<void(const double2 *, double2 *)> &prec, const double2 *b, double2 *x, double tol,
             int maxcyc, double2 *hist, int *nhist)
{
	typedef std::complex<double> zc;
	const int n = o.n;
	hipStream_t s = o.s;
	double2 *V = dal<double2>((size_t)(GM_M + 1) * n);
	double2 *Z = dal<double2>((size_t)GM_M * n);
	double2 *w = dal<double2>(n);
	double2 *r = dal<double2>(n);
	const double bnorm = std::sqrt(hdot(o, b, b, n).real());
	hipLaunchKernelGGL(k_zero, dim3(nb(n)), dim3(256), 0, s, x, (long)n);
	if (nhist) *nhist = 0;
	int total = 0, done = (bnorm <= 0);
	const double tolabs = tol * bnorm;
	for (int cyc = 0; (cyc <= maxcyc) && !done; cyc++) {
		hop_apply(o, x, r);
		hipLaunchKernelGGL(k_sub, dim3(nb(n)), dim3(256), 0, s, b, r, n);
		const double beta = std::sqrt(hdot(o, r, r, n).real());
		if (beta <= tolabs) {
			done = 1;
			break;
		}
		if (cyc == maxcyc) break;
		hipLaunchKernelGGL(k_scal, dim3(nb(n)), dim3(256), 0, s, (const double2 *)r, V, 1.0 / beta, n);
		std::vector<zc> H((size_t)(GM_M + 1) * GM_M), cs(GM_M), sn(GM_M), g(GM_M + 1);
		g[0] = beta;
		int j, k = 0;
		for (j = 0; j < GM_M; j++) {
			const double2 *vj = V + (long)j * n;
			double2 *zj = Z + (long)j * n;
			prec(vj, zj);
			hop_apply(o, zj, w);
			zc *Hc = &H[(size_t)j * (GM_M + 1)];
			const double hn = std::sqrt(orth_cgs2(o, V, j, w, Hc));
			Hc[j + 1] = hn;
			if (hn > 0) hipLaunchKernelGGL(k_scal, dim3(nb(n)), dim3(256), 0, s, (const double2 *)w, V + (long)(j + 1) * n, 1.0 / hn, n);
			for (int i = 0; i < j; i++) {
				const zc a = Hc[i], c = Hc[i + 1];
				Hc[i] = std::conj(cs[i]) * a + std::conj(sn[i]) * c;
				Hc[i + 1] = cs[i] * c - sn[i] * a;
			}
			const double rho = std::sqrt(std::norm(Hc[j]) + std::norm(Hc[j + 1]));
			if (rho <= 0) break;
			cs[j] = Hc[j] / rho;
			sn[j] = Hc[j + 1] / rho;
			Hc[j] = rho;
			Hc[j + 1] = 0;
			g[j + 1] = -sn[j] * g[j];
			g[j] = std::conj(cs[j]) * g[j];
			k = j + 1;
			total++;
			if ((std::abs(g[j + 1]) <= tolabs) || (hn <= 0)) break;
		}
		if ((cyc == 0) && hist && nhist) {
			HIPCHK(hipMemcpyAsync(hist, V, (size_t)k * n * sizeof(double2), hipMemcpyDeviceToDevice, s));
			*nhist = k;
		}
		std::vector<zc> y(k);
		for (int i = k - 1; i >= 0; i--) {
			zc acc = g[i];
			for (int l = i + 1; l < k; l++) acc -= H[(size_t)l * (GM_M + 1) + i] * y[l];
			y[i] = acc / H[(size_t)i * (GM_M + 1) + i];
		}
		for (int l = 0; l < k; l++)
			hipLaunchKernelGGL(k_axpy, dim3(nb(n)), dim3(256), 0, s, x, Z + (long)l * n, make_double2(y[l].real(), y[l].imag()), n);
	}
	HIPCHK(hipStreamSynchronize(s));
	HIPCHK(hipFree(V));
	HIPCHK(hipFree(Z));
	HIPCHK(hipFree(w));
	HIPCHK(hipFree(r));
	return done ? std::max(1, total) : -1;
}

// the system's compressed blocks, sparse part and operator at one frequency
struct hsys {
	double f = 0;
	hmat_t *hl = nullptr, *hp = nullptr;
	hblock_t blk[HBLK_MAX];
	int nblk = 0;
	mna_sparse_t sp{};
	hop o;
	bool ok = false;
	void build(peec_t *p, double freq, hipStream_t s)
	{
		f = freq;
		nblk = 0;
		hl = hmat_build(p, f, nullptr);
		if (hl == nullptr) return;
		blk[nblk++] = hblock_t{p->offS, p->nseg, hl, d_complex(0, -2 * PI * f)};
		if (p->capacitance && (p->ncell > 0)) {
			hp = hmat_build_pot(p, f, nullptr);
			if (hp == nullptr) return;
			blk[nblk++] = hblock_t{p->offQ, p->ncell, hp, d_complex(1, 0)};
		}
		mna_sparse(p, f, &sp);
		hop_init(o, p, &sp, blk, nblk, s);
		ok = true;
	}
	~hsys()
	{
		if (ok) hop_free(o);
		if (sp.ri) mna_sparse_free(&sp);
		hmat_free(hl);
		hmat_free(hp);
	}
};

// features of the system at f with the preconditioner built at fref
void lpc_features(const peec_t *p, const hsys &S, double fref, double *x)
{
	double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
	for (int i = 0; i < p->nseg; i++)
		for (int c = 0; c < 3; c++) {
			const double v = 0.5 * (p->seg[i].x1[c] + p->seg[i].x2[c]);
			lo[c] = std::min(lo[c], v);
			hi[c] = std::max(hi[c], v);
		}
	double D = 0;
	for (int c = 0; c < 3; c++) D += (hi[c] - lo[c]) * (hi[c] - lo[c]);
	D = std::sqrt(D);
	const double kr = p->retardation ? 2 * PI / C0 : 0.0;
	const double k = kr * S.f, k0 = kr * fref;
	double off = 0, dg = 1, offp = 0, dgp = 1;
	if (S.hl) hm_near_norms(S.hl, &off, &dg);
	if (S.hp) hm_near_norms(S.hp, &offp, &dgp);
	x[0] = k * D;
	x[1] = k0 * D;
	x[2] = (k - k0) * D;
	x[3] = k * D / std::sqrt((double)std::max(1, p->nseg));
	x[4] = (dg > 0) ? off / dg : 0;
	x[5] = (S.hp && (dgp > 0)) ? offp / dgp : 0;
	x[6] = std::log10((double)std::max(1, p->nunknown)) / 5;
	x[7] = (double)(p->capacitance ? p->ncell : 0) / std::max(1, p->nunknown);
}

std::complex<double> lpc_theta(hipStream_t s, const double *x)
{
	// PEEC_LPC_THETA="re,im" fixes theta (comparisons against the network's choice)
	const char *fix = getenv("PEEC_LPC_THETA");
	if (fix != nullptr) {
		double re = 0, im = 0;
		if (sscanf(fix, "%lf,%lf", &re, &im) >= 1) return {re, im};
	}
	if (g_lpc_model.size() != (size_t)MLP_NP) return {0, 0};
	double *dW = dal<double>(MLP_NP), *dX = dal<double>(MLP_NF), *dt = dal<double>(2);
	HIPCHK(hipMemcpyAsync(dW, g_lpc_model.data(), MLP_NP * sizeof(double), hipMemcpyHostToDevice, s));
	HIPCHK(hipMemcpyAsync(dX, x, MLP_NF * sizeof(double), hipMemcpyHostToDevice, s));
	mlp_forward_dev(s, dW, dX, 1, dt);
	double t[2];
	HIPCHK(hipMemcpyAsync(t, dt, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	for (double *q : {dW, dX, dt}) HIPCHK(hipFree(q));
	// the network was trained on labels in [LPC_GRID[0], LPC_GRID[last]] (real): no extrapolation
	return {std::min(std::max(t[0], 0.0), 0.6), std::min(std::max(t[1], -0.1), 0.1)};
}

}  // namespace

extern "C" {

void mna_sparse(const peec_t *p, double f, mna_sparse_t *sp)
{
	stamps_t st;
	mna_stamps(p, f, &st);
	sp->ri = st.ri;
	sp->ci = st.ci;
	sp->val = st.val;
	sp->nnz = st.n;
	sp->cap = st.n;
}

void mna_sparse_free(mna_sparse_t *sp)
{
	free(sp->ri);
	free(sp->ci);
	free(sp->val);
	memset(sp, 0, sizeof(*sp));
}

void mna_apply(const peec_t *p, const mna_sparse_t *sp, const hblock_t *blk, int nblk, const d_complex_t *x,
               d_complex_t *y, d_complex_t *work)
{
	(void)work;
	hipStream_t s = pgpu_hip_stream(pgpu_get(p));
	hop o;
	hop_init(o, p, sp, blk, nblk, s);
	double2 *dx = dal<double2>(o.n), *dy = dal<double2>(o.n);
	HIPCHK(hipMemcpyAsync(dx, x, (size_t)o.n * sizeof(double2), hipMemcpyHostToDevice, s));
	hop_apply(o, dx, dy);
	HIPCHK(hipMemcpyAsync(y, dy, (size_t)o.n * sizeof(double2), hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	HIPCHK(hipFree(dx));
	HIPCHK(hipFree(dy));
	hop_free(o);
}

struct precond_t *pc_build(const peec_t *p, const mna_sparse_t *sp, const hblock_t *blk, int nblk)
{
	return pc_build_dev(p, sp, blk, nblk, pgpu_hip_stream(pgpu_get(p)));
}

void pc_free(struct precond_t *pc)
{
	if (pc == nullptr) return;
	if (pc->s) HIPCHK(hipStreamSynchronize(pc->s));
	for (int b = 0; b < HBLK_MAX; b++) {
		pcblk &x = pc->blk[b];
		void *bb[] = {x.dL, x.dlidx, x.dD, x.brp, x.bcol, x.bval, x.crp, x.ccol, x.cval};
		for (void *q : bb)
			if (q) HIPCHK(hipFree(q));
	}
	void *bufs[] = {pc->Linv, pc->Uinv, pc->perm, pc->w1, pc->w2, pc->wb, pc->zb, pc->d_b};
	for (void *q : bufs)
		if (q) HIPCHK(hipFree(q));
	delete pc;
}

void pc_apply(const struct precond_t *pc_, d_complex_t *b)
{
	precond_t *pc = const_cast<precond_t *>(pc_);
	double2 *z = dal<double2>(pc->n);
	HIPCHK(hipMemcpyAsync(pc->d_b, b, (size_t)pc->n * sizeof(double2), hipMemcpyHostToDevice, pc->s));
	pc_apply_dev(pc, pc->d_b, z);
	HIPCHK(hipMemcpyAsync(b, z, (size_t)pc->n * sizeof(double2), hipMemcpyDeviceToHost, pc->s));
	HIPCHK(hipStreamSynchronize(pc->s));
	HIPCHK(hipFree(z));
}

double pc_memory_mb(const struct precond_t *pc)
{
	return pc ? pc->mb : 0;
}

int gmres_hmat(const peec_t *p, const mna_sparse_t *sp, const hblock_t *blk, int nblk, const struct precond_t *pc,
               d_complex_t *b, double tol)
{
	hipStream_t s = pgpu_hip_stream(pgpu_get(p));
	hop o;
	hop_init(o, p, sp, blk, nblk, s);
	double2 *db = dal<double2>(o.n), *dx = dal<double2>(o.n);
	HIPCHK(hipMemcpyAsync(db, b, (size_t)o.n * sizeof(double2), hipMemcpyHostToDevice, s));
	const int it = gmres_dev(o, const_cast<precond_t *>(pc), db, dx, tol);
	if (it > 0) HIPCHK(hipMemcpyAsync(b, dx, (size_t)o.n * sizeof(double2), hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	HIPCHK(hipFree(db));
	HIPCHK(hipFree(dx));
	hop_free(o);
	return it;
}

// One point of the reference's solve() on the compressed path (solve.c:246-300, 319-405 with
// compress = 1), device resident: the H matrices of Lp (and P), the sparse part and the
// preconditioner at this frequency, then one GMRES (true residual 1e-10, gmres_hmat) per
// right-hand side: rhs holds nr host vectors of n, x receives their solutions.  its (may be NULL)
// accumulates GMRES iterations, mem_mb (may be NULL) receives the point's H-matrix and
// preconditioner memory (hmat_memory_mb + pc_memory_mb).  Returns 0, 5 (H-matrix build), 6
// (preconditioner) or 4 (GMRES did not converge).
int pgpu_solve_point_compressed(pgpu_t *g, peec_t *p, int ifreq, int nr, const d_complex_t *rhs, d_complex_t *x, long *its,
                                double *mem_mb, pgpu_stats_t *st)
{
	hipStream_t s = pgpu_hip_stream(g);
	const int n = p->nunknown;
	const double f = freq_at(p, ifreq);
	const auto t0 = std::chrono::steady_clock::now();
	hblock_t blk[HBLK_MAX];
	int nblk = 0;
	hmat_t *hl = hmat_build(p, f, nullptr);
	if (hl == nullptr) {
		printf("*** compression : Lp H-matrix build failed at f = %.5e Hz\n", f);
		return 5;
	}
	blk[nblk++] = hblock_t{p->offS, p->nseg, hl, d_complex(0, -2 * PI * f)};
	hmat_t *hp = nullptr;
	if (p->capacitance && (p->ncell > 0)) {
		hp = hmat_build_pot(p, f, nullptr);
		if (hp == nullptr) {
			printf("*** compression : P H-matrix build failed at f = %.5e Hz\n", f);
			hmat_free(hl);
			return 5;
		}
		blk[nblk++] = hblock_t{p->offQ, p->ncell, hp, d_complex(1, 0)};
	}
	const auto t1 = std::chrono::steady_clock::now();
	mna_sparse_t sp;
	mna_sparse(p, f, &sp);
	precond_t *pc = pc_build_dev(p, &sp, blk, nblk, s);
	const auto t2 = std::chrono::steady_clock::now();
	int rc = 0;
	if (pc == nullptr) {
		printf("*** compression : preconditioner failed at f = %.5e Hz (floating node?)\n", f);
		rc = 6;
	}
	hop o;
	if (!rc) hop_init(o, p, &sp, blk, nblk, s);
	double2 *db = dal<double2>(n), *dx = dal<double2>(n);
	for (int r = 0; (r < nr) && !rc; r++) {
		HIPCHK(hipMemcpyAsync(db, rhs + (size_t)r * n, (size_t)n * sizeof(double2), hipMemcpyHostToDevice, s));
		const int it = gmres_dev(o, pc, db, dx, 1e-10);
		if (it < 0) {
			printf("*** compression : GMRES did not converge at f = %.5e Hz\n", f);
			rc = 4;
			break;
		}
		if (its) *its += it;
		if (st) st->gmres_its += it;
		HIPCHK(hipMemcpyAsync(x + (size_t)r * n, dx, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
	}
	const auto t3 = std::chrono::steady_clock::now();
	if (st) {
		st->fill_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
		st->lu_ms += std::chrono::duration<double, std::milli>(t2 - t1).count();
		st->solve_ms += std::chrono::duration<double, std::milli>(t3 - t2).count();
		st->total_ms += std::chrono::duration<double, std::milli>(t3 - t0).count();
	}
	if (mem_mb) *mem_mb = hmat_memory_mb(hl) + hmat_memory_mb(hp) + pc_memory_mb(pc);
	HIPCHK(hipFree(db));
	HIPCHK(hipFree(dx));
	if (pc != nullptr) hop_free(o);
	pc_free(pc);
	mna_sparse_free(&sp);
	hmat_free(hl);
	hmat_free(hp);
	return rc;
}

// Compressed sweep (solve.c:246-300 with compress = 1) of the ports alone: Z of every point.
int pgpu_sweep_compressed(pgpu_t *g, peec_t *p, int ifreq0, int nfreq, d_complex_t *zmat, pgpu_stats_t *st)
{
	{
		const auto t0 = std::chrono::steady_clock::now();
		pgpu_kf_band(g, p, freq_at(p, ifreq0), freq_at(p, ifreq0 + nfreq - 1));
		if (st) {
			st->kf_setup_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
			st->fill_ms += st->kf_setup_ms;
			st->total_ms += st->kf_setup_ms;
		}
	}
	const int np = p->nport, n = p->nunknown;
	std::vector<d_complex_t> rhs((size_t)n * np), x((size_t)n * np);
	for (int j = 0; j < np; j++) mna_rhs_port(p, j, &rhs[(size_t)j * n]);
	for (int k = 0; k < nfreq; k++) {
		const int rc = pgpu_solve_point_compressed(g, p, ifreq0 + k, np, rhs.data(), x.data(), nullptr, nullptr, st);
		if (rc) return rc;
		for (int j = 0; j < np; j++)
			for (int i = 0; i < np; i++) {
				const int i1 = p->nodemap[p->port[i].n1], i2 = p->nodemap[p->port[i].n2];
				const d_complex_t v1 = (i1 < 0) ? d_complex(0, 0) : x[(size_t)j * n + i1];
				const d_complex_t v2 = (i2 < 0) ? d_complex(0, 0) : x[(size_t)j * n + i2];
				zmat[((size_t)k * np + i) * np + j] = d_sub(v1, v2);
			}
	}
	return 0;
}


void peec_lpc_set_model(const double *W)
{
	if (W == nullptr) g_lpc_model.clear();
	else g_lpc_model.assign(W, W + MLP_NP);
}

int peec_lpc_get_model(double *W)
{
	if (g_lpc_model.size() != (size_t)MLP_NP) return 0;
	memcpy(W, g_lpc_model.data(), MLP_NP * sizeof(double));
	return 1;
}

// Training samples from the solver: for each pair (fref[i], f[i]) the preconditioner of
// precond.c is built at fref and the system at f is solved for port 1 by GMRES (true residual
// 1e-8) with the correction at every theta of LPC_GRID; the label theta* is the one with the
// fewest iterations (the smallest on ties).  X[8 i ..] are the features of the system at f with
// fref; coef[4 i ..] = (-theta*, 0, 1, 1 + theta*^2) puts the sample's loss
// (d + 2 Re(theta p) + |theta|^2 c) / d = (1 + |theta - theta*|^2) / (1 + theta*^2) in the
// network's quadratic form; its (may be NULL) receives the iteration counts, LPC_NGRID per pair.
// Returns 0, or 1 when a build fails.
constexpr int LPC_NGRID = 7;
constexpr double LPC_GRID[LPC_NGRID] = {0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6};

int peec_lpc_samples(peec_t *p, const double *fref, const double *f, int npairs, double *X, double *coef, int *its)
{
	if (!p->compress) return 1;
	pgpu_t *g = pgpu_get(p);
	hipStream_t s = pgpu_hip_stream(g);
	const int n = p->nunknown;
	{
		double lo = 1e300, hi = 0;
		for (int k = 0; k < npairs; k++) {
			lo = std::min(lo, std::min(fref[k], f[k]));
			hi = std::max(hi, std::max(fref[k], f[k]));
		}
		if (npairs > 0) pgpu_kf_band(g, p, lo, hi);
	}
	std::vector<d_complex_t> hb(n);
	double2 *db = dal<double2>(n), *dx = dal<double2>(n);
	lpc_ws ws;
	ws.ensure(n);
	int rc = 0;
	for (int k = 0; (k < npairs) && !rc; k++) {
		precond_t *pc = nullptr;
		{
			hsys R;
			R.build(p, fref[k], s);
			if (!R.ok) {
				printf("*** learned preconditioner samples: compressed build failed at f = %.5e Hz\n", fref[k]);
				rc = 1;
				break;
			}
			pc = pc_build_dev(p, &R.sp, R.blk, R.nblk, s);
		}
		if (pc == nullptr) {
			printf("*** learned preconditioner samples: preconditioner failed at f = %.5e Hz\n", fref[k]);
			rc = 1;
			break;
		}
		hsys S;
		S.build(p, f[k], s);
		if (!S.ok) {
			pc_free(pc);
			rc = 1;
			break;
		}
		mna_rhs_port(p, 0, hb.data());
		HIPCHK(hipMemcpyAsync(db, hb.data(), (size_t)n * sizeof(double2), hipMemcpyHostToDevice, s));
		int best = -1, bestits = 0;
		for (int q = 0; q < LPC_NGRID; q++) {
			const std::complex<double> th(LPC_GRID[q], 0);
			int it = gmres_fn(S.o, [&](const double2 *v, double2 *z) { lpc_apply(S.o, pc, th, ws, v, z); }, db, dx, 1e-8, GM_CYC,
			                  nullptr, nullptr);
			if (it < 0) it = (GM_CYC + 1) * GM_M;
			if (its) its[(long)k * LPC_NGRID + q] = it;
			if ((best < 0) || (it < bestits)) {
				best = q;
				bestits = it;
			}
		}
		const double t = LPC_GRID[best];
		lpc_features(p, S, fref[k], X + (long)MLP_NF * k);
		coef[4 * k] = -t;
		coef[4 * k + 1] = 0;
		coef[4 * k + 2] = 1;
		coef[4 * k + 3] = 1 + t * t;
		pc_free(pc);
	}
	HIPCHK(hipFree(db));
	HIPCHK(hipFree(dx));
	return rc;
}

// Compressed sweep with the learned correction: the preconditioner is built at a reference
// frequency and kept while GMRES needs at most max(1.5 x, +10) the iterations of the first solve
// after the build; theta from the network (0 without a model: precond.c's method under the same
// rule); tolerance tol on the true residual; a solve that fails is redone with a fresh
// preconditioner and theta = 0.  The schedule depends on iteration counts only, so the
// preconditioner used at each point (and Z at the tolerance level) is the same on every run.
int pgpu_sweep_learned(pgpu_t *g, peec_t *p, int ifreq0, int nfreq, d_complex_t *zmat, pgpu_stats_t *st, double tol)
{
	hipStream_t s = pgpu_hip_stream(g);
	{
		const auto t0 = std::chrono::steady_clock::now();
		pgpu_kf_band(g, p, freq_at(p, ifreq0), freq_at(p, ifreq0 + nfreq - 1));
		if (st) {
			st->kf_setup_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
			st->fill_ms += st->kf_setup_ms;
			st->total_ms += st->kf_setup_ms;
		}
	}
	const int np = p->nport, n = p->nunknown;
	std::vector<d_complex_t> hb(n);
	std::vector<double2> hx(n);
	double2 *db = dal<double2>(n), *dx = dal<double2>(n);
	lpc_ws ws;
	ws.ensure(n);
	precond_t *pc = nullptr;
	double fref = 0;
	int its_fresh = -1;
	bool rebuild = true;
	int rc = 0;
	for (int k = 0; (k < nfreq) && !rc; k++) {
		const double f = freq_at(p, ifreq0 + k);
		const auto t0 = std::chrono::steady_clock::now();
		hsys S;
		S.build(p, f, s);
		if (!S.ok) {
			rc = 5;
			break;
		}
		const auto t1 = std::chrono::steady_clock::now();
		if (rebuild || (pc == nullptr)) {
			if (pc) pc_free(pc);
			pc = pc_build_dev(p, &S.sp, S.blk, S.nblk, s);
			if (pc == nullptr) {
				printf("*** compression : preconditioner failed at f = %.5e Hz (floating node?)\n", f);
				rc = 6;
				break;
			}
			fref = f;
			its_fresh = -1;
			rebuild = false;
			if (st) st->pc_builds++;
		}
		double x[MLP_NF];
		lpc_features(p, S, fref, x);
		const std::complex<double> theta = lpc_theta(s, x);
		if (st) st->theta_abs_sum += std::abs(theta);
		const auto t2 = std::chrono::steady_clock::now();
		for (int j = 0; (j < np) && !rc; j++) {
			mna_rhs_port(p, j, hb.data());
			HIPCHK(hipMemcpyAsync(db, hb.data(), (size_t)n * sizeof(double2), hipMemcpyHostToDevice, s));
			int it = gmres_fn(S.o, [&](const double2 *v, double2 *z) { lpc_apply(S.o, pc, theta, ws, v, z); }, db, dx, tol,
			                  GM_CYC, nullptr, nullptr);
			if (it < 0) {
				// fallback: precond.c's preconditioner at this frequency, theta = 0
				if (st) st->lpc_fallbacks++;
				pc_free(pc);
				pc = pc_build_dev(p, &S.sp, S.blk, S.nblk, s);
				if (pc == nullptr) {
					rc = 6;
					break;
				}
				if (st) st->pc_builds++;
				fref = f;
				its_fresh = -1;
				it = gmres_fn(S.o, [&](const double2 *v, double2 *z) { pc_apply_dev(pc, v, z); }, db, dx, tol, GM_CYC, nullptr,
				              nullptr);
				if (it < 0) {
					printf("*** GMRES did not converge at f = %.5e Hz\n", f);
					rc = 4;
					break;
				}
			}
			if (st) st->gmres_its += it;
			if (its_fresh < 0) its_fresh = it;
			else if (it > std::max(3 * its_fresh / 2, its_fresh + 10)) rebuild = true;
			HIPCHK(hipMemcpyAsync(hx.data(), dx, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost, s));
			HIPCHK(hipStreamSynchronize(s));
			for (int i = 0; i < np; i++) {
				const int i1 = p->nodemap[p->port[i].n1], i2 = p->nodemap[p->port[i].n2];
				const double2 v1 = (i1 < 0) ? make_double2(0, 0) : hx[i1];
				const double2 v2 = (i2 < 0) ? make_double2(0, 0) : hx[i2];
				zmat[((size_t)k * np + i) * np + j] = d_complex(v1.x - v2.x, v1.y - v2.y);
			}
		}
		const auto t3 = std::chrono::steady_clock::now();
		if (st) {
			st->fill_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
			st->lu_ms += std::chrono::duration<double, std::milli>(t2 - t1).count();
			st->solve_ms += std::chrono::duration<double, std::milli>(t3 - t2).count();
			st->total_ms += std::chrono::duration<double, std::milli>(t3 - t0).count();
		}
	}
	if (pc) pc_free(pc);
	HIPCHK(hipFree(db));
	HIPCHK(hipFree(dx));
	return rc;
}

}  // extern "C"
