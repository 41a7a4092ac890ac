// hmat.hip — H-matrix compression of Lp and P on the device (compression = 1).
//
// The block structure (cluster tree, far/near blocks, leaf-row grouping) is the reference's
// (hmatrix.c:43-227, 483-586) and is built on the host (hmat_host.c).  Everything that touches
// matrix entries runs here:
//   * ACA, one 256-thread workgroup per far block (k_aca): the reference's partial-pivot
//     adaptive cross approximation with its restarts from unrepresented rows (aca_block,
//     hmatrix.c:250-395), step for step, plus a column check where it stops (column_check:
//     it only acts on blocks left far outside the tolerance, so elsewhere the ranks and
//     factors are the reference's); the entry evaluations of a residual row/column and
//     the norms are spread over the workgroup, reductions run in a fixed tree order, pivot
//     searches keep the reference's first-extremum rule;
//   * near blocks and far blocks that did not compress are evaluated densely (k_hm_dense);
//   * when the whole matrix fits in a quarter of free device memory, every entry is evaluated
//     first, in bulk, and ACA and the dense blocks read them from there: over a sweep by the
//     k-factored fill of the band (records read in slot order, P aggregated from the sub-cell
//     matrix in the order hm_entry sums), otherwise one thread per entry (k_hm_full).  The
//     entries are the same numbers either way (same device arithmetic, same summation order);
//     what changes is that the costly part of an entry (a Horner per sub-cell pair, or the
//     element code) runs at full occupancy instead of inside ACA's sequential rank steps;
//   * y = H x in two stages as the reference's hmat_matvec (hmatrix.c:680-778): t_b = V_b^T x
//     per far block, then one workgroup per leaf row cluster adds U_b t_b and the dense rows of
//     its blocks in the registered order, so every output row is written by one workgroup.
// Entries come from the same device code as the dense fills (lp_entry / the aggregated
// pot_entry rule, fill_entries.hpp), so compressed and dense paths see identical numbers.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>
#include "elements.hpp"
#include "fill_entries.hpp"
#include "fill_kernels.hpp"
#include "hip_check.hpp"
#include "kf_eval.hpp"
#include "hmat_dev.hpp"

namespace pe {

// one entry of the kernel matrix in original numbering: Lp (kind 0) or P (kind 1, summed over
// the sub-cells of both capacitive cells, pot_entry potential.c:124-150)
// An entry from the k-factored records of the sweep band when they are given and the pair has
// one (one Horner per pair), else from the element code.
template <int KM>
__device__ cplx hm_entry(const hent &E, int i, int j)
{
	if (E.kind == 0) {
		if (i > j) {
			const int t = i;
			i = j;
			j = t;
		}
		if (E.rec != nullptr) {
			const long sl = kf_slot(i, j, E.f.nt);
			if (kf_has(E.rec, E.nslots, sl)) {
				double vr, vi;
				kf_value(E.rec, E.nslots, sl, E.T, E.f.kw, vr, vi);
				return mk(vr, vi);
			}
		}
		return lp_entry<KM>(E.f.S, E.f.pS, E.f.SM, E.f.pSM, i, j, E.f.kw, E.f.coef, E.f.gp);
	}
	double sr = 0, si = 0;
	for (int p = E.csoff[i]; p < E.csoff[i + 1]; p++) {
		const int h1 = E.csidx[p];
		for (int q = E.csoff[j]; q < E.csoff[j + 1]; q++) {
			const int h2 = E.csidx[q];
			const int lo = (h1 <= h2) ? h1 : h2, hi = (h1 <= h2) ? h2 : h1;
			cplx v;
			const long sl = (E.rec != nullptr) ? kf_slot(lo, hi, E.f.nt) : 0;
			if ((E.rec != nullptr) && kf_has(E.rec, E.nslots, sl)) {
				double vr, vi;
				kf_value(E.rec, E.nslots, sl, E.T, E.f.kw, vr, vi);
				v = mk(vr, vi);
			}
			else {
				v = ig_entry<KM>(E.f.S, E.f.pS, E.f.SM, E.f.pSM, lo, hi, E.f.kw, E.f.gp);
			}
			sr = sr + v.r;
			si = si + v.i;
		}
	}
	const double s = 1 / (E.f.k4pe * E.f.carea[i] * E.f.carea[j]);
	return mk(s * sr, s * si);
}

// entry at permuted positions (a, b): A(perm[a], perm[b]), from the bulk-evaluated matrix when
// there is one
template <int KM>
__device__ __forceinline__ cplx hm_pos(const hent &E, const int *perm, int a, int b)
{
	if (E.full != nullptr) {
		const double2 v = E.full[(long)perm[a] * E.nfull + perm[b]];
		return mk(v.x, v.y);
	}
	return hm_entry<KM>(E, perm[a], perm[b]);
}

constexpr int AT = 256;

// Workgroup reductions: a butterfly inside each wave (every lane ends with the same bits: each
// step adds the same two partial values, a + b = b + a), then the four wave results combined
// in a fixed order by every thread: two barriers per reduction instead of a 256-wide tree.
constexpr int NW = AT / 64;
static_assert(NW == 4, "the wave results are combined as four");

__device__ __forceinline__ double wg_sum(double x, double *red)
{
	for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
	if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
	__syncthreads();
	const double r = (red[0] + red[1]) + (red[2] + red[3]);
	__syncthreads();
	return r;
}

__device__ __forceinline__ double wg_max(double x, double *red)
{
	for (int off = 32; off > 0; off >>= 1) x = fmax(x, __shfl_xor(x, off));
	if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
	__syncthreads();
	const double r = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
	__syncthreads();
	return r;
}

// (v, i) beats (w, j): an index beats none; a larger value (a smaller one when MIN); equal
// values go to the smaller index.  A total order, so every combination order gives the same
// winner.
template <bool MIN>
__device__ __forceinline__ bool beats(double v, int i, double w, int j)
{
	if (i < 0) return false;
	if (j < 0) return true;
	if (v != w) return MIN ? (v < w) : (v > w);
	return i < j;
}

template <bool MIN>
__device__ __forceinline__ void wg_argext(double &v, int &idx, double *rv, int *ri)
{
	for (int off = 32; off > 0; off >>= 1) {
		const double ov = __shfl_xor(v, off);
		const int oi = __shfl_xor(idx, off);
		if (beats<MIN>(ov, oi, v, idx)) {
			v = ov;
			idx = oi;
		}
	}
	if ((threadIdx.x & 63) == 0) {
		rv[threadIdx.x >> 6] = v;
		ri[threadIdx.x >> 6] = idx;
	}
	__syncthreads();
	v = rv[0];
	idx = ri[0];
	for (int w = 1; w < NW; w++)
		if (beats<MIN>(rv[w], ri[w], v, idx)) {
			v = rv[w];
			idx = ri[w];
		}
	__syncthreads();
}

// largest value, ties to the smaller index; (0, -1) when nothing positive was seen.  Each thread
// scanned its indices in increasing order keeping only strictly larger values, so the result
// is the first index of the maximum, as a serial scan with ">" would find.
__device__ __forceinline__ void wg_argmax(double &v, int &idx, double *rv, int *ri)
{
	wg_argext<false>(v, idx, rv, ri);
}

// unused row with the smallest representation (first index among equals), if at most thr
// (aca_unrepresented, hmatrix.c:233-248)
__device__ int unrepresented(const int *rused, const double *urep, int m, double thr, double *rv, int *ri)
{
	double best = 0;
	int nxt = -1;
	for (int i = threadIdx.x; i < m; i += AT) {
		if (rused[i]) continue;
		if ((nxt < 0) || (urep[i] < best)) {
			best = urep[i];
			nxt = i;
		}
	}
	wg_argext<true>(best, nxt, rv, ri);
	return ((nxt >= 0) && (best <= thr)) ? nxt : -1;
}

// a - u v  (d_sub(a, d_mul(u, v)))
__device__ __forceinline__ void sub_uv(double &ar, double &ai, double2 u, double2 v)
{
	const double pr = (u.x * v.x) - (u.y * v.y);
	const double pi = (u.x * v.y) + (u.y * v.x);
	ar = ar - pr;
	ai = ai - pi;
}

// Column check where the reference's ACA stops (not part of the reference): the unused column
// least represented in the approximation (smallest max_l |v_l[j]|, first index among equals) is
// marked and its residual A(:, jc) - sum_l u_l v_l[jc] formed in the free factor slot U[k].  A
// block whose rows fall into groups that are exact multiples of each other can stop with a
// residual row of zero while the pivots never reached the columns that tell the groups apart;
// the reference then keeps an O(1) error.  Such a column's residual exceeds HM_COLCHECK times
// the tolerance (a converged block stays within a few tol ||A_k||_F, so it never fires there
// and the result is the reference's); the ACA then resumes at the unused row of largest
// residual in that column.  Returns that row, or -1 to stop.  k = 0 and k = lim are left to
// the reference's rule (stored dense).
template <int KM>
__device__ int column_check(const aca_args &a, const hfar &B, double2 *U, const double2 *V, int k, int lim,
                            const int *rused, const double *vrep, int *cused, double fro2, double *rv, int *ri)
{
	if ((k <= 0) || (k >= lim)) return -1;
	const int m = B.m, nn = B.nn;
	const int jc = unrepresented(cused, vrep, nn, INFINITY, rv, ri);
	if (jc < 0) return -1;
	if (threadIdx.x == 0) cused[jc] = 1;
	double r2 = 0, rmax = 0;
	int imax = -1;
	for (int i = threadIdx.x; i < m; i += AT) {
		const cplx e = hm_pos<KM>(a.E, a.perm, B.r0 + i, B.c0 + jc);
		double ar = e.r, ai = e.i;
		for (int l = 0; l < k; l++) sub_uv(ar, ai, U[(long)l * m + i], V[(long)l * nn + jc]);
		U[(long)k * m + i] = make_double2(ar, ai);
		const double t = sqrt((ar * ar) + (ai * ai));
		r2 += t * t;
		if (!rused[i] && (t > rmax)) {
			rmax = t;
			imax = i;
		}
	}
	r2 = wg_sum(r2, rv);
	wg_argmax(rmax, imax, rv, ri);
	return ((imax >= 0) && (sqrt(r2) > HM_COLCHECK * a.tol * sqrt(fro2))) ? imax : -1;
}

template <int KM>
__global__ __launch_bounds__(AT) void k_aca(aca_args a)
{
	__shared__ double rv[AT];
	__shared__ int ri[AT];
	__shared__ double2 cdu[HM_MAXIT], cdv[HM_MAXIT];
	__shared__ double sh_cross;
	const int b = a.order[blockIdx.x], tid = threadIdx.x;
	const int wave = tid >> 6, lane = tid & 63;
	const hfar B = a.far[b];
	const int r0 = B.r0, m = B.m, c0 = B.c0, nn = B.nn;
	const int kmax = (m < nn) ? m : nn;
	const int lim = (kmax < HM_MAXIT) ? kmax : HM_MAXIT;
	double2 *U = a.scr + a.soff[b];         // [lim][m]
	double2 *V = U + (long)m * lim;         // [lim][nn]
	double *urep = a.urep + a.moff[b], *vrep = urep + m;
	int *rused = a.rused + a.moff[b], *cused = rused + m;
	for (int i = tid; i < m + nn; i += AT) {
		urep[i] = 0;
		rused[i] = 0;
	}
	__syncthreads();
	double fro2 = 0, urmax = 0;
	int k = 0, irow = 0, probe = 0;
	while (k < lim) {
		// residual row  v = A(irow, :) - sum_l u_l[irow] v_l
		double amax = 0, rn2 = 0;
		int jmax = -1;
		for (int j = tid; j < nn; j += AT) {
			const cplx e = hm_pos<KM>(a.E, a.perm, r0 + irow, c0 + j);
			double ar = e.r, ai = e.i;
			for (int l = 0; l < k; l++) sub_uv(ar, ai, U[(long)l * m + irow], V[(long)l * nn + j]);
			V[(long)k * nn + j] = make_double2(ar, ai);
			const double t = sqrt((ar * ar) + (ai * ai));
			rn2 += t * t;
			if (t > amax) {
				amax = t;
				jmax = j;
			}
		}
		rn2 = wg_sum(rn2, rv);
		wg_argmax(amax, jmax, rv, ri);
		if (tid == 0) rused[irow] = 1;
		__syncthreads();
		if ((jmax < 0) || (amax <= 0) || (sqrt(rn2) <= a.tol * sqrt(fro2))) {
			int nxt = (++probe > HM_PROBE) ? -1 : unrepresented(rused, urep, m, a.tol * urmax, rv, ri);
			if (nxt < 0) {
				nxt = column_check<KM>(a, B, U, V, k, lim, rused, vrep, cused, fro2, rv, ri);
				if (nxt < 0) break;
				probe = 0;
			}
			irow = nxt;
			continue;
		}
		probe = 0;
		// normalise the row by its pivot (d_div)
		const double2 piv = V[(long)k * nn + jmax];
		__syncthreads();
		const double den = (piv.x * piv.x) + (piv.y * piv.y);
		for (int j = tid; j < nn; j += AT) {
			const double2 x = V[(long)k * nn + j];
			const double2 y = ((fabs(piv.x) <= 0) && (fabs(piv.y) <= 0))
			                      ? make_double2(0, 0)
			                      : make_double2(((x.x * piv.x) + (x.y * piv.y)) / den, ((x.y * piv.x) - (x.x * piv.y)) / den);
			V[(long)k * nn + j] = y;
			const double t = sqrt((y.x * y.x) + (y.y * y.y));
			if (t > vrep[j]) vrep[j] = t;
		}
		if (tid == 0) cused[jmax] = 1;
		__syncthreads();
		// residual column  u = A(:, jmax) - sum_l u_l v_l[jmax]
		for (int i = tid; i < m; i += AT) {
			const cplx e = hm_pos<KM>(a.E, a.perm, r0 + i, c0 + jmax);
			double ar = e.r, ai = e.i;
			for (int l = 0; l < k; l++) sub_uv(ar, ai, U[(long)l * m + i], V[(long)l * nn + jmax]);
			U[(long)k * m + i] = make_double2(ar, ai);
		}
		__syncthreads();
		// norms of u_k, v_k and the cross terms of ||A_k||_F^2
		double un = 0, vn = 0;
		for (int i = tid; i < m; i += AT) {
			const double2 x = U[(long)k * m + i];
			un += (x.x * x.x) + (x.y * x.y);
		}
		for (int j = tid; j < nn; j += AT) {
			const double2 x = V[(long)k * nn + j];
			vn += (x.x * x.x) + (x.y * x.y);
		}
		un = sqrt(wg_sum(un, rv));
		vn = sqrt(wg_sum(vn, rv));
		for (int l = wave; l < k; l += AT / 64) {
			double dur = 0, dui = 0, dvr = 0, dvi = 0;
			for (int i = lane; i < m; i += 64) {
				const double2 p = U[(long)l * m + i], q = U[(long)k * m + i];   // conj(p) q
				dur += (p.x * q.x) + (p.y * q.y);
				dui += (p.x * q.y) - (p.y * q.x);
			}
			for (int j = lane; j < nn; j += 64) {
				const double2 p = V[(long)l * nn + j], q = V[(long)k * nn + j];
				dvr += (p.x * q.x) + (p.y * q.y);
				dvi += (p.x * q.y) - (p.y * q.x);
			}
			for (int off = 32; off > 0; off >>= 1) {
				dur += __shfl_xor(dur, off);
				dui += __shfl_xor(dui, off);
				dvr += __shfl_xor(dvr, off);
				dvi += __shfl_xor(dvi, off);
			}
			if (lane == 0) {
				cdu[l] = make_double2(dur, dui);
				cdv[l] = make_double2(dvr, dvi);
			}
		}
		__syncthreads();
		if (tid == 0) {
			double cross = 0;
			for (int l = 0; l < k; l++) {
				const double2 p = cdu[l], q = cdv[l];
				const double pr = (p.x * q.x) - (p.y * q.y), pi = (p.x * q.y) + (p.y * q.x);
				cross += 2 * sqrt((pr * pr) + (pi * pi));
			}
			sh_cross = cross;
		}
		__syncthreads();
		fro2 += (un * un * vn * vn) + sh_cross;
		k++;
		// how strongly each row appears in the approximation (reducible blocks)
		double lmax = 0;
		for (int i = tid; i < m; i += AT) {
			const double2 x = U[(long)(k - 1) * m + i];
			const double t = sqrt((x.x * x.x) + (x.y * x.y));
			if (t > urep[i]) urep[i] = t;
			if (t > lmax) lmax = t;
		}
		lmax = wg_max(lmax, rv);
		if (lmax > urmax) urmax = lmax;
		// next pivot row: largest unused entry of u_k
		int inext = -1;
		if ((un * vn) > a.tol * sqrt((fro2 > 0) ? fro2 : 1)) {
			double umax = 0;
			for (int i = tid; i < m; i += AT) {
				if (rused[i]) continue;
				const double2 x = U[(long)(k - 1) * m + i];
				const double t = sqrt((x.x * x.x) + (x.y * x.y));
				if (t > umax) {
					umax = t;
					inext = i;
				}
			}
			wg_argmax(umax, inext, rv, ri);
		}
		if (inext < 0) {
			if (++probe <= HM_PROBE) inext = unrepresented(rused, urep, m, a.tol * urmax, rv, ri);
			if (inext < 0) {
				inext = column_check<KM>(a, B, U, V, k, lim, rused, vrep, cused, fro2, rv, ri);
				if (inext < 0) break;
				probe = 0;
			}
		}
		irow = inext;
	}
	if (tid == 0) a.rank[b] = ((k <= 0) || (k >= lim)) ? -1 : k;
}

// dense blocks (near, and far blocks ACA could not compress): D[i][j] = A(perm[r0+i], perm[c0+j])
template <int KM>
__global__ __launch_bounds__(AT) void k_hm_dense(hent E, const int *perm, const hdense *blk, double2 *D)
{
	const hdense B = blk[blockIdx.x];
	const long tot = (long)B.m * B.nn;
	for (long e = threadIdx.x; e < tot; e += AT) {
		const int i = (int)(e / B.nn), j = (int)(e - (long)i * B.nn);
		const cplx v = hm_pos<KM>(E, perm, B.r0 + i, B.c0 + j);
		D[B.off + e] = make_double2(v.r, v.i);
	}
}

// every entry of the matrix: F[i * n + j] = A(i, j); one thread per entry, workgroup q covers
// row q / nchunk, columns 256 (q mod nchunk) onwards
template <int KM>
__global__ __launch_bounds__(AT) void k_hm_full(hent E, int n, int nchunk, double2 *F)
{
	const long q = blockIdx.x;
	const int i = (int)(q / nchunk);
	const int j = (int)(q - (long)i * nchunk) * AT + threadIdx.x;
	if (j >= n) return;
	const cplx v = hm_entry<KM>(E, i, j);
	F[(long)i * n + j] = make_double2(v.r, v.i);
}

// compact the ACA factors of block b from the scratch into the U / V pools
__global__ void k_aca_compact(const hfar *far, const long *soff, const double2 *scr, double2 *Upool, double2 *Vpool)
{
	const hfar B = far[blockIdx.x];
	if (B.rank <= 0) return;
	const int kmax = (B.m < B.nn) ? B.m : B.nn;
	const int lim = (kmax < HM_MAXIT) ? kmax : HM_MAXIT;
	const double2 *U = scr + soff[blockIdx.x];
	const double2 *V = U + (long)B.m * lim;
	for (long e = threadIdx.x; e < (long)B.rank * B.m; e += blockDim.x) Upool[B.uo + e] = U[e];
	for (long e = threadIdx.x; e < (long)B.rank * B.nn; e += blockDim.x) Vpool[B.vo + e] = V[e];
}

// xp[j] = x[perm[j]]
__global__ void k_hm_gather(const int *perm, const double2 *x, double2 *xp, int n)
{
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e < n) xp[e] = x[perm[e]];
}

// stage 1: ft[fo_b + l] = v_l^T xp over the block's columns, one wave per (far block, rank)
// pair: item q = fo_b + l of block b = item_blk[q], lanes over columns, a fixed xor reduction
// (the pairs of one large high-rank block run in parallel instead of queueing in one workgroup)
__global__ __launch_bounds__(AT) void k_hm_vt_items(const hfar *far, const int *item_blk, long nitem, const double2 *Vpool,
                                                    const double2 *xp, double2 *ft)
{
	const long q = (long)blockIdx.x * (AT / 64) + (threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (q >= nitem) return;
	const hfar B = far[item_blk[q]];
	const int l = (int)(q - B.fo);
	const double2 *v = Vpool + B.vo + (long)l * B.nn;
	double sr = 0, si = 0;
	for (int j = lane; j < B.nn; j += 64) {
		const double2 a = v[j], x = xp[B.c0 + j];
		sr += (a.x * x.x) - (a.y * x.y);
		si += (a.x * x.y) + (a.y * x.x);
	}
	for (int off = 32; off > 0; off >>= 1) {
		sr += __shfl_xor(sr, off);
		si += __shfl_xor(si, off);
	}
	if (lane == 0) ft[q] = make_double2(sr, si);
}

// stage 2: the rows of each leaf row cluster (at most HM_LEAF = 32 rows) times the blocks
// listed for it (far blocks in registered order, then near blocks; NEAR_OFF: the near blocks
// other than the leaf's self-block).  HM_SPLIT workgroups share a leaf: workgroup s of leaf r
// takes blocks e = 4 s + wave, 4 s + wave + 4 HM_SPLIT, ...; in a wave lane & 31 is the row and
// lane >> 5 takes every other rank (U is stored rank by rank, so the 32 rows of one rank are one
// coalesced read) or every other column of a dense row.  The workgroup's 8 partial sums of a
// row are added in a fixed order into part[s][row]; k_hm_rowsum adds the HM_SPLIT of them in
// order, so the product is deterministic.
constexpr int HM_SPLIT = 8;

template <bool NEAR_OFF>
__global__ __launch_bounds__(AT) void k_hm_rows(hleafs L, const hfar *far, const hdense *near, const double2 *Upool,
                                                const double2 *Dpool, const double2 *ft, const double2 *xp, int n,
                                                double2 *part)
{
	__shared__ double2 lp[2 * (AT / 64)][32];
	const int r = blockIdx.x / HM_SPLIT, sp = blockIdx.x - r * HM_SPLIT;
	const int i0 = L.lo[r], i1 = L.hi[r];
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	const int row = lane & 31, sub = lane >> 5;
	const int f0 = L.rfoff[r], nf = NEAR_OFF ? 0 : (L.rfoff[r + 1] - f0);
	const int n0 = L.rnoff[r], nb = nf + (L.rnoff[r + 1] - n0);
	for (int c0 = i0; c0 < i1; c0 += 32) {
		const int i = c0 + row;
		const bool live = i < i1;
		double sr = 0, si = 0;
		for (int e = (AT / 64) * sp + wave; e < nb; e += (AT / 64) * HM_SPLIT) {
			int br0, bnn, bc0;
			long off;
			if (e < nf) {
				const int b = __builtin_amdgcn_readfirstlane(L.rfidx[f0 + e]);
				const hfar B = far[b];
				br0 = B.r0;
				bnn = B.nn;
				bc0 = B.c0;
				off = B.uo;
				if (B.rank > 0) {
					if (live) {
						const double2 *u = Upool + off + (i - br0);
						const double2 *t = ft + B.fo;
						const long m = B.m;
						for (int l = sub; l < B.rank; l += 2) {
							const double2 a = u[l * m], x = t[l];
							sr += (a.x * x.x) - (a.y * x.y);
							si += (a.x * x.y) + (a.y * x.x);
						}
					}
					continue;
				}
			}
			else {
				const int b = __builtin_amdgcn_readfirstlane(L.rnidx[n0 + (e - nf)]);
				const hdense B = near[b];
				if (NEAR_OFF && (B.r0 == B.c0) && (B.m == B.nn)) continue;
				br0 = B.r0;
				bnn = B.nn;
				bc0 = B.c0;
				off = B.off;
			}
			if (live) {
				const double2 *a = Dpool + off + (long)(i - br0) * bnn;
				const double2 *x = xp + bc0;
				for (int j = sub; j < bnn; j += 2) {
					const double2 v = a[j], w = x[j];
					sr += (v.x * w.x) - (v.y * w.y);
					si += (v.x * w.y) + (v.y * w.x);
				}
			}
		}
		lp[2 * wave + sub][row] = make_double2(sr, si);
		__syncthreads();
		if ((threadIdx.x < 32) && live) {
			double tr = 0, ti = 0;
#pragma unroll
			for (int q = 0; q < 2 * (AT / 64); q++) {
				tr += lp[q][threadIdx.x].x;
				ti += lp[q][threadIdx.x].y;
			}
			part[(long)sp * n + i] = make_double2(tr, ti);
		}
		__syncthreads();
	}
}

// y[perm[i]] = sum over s of part[s][i], s in increasing order
__global__ void k_hm_rowsum(const double2 *part, int n, const int *perm, double2 *y)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	double tr = 0, ti = 0;
#pragma unroll
	for (int sp = 0; sp < HM_SPLIT; sp++) {
		const double2 v = part[(long)sp * n + i];
		tr += v.x;
		ti += v.y;
	}
	y[perm[i]] = make_double2(tr, ti);
}

// sum of |entry|^2 of each near block (out[2 q] diagonal self-blocks, out[2 q + 1] the others)
__global__ __launch_bounds__(256) void k_hm_near_sq(const hdense *near, int nnear, const double2 *Dpool, double *out)
{
	__shared__ double red[256];
	const int q = blockIdx.x;
	const hdense B = near[q];
	const long cnt = (long)B.m * B.nn;
	double acc = 0;
	for (long e = threadIdx.x; e < cnt; e += 256) {
		const double2 a = Dpool[B.off + e];
		acc += (a.x * a.x) + (a.y * a.y);
	}
	red[threadIdx.x] = acc;
	__syncthreads();
	for (int w = 128; w > 0; w >>= 1) {
		if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		const bool diag = (B.r0 == B.c0) && (B.m == B.nn);
		out[2 * q] = diag ? red[0] : 0.0;
		out[2 * q + 1] = diag ? 0.0 : red[0];
	}
}

}  // namespace pe

// ------------------------------------------------------------------------------------------
// host side

namespace {

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

int km_class_hm(int km)
{
	if ((km & ~pe::KM_WIRE) == 0) return pe::KM_WIRE;
	if ((km & ~pe::KM_RIBBON) == 0) return pe::KM_RIBBON;
	if ((km & ~(pe::KM_RIBBON | pe::KM_BAR)) == 0) return pe::KM_RIBBON | pe::KM_BAR;
	if ((km & ~(pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR)) == 0) return pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR;
	return 15;
}

template <int KM>
void launch_aca(hipStream_t s, const pe::aca_args &a, int nfar)
{
	if (nfar > 0) hipLaunchKernelGGL(pe::k_aca<KM>, dim3(nfar), dim3(pe::AT), 0, s, a);
}

template <int KM>
void launch_dense(hipStream_t s, const pe::hent &E, const int *perm, const pe::hdense *blk, int nb, double2 *D)
{
	if (nb > 0) hipLaunchKernelGGL(pe::k_hm_dense<KM>, dim3(nb), dim3(pe::AT), 0, s, E, perm, blk, D);
}

void aca_dispatch(int km, hipStream_t s, const pe::aca_args &a, int nfar)
{
	switch (km_class_hm(km)) {
	case pe::KM_WIRE: launch_aca<pe::KM_WIRE>(s, a, nfar); break;
	case pe::KM_RIBBON: launch_aca<pe::KM_RIBBON>(s, a, nfar); break;
	case pe::KM_RIBBON | pe::KM_BAR: launch_aca<pe::KM_RIBBON | pe::KM_BAR>(s, a, nfar); break;
	case pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR: launch_aca<pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR>(s, a, nfar); break;
	default: launch_aca<15>(s, a, nfar); break;
	}
	HIPCHK(hipGetLastError());
}

template <int KM>
void launch_full(hipStream_t s, const pe::hent &E, int n, double2 *F)
{
	const int nchunk = (n + pe::AT - 1) / pe::AT;
	hipLaunchKernelGGL(pe::k_hm_full<KM>, dim3((unsigned)((long)n * nchunk)), dim3(pe::AT), 0, s, E, n, nchunk, F);
}

void full_dispatch(int km, hipStream_t s, const pe::hent &E, int n, double2 *F)
{
	switch (km_class_hm(km)) {
	case pe::KM_WIRE: launch_full<pe::KM_WIRE>(s, E, n, F); break;
	case pe::KM_RIBBON: launch_full<pe::KM_RIBBON>(s, E, n, F); break;
	case pe::KM_RIBBON | pe::KM_BAR: launch_full<pe::KM_RIBBON | pe::KM_BAR>(s, E, n, F); break;
	case pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR: launch_full<pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR>(s, E, n, F); break;
	default: launch_full<15>(s, E, n, F); break;
	}
	HIPCHK(hipGetLastError());
}

// bulk evaluation of all entries before the build (PEEC_HM_FULL=0 evaluates them inside ACA
// and the dense blocks instead): only when n^2 entries take at most a quarter of free memory
bool hm_full_wanted(int n)
{
	const char *e = getenv("PEEC_HM_FULL");
	if ((e != nullptr) && (atoi(e) == 0)) return false;
	size_t fr = 0, tot = 0;
	if (hipMemGetInfo(&fr, &tot) != hipSuccess) return false;
	return (double)n * n * sizeof(double2) <= 0.25 * (double)fr;
}

void dense_dispatch(int km, hipStream_t s, const pe::hent &E, const int *perm, const pe::hdense *blk, int nb, double2 *D)
{
	switch (km_class_hm(km)) {
	case pe::KM_WIRE: launch_dense<pe::KM_WIRE>(s, E, perm, blk, nb, D); break;
	case pe::KM_RIBBON: launch_dense<pe::KM_RIBBON>(s, E, perm, blk, nb, D); break;
	case pe::KM_RIBBON | pe::KM_BAR: launch_dense<pe::KM_RIBBON | pe::KM_BAR>(s, E, perm, blk, nb, D); break;
	case pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR: launch_dense<pe::KM_WIRE | pe::KM_RIBBON | pe::KM_BAR>(s, E, perm, blk, nb, D); break;
	default: launch_dense<15>(s, E, perm, blk, nb, D); break;
	}
	HIPCHK(hipGetLastError());
}

}  // namespace

// Build the compressed kernel matrix of a cell set (kind 0 Lp over the current cells, 1 P over
// the capacitive cells) at wavenumber kw.  pts: representative points in original numbering.
struct hmat_t *hm_build_dev(pgpu_t *g, int kind, int n, const double *pts, double kw, double tol, const char *label,
                            FILE *fp_log, const double2 *host_full)
{
	if (n <= 0) return nullptr;
	hmat_t *h = new hmat_t();
	h->n = n;
	h->kind = kind;
	h->kw = kw;
	h->tol = tol;
	h->g = g;
	if (g != nullptr) {
		h->s = pgpu_hip_stream(g);
	}
	else {
		HIPCHK(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
		h->own_stream = true;
	}
	hipStream_t s = h->s;
	if (hm_tree_build(&h->t, n, pts)) {
		if (fp_log) fprintf(fp_log, "*** compression : leaf clusters do not tile [0, %d)\n", n);
		if (h->own_stream) HIPCHK(hipStreamDestroy(h->s));
		delete h;
		return nullptr;
	}
	const hm_tree_t &t = h->t;
	int km = 15;
	pe::hent E;
	memset(&E, 0, sizeof(E));
	if (g != nullptr) {
		const int *csoff = nullptr, *csidx = nullptr;
		E.f = pgpu_set_args(g, kind, kw, &km, &csoff, &csidx);
		E.csoff = csoff;
		E.csidx = csidx;
		E.rec = pgpu_kf_records(g, kind, kw, &E.nslots, &E.T);
	}
	E.kind = kind;
	E.full = nullptr;
	E.nfull = n;
	h->km = km;
	h->d_perm = up(s, std::vector<int>(t.perm, t.perm + n));
	double2 *d_full = nullptr;
	if (host_full != nullptr) {
		// the caller's entries: ACA and the dense blocks read them instead of the element code
		d_full = dal<double2>((size_t)n * n);
		HIPCHK(hipMemcpyAsync(d_full, host_full, (size_t)n * n * sizeof(double2), hipMemcpyHostToDevice, s));
		E.full = d_full;
	}
	else if (hm_full_wanted(n) && (hipMalloc((void **)&d_full, (size_t)n * n * sizeof(double2)) == hipSuccess)) {
		if ((E.rec == nullptr) || (pgpu_kf_fill_kind(g, kind, kw, (d_complex_t *)d_full) != 0)) full_dispatch(km, s, E, n, d_full);
		E.full = d_full;
	}
	else {
		(void)hipGetLastError();
		d_full = nullptr;
	}
	// far blocks and their ACA scratch
	std::vector<pe::hfar> far(t.nfar);
	std::vector<long> soff(t.nfar + 1, 0), moff(t.nfar + 1, 0);
	for (int b = 0; b < t.nfar; b++) {
		pe::hfar &B = far[b];
		B.r0 = t.lo[t.fr[b]];
		B.m = t.hi[t.fr[b]] - B.r0;
		B.c0 = t.lo[t.fc[b]];
		B.nn = t.hi[t.fc[b]] - B.c0;
		B.rank = 0;
		B.uo = B.vo = B.fo = 0;
		const int kmax = std::min(B.m, B.nn), lim = std::min(kmax, HM_MAXIT);
		soff[b + 1] = soff[b] + (long)(B.m + B.nn) * lim;
		moff[b + 1] = moff[b] + B.m + B.nn;
	}
	std::vector<int> rank(t.nfar, 0);
	pe::hfar *d_far = up(s, far);
	if (t.nfar > 0) {
		// launch the blocks with the most work per rank step first, so the longest sequential
		// rank loops do not start last (scheduling only: each block's result is its own)
		std::vector<int> order(t.nfar);
		for (int b = 0; b < t.nfar; b++) order[b] = b;
		std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return soff[x + 1] - soff[x] > soff[y + 1] - soff[y]; });
		int *d_order = up(s, order);
		long *d_soff = up(s, soff), *d_moff = up(s, moff);
		double2 *scr = dal<double2>(soff[t.nfar]);
		double *urep = dal<double>(moff[t.nfar]);
		int *rused = dal<int>(moff[t.nfar]);
		int *d_rank = dal<int>(t.nfar);
		pe::aca_args a;
		a.E = E;
		a.perm = h->d_perm;
		a.far = d_far;
		a.order = d_order;
		a.soff = d_soff;
		a.moff = d_moff;
		a.scr = scr;
		a.urep = urep;
		a.rused = rused;
		a.rank = d_rank;
		a.tol = tol;
		aca_dispatch(km, s, a, t.nfar);
		HIPCHK(hipMemcpyAsync(rank.data(), d_rank, t.nfar * sizeof(int), hipMemcpyDeviceToHost, s));
		HIPCHK(hipStreamSynchronize(s));
		// compact pools, dense storage for blocks that did not compress
		long nu = 0, nv = 0, nd = 0, nf = 0;
		for (int b = 0; b < t.nfar; b++) {
			pe::hfar &B = far[b];
			B.rank = (rank[b] > 0) ? rank[b] : 0;
			if (B.rank > 0) {
				B.uo = nu;
				B.vo = nv;
				B.fo = nf;
				nu += (long)B.rank * B.m;
				nv += (long)B.rank * B.nn;
				nf += B.rank;
			}
		}
		h->nU = nu;
		h->nV = nv;
		h->ftlen = nf;
		h->d_U = dal<double2>(nu);
		h->d_V = dal<double2>(nv);
		HIPCHK(hipMemcpyAsync(d_far, far.data(), far.size() * sizeof(pe::hfar), hipMemcpyHostToDevice, s));
		hipLaunchKernelGGL(pe::k_aca_compact, dim3(t.nfar), dim3(256), 0, s, (const pe::hfar *)d_far, (const long *)d_soff,
		                   (const double2 *)scr, h->d_U, h->d_V);
		HIPCHK(hipStreamSynchronize(s));
		HIPCHK(hipFree(scr));
		HIPCHK(hipFree(urep));
		HIPCHK(hipFree(rused));
		HIPCHK(hipFree(d_rank));
		HIPCHK(hipFree(d_soff));
		HIPCHK(hipFree(d_moff));
		HIPCHK(hipFree(d_order));
		(void)nd;
	}
	// dense storage: far blocks that did not compress, then the near blocks
	std::vector<pe::hdense> dense;
	long nd = 0;
	for (int b = 0; b < t.nfar; b++) {
		if (far[b].rank > 0) continue;
		far[b].uo = nd;
		pe::hdense D{far[b].r0, far[b].m, far[b].c0, far[b].nn, nd};
		dense.push_back(D);
		nd += (long)far[b].m * far[b].nn;
	}
	std::vector<pe::hdense> near(t.nnear);
	for (int q = 0; q < t.nnear; q++) {
		pe::hdense &B = near[q];
		B.r0 = t.lo[t.nr[q]];
		B.m = t.hi[t.nr[q]] - B.r0;
		B.c0 = t.lo[t.nc2[q]];
		B.nn = t.hi[t.nc2[q]] - B.c0;
		B.off = nd;
		nd += (long)B.m * B.nn;
		dense.push_back(B);
	}
	h->nD = nd;
	h->d_D = dal<double2>(nd);
	pe::hdense *d_dense = up(s, dense);
	dense_dispatch(km, s, E, h->d_perm, d_dense, (int)dense.size(), h->d_D);
	HIPCHK(hipMemcpyAsync(d_far, far.data(), far.size() * sizeof(pe::hfar), hipMemcpyHostToDevice, s));
	h->d_far = d_far;
	h->d_near = up(s, near);
	HIPCHK(hipStreamSynchronize(s));
	HIPCHK(hipFree(d_dense));
	if (d_full != nullptr) HIPCHK(hipFree(d_full));
	h->rank = rank;
	h->far = far;
	h->near = near;
	// leaf-row lists
	std::vector<int> llo(t.nleaf), lhi(t.nleaf);
	for (int r = 0; r < t.nleaf; r++) {
		llo[r] = t.lo[t.leafid[r]];
		lhi[r] = t.hi[t.leafid[r]];
	}
	h->L.lo = up(s, llo);
	h->L.hi = up(s, lhi);
	h->L.rfoff = up(s, std::vector<int>(t.rfoff, t.rfoff + t.nleaf + 1));
	h->L.rfidx = up(s, std::vector<int>(t.rfidx, t.rfidx + std::max(1, t.rfoff[t.nleaf])));
	h->L.rnoff = up(s, std::vector<int>(t.rnoff, t.rnoff + t.nleaf + 1));
	h->L.rnidx = up(s, std::vector<int>(t.rnidx, t.rnidx + std::max(1, t.rnoff[t.nleaf])));
	h->d_ft = dal<double2>(std::max(1L, h->ftlen));
	// far block of each V^T x item (kept alive until the stream is synchronised below)
	std::vector<int> item((size_t)std::max(1L, h->ftlen), 0);
	for (int b = 0; b < t.nfar; b++)
		for (int l = 0; l < far[b].rank; l++) item[(size_t)far[b].fo + l] = b;
	h->d_item = up(s, item);
	h->d_xp = dal<double2>(n);
	h->d_part = dal<double2>((size_t)pe::HM_SPLIT * n);
	h->d_x = dal<double2>(n);
	h->d_y = dal<double2>(n);
	HIPCHK(hipStreamSynchronize(s));
	if (fp_log) {
		const double mb = hmat_memory_mb(h), dm = hmat_dense_mb(n);
		fprintf(fp_log, "compression %s : %d cells, %d far blocks + %d near blocks, tol = %.1e, %.2f MB (dense %.2f MB, %.1fx)\n",
		        label, n, t.nfar, t.nnear, tol, mb, dm, (mb > 0) ? (dm / mb) : 0);
		fflush(fp_log);
	}
	return h;
}

// y = H x with x, y device vectors in original numbering, on the H-matrix's stream
void hm_matvec_dev(const hmat_t *h, const double2 *x, double2 *y)
{
	hipStream_t s = h->s;
	const int n = h->n;
	hipLaunchKernelGGL(pe::k_hm_gather, dim3((n + 255) / 256), dim3(256), 0, s, (const int *)h->d_perm, x, h->d_xp, n);
	if (h->ftlen > 0)
		hipLaunchKernelGGL(pe::k_hm_vt_items, dim3((unsigned)((h->ftlen + pe::AT / 64 - 1) / (pe::AT / 64))), dim3(pe::AT), 0, s,
		                   (const pe::hfar *)h->d_far, (const int *)h->d_item, h->ftlen, (const double2 *)h->d_V,
		                   (const double2 *)h->d_xp, h->d_ft);
	hipLaunchKernelGGL(pe::k_hm_rows<false>, dim3(h->t.nleaf * pe::HM_SPLIT), dim3(pe::AT), 0, s, h->L, (const pe::hfar *)h->d_far,
	                   (const pe::hdense *)h->d_near, (const double2 *)h->d_U, (const double2 *)h->d_D, (const double2 *)h->d_ft,
	                   (const double2 *)h->d_xp, n, h->d_part);
	hipLaunchKernelGGL(pe::k_hm_rowsum, dim3((n + 255) / 256), dim3(256), 0, s, (const double2 *)h->d_part, n,
	                   (const int *)h->d_perm, y);
}

// y = N x: near-field blocks off the block diagonal (x, y in cell order, device)
void hm_near_off_dev(const hmat_t *h, const double2 *x, double2 *y)
{
	hipStream_t s = h->s;
	const int n = h->n;
	hipLaunchKernelGGL(pe::k_hm_gather, dim3((n + 255) / 256), dim3(256), 0, s, (const int *)h->d_perm, x, h->d_xp, n);
	hipLaunchKernelGGL(pe::k_hm_rows<true>, dim3(h->t.nleaf * pe::HM_SPLIT), dim3(pe::AT), 0, s, h->L, (const pe::hfar *)h->d_far,
	                   (const pe::hdense *)h->d_near, (const double2 *)h->d_U, (const double2 *)h->d_D, (const double2 *)h->d_ft,
	                   (const double2 *)h->d_xp, n, h->d_part);
	hipLaunchKernelGGL(pe::k_hm_rowsum, dim3((n + 255) / 256), dim3(256), 0, s, (const double2 *)h->d_part, n,
	                   (const int *)h->d_perm, y);
}

// Frobenius norms of the near field: off the block diagonal and of the leaf self-blocks
void hm_near_norms(const hmat_t *h, double *off, double *diag)
{
	const int nn = (int)h->near.size();
	*off = *diag = 0;
	if (nn == 0) return;
	double *d = nullptr;
	HIPCHK(hipMalloc((void **)&d, (size_t)2 * nn * sizeof(double)));
	hipLaunchKernelGGL(pe::k_hm_near_sq, dim3(nn), dim3(256), 0, h->s, (const pe::hdense *)h->d_near, nn,
	                   (const double2 *)h->d_D, d);
	std::vector<double> hsq((size_t)2 * nn);
	HIPCHK(hipMemcpyAsync(hsq.data(), d, hsq.size() * sizeof(double), hipMemcpyDeviceToHost, h->s));
	HIPCHK(hipStreamSynchronize(h->s));
	HIPCHK(hipFree(d));
	double so = 0, sd = 0;
	for (int q = 0; q < nn; q++) {
		sd += hsq[2 * q];
		so += hsq[2 * q + 1];
	}
	*off = std::sqrt(so);
	*diag = std::sqrt(sd);
}

// ------------------------------------------------------------------------------------------
// the reference's C interface (peec.h)

extern "C" {

// hmatrix.c:403-614 for any kernel: the cluster tree and partition of the n points, then ACA
// per far block and the dense near blocks on the device.  The caller's fn is a host function,
// so its n^2 entries are evaluated once on the host (in parallel: the reference requires fn to
// be re-entrant, hmatrix.c:427-454), uploaded, and read by the device ACA exactly where the
// library's own kernels read their bulk entries.  The library's own lp_entry / pot_entry are
// recognised and evaluated on the device instead.  NULL when n^2 entries exceed
// PEEC_HM_GEN_GB (default 16) GiB of host memory.
struct hmat_t *hmat_build_gen(int n, const double *pts, hmat_entry_fn fn, void *ctx, double kw, double tol,
                              const char *label, FILE *fp_log)
{
	if ((n <= 0) || (fn == nullptr)) return nullptr;
	const char *lab = label ? label : "";
	const peec_t *p = (const peec_t *)ctx;
	if ((fn == (hmat_entry_fn)lp_entry) && (n == p->nseg)) return hm_build_dev(pgpu_get(p), 0, n, pts, kw, tol, lab, fp_log);
	if ((fn == (hmat_entry_fn)pot_entry) && (n == p->ncell)) return hm_build_dev(pgpu_get(p), 1, n, pts, kw, tol, lab, fp_log);
	const char *env = getenv("PEEC_HM_GEN_GB");
	const double cap_gb = env ? atof(env) : 16.0;
	if ((double)n * n * sizeof(double2) > cap_gb * 1073741824.0) {
		if (fp_log) fprintf(fp_log, "*** compression %s : %d x %d caller entries exceed %.0f GiB of host memory\n", lab, n, n, cap_gb);
		return nullptr;
	}
	if (!pgpu_device_ok()) pgpu_fatal("no HIP device");
	std::vector<double2> full((size_t)n * n);
	const int nth = std::max(1, std::min((int)std::thread::hardware_concurrency(), 64));
	std::vector<std::thread> pool;
	for (int w = 0; w < nth; w++)
		pool.emplace_back([&, w]() {
			for (int i = w; i < n; i += nth)
				for (int j = 0; j < n; j++) {
					const d_complex_t v = fn(ctx, i, j, kw);
					full[(size_t)i * n + j] = make_double2(v.r, v.i);
				}
		});
	for (std::thread &t : pool) t.join();
	return hm_build_dev(nullptr, 0, n, pts, kw, tol, lab, fp_log, full.data());
}

struct hmat_t *hmat_build(peec_t *p, double f, FILE *fp_log)
{
	const int n = p->nseg;
	if (n <= 0) return nullptr;
	std::vector<double> pts((size_t)3 * n);
	for (int i = 0; i < n; i++)
		for (int c = 0; c < 3; c++) pts[3 * i + c] = 0.5 * (p->seg[i].x1[c] + p->seg[i].x2[c]);
	const double kw = p->retardation ? (2 * PI * f / C0) : 0;
	return hm_build_dev(pgpu_get(p), 0, n, pts.data(), kw, p->ctol, "Lp", fp_log);
}

struct hmat_t *hmat_build_pot(peec_t *p, double f, FILE *fp_log)
{
	if (pot_cells(p)) return nullptr;
	if (p->ncell <= 0) return nullptr;
	const double kw = p->retardation ? (2 * PI * f / C0) : 0;
	return hm_build_dev(pgpu_get(p), 1, p->ncell, p->cpt, kw, p->ctol, "P", fp_log);
}

void hmat_free(struct hmat_t *h)
{
	if (h == nullptr) return;
	HIPCHK(hipStreamSynchronize(h->s));
	const void *bufs[] = {h->d_perm, h->d_far, h->d_near, h->d_U, h->d_V, h->d_D, h->L.lo, h->L.hi, h->L.rfoff, h->L.rfidx,
	                      h->L.rnoff, h->L.rnidx, h->d_ft, h->d_xp, h->d_x, h->d_y, h->d_part, h->d_item};
	for (const void *b : bufs)
		if (b) HIPCHK(hipFree(const_cast<void *>(b)));
	if (h->own_stream) HIPCHK(hipStreamDestroy(h->s));
	hm_tree_free(&h->t);
	delete h;
}

// y = H x for host vectors (one copy each way)
void hmat_matvec(const struct hmat_t *h, const d_complex_t *x, d_complex_t *y)
{
	const int n = h->n;
	HIPCHK(hipMemcpyAsync(h->d_x, x, (size_t)n * sizeof(double2), hipMemcpyHostToDevice, h->s));
	hm_matvec_dev(h, h->d_x, h->d_y);
	HIPCHK(hipMemcpyAsync(y, h->d_y, (size_t)n * sizeof(double2), hipMemcpyDeviceToHost, h->s));
	HIPCHK(hipStreamSynchronize(h->s));
}

int hmat_nleaf(const struct hmat_t *h)
{
	return (h != nullptr) ? h->t.nleaf : 0;
}

// leaf r: its cells (original numbers) and its diagonal block (host copy, m x m row-major)
const d_complex_t *hmat_leaf_block(const struct hmat_t *h, int r, const int **idx, int *m)
{
	hmat_t *hh = const_cast<hmat_t *>(h);
	const hm_tree_t &t = h->t;
	const int a = t.leafid[r];
	*idx = &t.perm[t.lo[a]];
	*m = t.hi[a] - t.lo[a];
	if (hh->leafblk.empty()) {
		hh->leafoff.assign(t.nleaf + 1, -1);
		std::vector<int> which(t.nleaf, -1);
		long tot = 0;
		for (int rr = 0; rr < t.nleaf; rr++) {
			const int aa = t.leafid[rr];
			for (int e = t.rnoff[rr]; e < t.rnoff[rr + 1]; e++) {
				const int q = t.rnidx[e];
				if ((t.nr[q] == aa) && (t.nc2[q] == aa)) which[rr] = q;
			}
			hh->leafoff[rr] = tot;
			const int mm = t.hi[aa] - t.lo[aa];
			tot += (long)mm * mm;
		}
		hh->leafblk.resize(std::max(1L, tot));
		for (int rr = 0; rr < t.nleaf; rr++) {
			if (which[rr] < 0) continue;
			const pe::hdense &B = h->near[which[rr]];
			HIPCHK(hipMemcpyAsync(&hh->leafblk[hh->leafoff[rr]], h->d_D + B.off, (size_t)B.m * B.nn * sizeof(double2),
			                      hipMemcpyDeviceToHost, h->s));
		}
		HIPCHK(hipStreamSynchronize(h->s));
		hh->leafhas = which;
	}
	if (hh->leafhas[r] < 0) return nullptr;
	return (const d_complex_t *)&hh->leafblk[hh->leafoff[r]];
}

double hmat_memory_mb(const struct hmat_t *h)
{
	if (h == nullptr) return 0;
	return (double)(h->nU + h->nV + h->nD) * sizeof(d_complex_t) / (1024.0 * 1024.0);
}

double hmat_dense_mb(int n)
{
	return (double)n * n * sizeof(d_complex_t) / (1024.0 * 1024.0);
}

// block structure and ranks for tests: counts[0..3] = nfar, nnear, nleaf, n; perm, ranks
int hmat_info(const struct hmat_t *h, int *counts, int *perm, int *ranks)
{
	if (h == nullptr) return 1;
	counts[0] = h->t.nfar;
	counts[1] = h->t.nnear;
	counts[2] = h->t.nleaf;
	counts[3] = h->n;
	if (perm) memcpy(perm, h->t.perm, (size_t)h->n * sizeof(int));
	if (ranks)
		for (int b = 0; b < h->t.nfar; b++) ranks[b] = h->rank[b];
	return 0;
}

}  // extern "C"
