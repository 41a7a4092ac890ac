// lu.hip — dense complex FP64 LU with partial pivoting on gfx950, and the triangular solves.
//
// Replaces lu_decomp / lu_solve (OpenPEEC src/lu.c:16-96).  Same contract: row-major matrix,
// in-place Doolittle factors (unit L below the diagonal, U on and above), piv[k] = row swapped
// with row k at step k, pivot = first row of maximal |re| + |im| in the column (ties to the
// lowest index, lu.c:19-28), failure when that maximum is below 1e-30 (lu.c:30).
//
// Structure (right-looking, blocked by NB columns):
//   k_panel     one cooperative launch per panel: the panel rows are split over P workgroups
//               held in LDS; per column every workgroup publishes its local pivot candidate
//               (value, row index, row contents) and the current diagonal row, reads every
//               workgroup's candidate value from a ring of slots (the wait and the read in one
//               poll), then every workgroup derives the same winner, swaps and eliminates.
//               That exchange per column is the whole latency cost of partial pivoting.
//   k_laswp     applies the panel's swaps to the columns left and right of the panel and to
//               the running row permutation (used by the solves as one gather).
//   k_trsm_lu   U12 = L11^-1 A12 (unit lower, LDS-staged).
//   GEMM        A22 -= L21 U12: rocBLAS zgemm / cgemm by default; PEEC_GEMM=hip selects the
//               hand-written k_zgemm (FP64 MFMA, v_mfma_f64_16x16x4) / k_cgemm32 (FP32 MFMA),
//               complex as 4 real products.
// The trailing update's summation order differs from the reference's rank-1 sequence, so the
// factors agree with the reference to rounding, not bitwise; pivots agree except on near-ties.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>
#include "../peec_internal.h"
#include "hip_check.hpp"

namespace {

constexpr int NB = 64;          // panel width (GEMM K)
constexpr int PANEL_T = 256;    // threads per panel workgroup
constexpr int PANEL_RMAX = 144; // panel rows per workgroup held in LDS (144 * 65 * 16 B = 146 KiB)
constexpr int PANEL_PMAX = 256; // workgroups in a panel launch (at most one per CU)
constexpr int PANEL_RING = 4;   // columns of candidate slots in flight (see k_panel)
constexpr unsigned PANEL_SPIN_LIMIT = 1u << 22;   // polls before the exchange gives up (seconds)

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
	return make_double2((a.x * b.x) - (a.y * b.y), (a.x * b.y) + (a.y * b.x));
}

__device__ __forceinline__ double2 cinv(double2 z)
{
	const double n2 = (z.x * z.x) + (z.y * z.y);
	if (n2 <= 0) return make_double2(0, 0);
	return make_double2(z.x / n2, -z.y / n2);
}

// ------------------------------------------------------------------------------------------
// Pivot exchange of the cooperative panel: per column every workgroup publishes its candidate
// (row index, row contents, and the value max |re| + |im|) and reads all candidates' values.
// The values go to a ring of PANEL_RING columns of slots, one per workgroup, NaN meaning "not
// yet written" (the ring is filled with 0xFF bytes, a NaN, before each launch): the reader polls
// the slots with coherent loads until none is NaN, so the wait and the read of the values are
// one round trip, where a counter barrier needed an atomic, a poll and then the read.  The
// writer publishes the row with plain stores and a release fence before its slot; the reader
// fences (acquire) after its poll.  A workgroup clears its own slot two columns ahead once it
// has seen every slot of the current column: all workgroups have then left the column two
// back.  A two-level counter barrier measured 5.2 us per column at 117 workgroups; the first
// flat counter 4.7-6.5 us (same-address atomics serialise at the memory side).
constexpr int SLOT_BYTES = PANEL_RING * PANEL_PMAX * (int)sizeof(double);

// (value, index) arg-max with ties to the smaller index; value < 0 marks "no candidate"
__device__ __forceinline__ void amax_merge(double &v, int &i, double ov, int oi)
{
	if ((ov > v) || ((ov == v) && (oi < i))) {
		v = ov;
		i = oi;
	}
}

// storage <-> arithmetic: factors may be stored in FP64 (double2) or FP32 (float2); every
// operation below runs in FP64
__device__ __forceinline__ double2 ld2(const double2 &v) { return v; }
__device__ __forceinline__ double2 ld2(const float2 &v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ void st2(double2 &d, double2 v) { d = v; }
__device__ __forceinline__ void st2(float2 &d, double2 v) { d = make_float2((float)v.x, (float)v.y); }

struct panel_ws {
	double  *slot;      // [PANEL_RING][PANEL_PMAX] local max |re|+|im| per column, NaN until written
	int     *pubi;      // [2][P]     its global row
	double2 *pubrow;    // [2][P][NB] that row's panel entries
	double2 *diagrow;   // [2][NB]    current row kk
	int     *info;      // [0] singular row or -1; [1] column whose pivot exchange timed out, or -1
};

constexpr int SROW = NB + 1;    // padded LDS row stride (complex): spreads row-per-thread access over banks

// Panel factorisation.  Thread t of workgroup w owns panel row r0 + t (rows held in LDS with a
// padded stride); per column: wave 0 finds and publishes the local candidate, one grid barrier,
// wave 0 merges the P candidates and fetches the pivot row and the current diagonal row, then
// every thread swaps/eliminates its own row.  Four workgroup barriers and two global round
// trips per column.
template <typename TS>
__global__ __launch_bounds__(PANEL_T) void k_panel(TS *A, long lda, int n, int k0, int nb, int R, int *piv,
                                                    panel_ws ws)
{
	extern __shared__ double2 S[];   // [R][SROW]
	__shared__ double2 U[NB], D[NB];
	__shared__ int s_gi, s_stop;
	const int P = gridDim.x, w = blockIdx.x, tid = threadIdx.x;
	const int wave = tid >> 6, lane = tid & 63;
	// a singular column in an earlier panel ends the factorisation: every workgroup reads the
	// same flag (written before this launch) and leaves before the first barrier
	if (*ws.info >= 0) return;
	const int r0 = k0 + w * R;
	const int rc = max(0, min(R, n - r0));

	for (int e = tid; e < rc * nb; e += PANEL_T) {
		const int r = e / nb, c = e - r * nb;
		S[r * SROW + c] = ld2(A[(long)(r0 + r) * lda + k0 + c]);
	}
	__syncthreads();

	for (int jj = 0; jj < nb; jj++) {
		const int kk = k0 + jj, par = jj & 1, ring = jj & (PANEL_RING - 1);
		if (wave == 0) {
			// local candidate: the first row of largest |re| + |im| among this workgroup's rows
			double bv = -1;
			int bi = INT_MAX;
			for (int r = lane; r < rc; r += 64) {
				const int g = r0 + r;
				if (g >= kk) {
					const double2 a = S[r * SROW + jj];
					amax_merge(bv, bi, fabs(a.x) + fabs(a.y), g);
				}
			}
			for (int off = 32; off > 0; off >>= 1) {
				const double ov = __shfl_xor(bv, off);
				const int oi = __shfl_xor(bi, off);
				amax_merge(bv, bi, ov, oi);
			}
			// publish: row index and row contents (and the current diagonal row), then, after a
			// release fence, the candidate's value in this workgroup's slot of the ring
			if (lane == 0) ws.pubi[par * P + w] = bi;
			if ((bv >= 0) && (lane < nb)) ws.pubrow[((long)par * P + w) * NB + lane] = S[(bi - r0) * SROW + lane];
			if ((kk >= r0) && (kk < r0 + rc) && (lane < nb)) ws.diagrow[par * NB + lane] = S[(kk - r0) * SROW + lane];
			// release only (L2 write-back): this side reads nothing the other workgroups wrote
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
			if (lane == 0) __hip_atomic_store(&ws.slot[ring * PANEL_PMAX + w], bv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
			// wait until every workgroup's slot of this column holds a value (NaN: not yet),
			// reading the values themselves; bounded, so a fault in the exchange ends the
			// factorisation (reported as a zero pivot) instead of hanging
			double qv[PANEL_PMAX / 64];
			unsigned spins = 0;
			bool timeout = false;
			for (;;) {
				// the lane's slots read together (one round trip), then tested
#pragma unroll
				for (int i = 0; i < PANEL_PMAX / 64; i++) {
					const int q = lane + 64 * i;
					qv[i] = (q < P) ? __hip_atomic_load(&ws.slot[ring * PANEL_PMAX + min(q, P - 1)], __ATOMIC_RELAXED,
					                                    __HIP_MEMORY_SCOPE_AGENT)
					                : -1.0;
				}
				bool ok = true;
#pragma unroll
				for (int i = 0; i < PANEL_PMAX / 64; i++)
					if (qv[i] != qv[i]) ok = false;
				if (__all(ok)) break;
				if (++spins > PANEL_SPIN_LIMIT) {
					timeout = true;
					break;
				}
				__builtin_amdgcn_s_sleep(1);
			}
			// acquire only (L2 invalidate): the candidates' rows are read after their slots
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
			// the winner: largest value, ties to the smaller workgroup (whose rows come first)
			double gv = -1;
			int gw = INT_MAX;
#pragma unroll
			for (int i = 0; i < PANEL_PMAX / 64; i++) {
				const int q = lane + 64 * i;
				if ((q < P) && ((qv[i] > gv) || ((qv[i] == gv) && (q < gw)))) {
					gv = qv[i];
					gw = q;
				}
			}
			for (int off = 32; off > 0; off >>= 1) {
				const double ov = __shfl_xor(gv, off);
				const int ow = __shfl_xor(gw, off);
				if ((ov > gv) || ((ov == gv) && (ow < gw))) {
					gv = ov;
					gw = ow;
				}
			}
			const int stop = timeout || !(gv >= 1e-30);
			const int gi = (gw < P) ? ws.pubi[par * P + gw] : INT_MAX;
			if (!stop && (lane < nb)) {
				U[lane] = ws.pubrow[((long)par * P + gw) * NB + lane];
				D[lane] = ws.diagrow[par * NB + lane];
			}
			// this workgroup's slot for column jj + 2 back to "not yet": every workgroup has
			// published column jj, so none still reads column jj - 2's values
			if (lane == 0) {
				__hip_atomic_store(&ws.slot[((jj + 2) & (PANEL_RING - 1)) * PANEL_PMAX + w], __builtin_nan(""),
				                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
				s_gi = gi;
				s_stop = stop;
				if (timeout) printf("*** LU panel: pivot exchange timed out (column %d)\n", kk);
				if (w == 0) {
					piv[kk] = stop ? ((gi == INT_MAX) ? kk : gi) : gi;
					if (stop) *ws.info = kk;
					if (timeout) ws.info[1] = kk;   // a failed exchange, not a zero pivot
				}
			}
		}
		__syncthreads();
		if (s_stop) return;
		const int gi = s_gi;
		// the interchange, one lane per column: row kk takes the pivot row, the pivot row's
		// place takes row kk (a serial copy by the two owning threads was the column's longest path)
		if (tid < nb) {
			if ((kk >= r0) && (kk < r0 + rc)) S[(kk - r0) * SROW + tid] = U[tid];
			if ((gi != kk) && (gi >= r0) && (gi < r0 + rc)) S[(gi - r0) * SROW + tid] = D[tid];
		}
		__syncthreads();
		if (tid < rc) {
			double2 *row = &S[tid * SROW];
			const int g = r0 + tid;
			if (g > kk) {
				const double2 l = cmul(row[jj], cinv(U[jj]));
				row[jj] = l;
#pragma unroll 4
				for (int c = jj + 1; c < nb; c++) {
					const double2 p = cmul(l, U[c]);
					double2 a = row[c];
					a.x = a.x - p.x;
					a.y = a.y - p.y;
					row[c] = a;
				}
			}
		}
		__syncthreads();
	}
	for (int e = tid; e < rc * nb; e += PANEL_T) {
		const int r = e / nb, c = e - r * nb;
		st2(A[(long)(r0 + r) * lda + k0 + c], S[r * SROW + c]);
	}
}

// apply piv[k0 .. k0+nb) to every column outside the panel, and to the running permutation
template <typename TS>
__global__ void k_laswp(TS *A, long lda, int n, int k0, int nb, const int *piv, int *perm, const int *info)
{
	if (*info >= 0) return;
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	const int nout = n - nb;
	if (t < nout) {
		const int j = (t < k0) ? t : (t + nb);
		for (int kk = k0; kk < k0 + nb; kk++) {
			const int p = piv[kk];
			if (p != kk) {
				const TS x = A[(long)kk * lda + j];
				A[(long)kk * lda + j] = A[(long)p * lda + j];
				A[(long)p * lda + j] = x;
			}
		}
	}
	else if ((t == nout) && (perm != nullptr)) {
		for (int kk = k0; kk < k0 + nb; kk++) {
			const int p = piv[kk];
			if (p != kk) {
				const int x = perm[kk];
				perm[kk] = perm[p];
				perm[p] = x;
			}
		}
	}
}

// X <- T^-1 X for an nb x nb triangle T (unit lower if LOWER, else non-unit upper) at T[0][0]
// and the nb x ncol block X; one thread per column, T and X staged in LDS.
template <bool LOWER, typename TS>
__global__ __launch_bounds__(64) void k_trsm(const TS *T, long ldt, TS *X, long ldx, int nb, int ncol)
{
	__shared__ double2 St[NB * NB];
	__shared__ double2 Sx[NB * 64];
	const int tid = threadIdx.x;
	const int j = blockIdx.x * 64 + tid;
	for (int e = tid; e < nb * nb; e += 64) {
		const int r = e / nb, c = e - r * nb;
		St[r * NB + c] = ld2(T[(long)r * ldt + c]);
	}
	if (j < ncol)
		for (int r = 0; r < nb; r++) Sx[r * 64 + tid] = ld2(X[(long)r * ldx + j]);
	__syncthreads();
	if (j >= ncol) return;
	if (LOWER) {
		for (int r = 1; r < nb; r++) {
			double2 sacc = Sx[r * 64 + tid];
			for (int q = 0; q < r; q++) {
				const double2 p = cmul(St[r * NB + q], Sx[q * 64 + tid]);
				sacc.x -= p.x;
				sacc.y -= p.y;
			}
			Sx[r * 64 + tid] = sacc;
		}
	}
	else {
		for (int r = nb - 1; r >= 0; r--) {
			double2 sacc = Sx[r * 64 + tid];
			for (int q = r + 1; q < nb; q++) {
				const double2 p = cmul(St[r * NB + q], Sx[q * 64 + tid]);
				sacc.x -= p.x;
				sacc.y -= p.y;
			}
			Sx[r * 64 + tid] = cmul(sacc, cinv(St[r * NB + r]));
		}
	}
	for (int r = 0; r < nb; r++) st2(X[(long)r * ldx + j], Sx[r * 64 + tid]);
}

// C[M x N] -= A[M x K] B[K x N] on FP64 MFMA.  Wave tile 32 x 64 (2 x 4 blocks of 16 x 16),
// workgroup 2 x 2 waves = 64 x 128.  v_mfma_f64_16x16x4 lane layout (checked on gfx950 with an
// in-bounds probe): A row = lane%16, k = lane/16; B col = lane%16, k = lane/16; D register v of
// lane l holds row lane/16 + 4 v, col lane%16.  Complex = 4 real MFMAs into (re, im)
// accumulators.  The fragments of step k+4 are loaded before the MFMAs of step k, so each
// wave keeps its loads in flight behind 64 MFMAs.
// Tile of this workgroup, XCD-aware.  Workgroups are dispatched round-robin over the 8 XCDs
// (linear id w runs on XCD w mod 8), each with its own L2.  The linear id is remapped so that
// every XCD takes a contiguous run of tiles, walked in groups of GT tile rows (down the rows of a
// group first), so the tiles in flight on one XCD share the rows of A and the columns of B they
// read.  A bijection of [0, gx gy).
__device__ __forceinline__ void xcd_tile(int &tx, int &ty)
{
	constexpr int NX = 8, GT = 8;
	const int gx = gridDim.x, gy = gridDim.y, total = gx * gy;
	const int pid = blockIdx.y * gx + blockIdx.x;
	const int per = (total + NX - 1) / NX, tall = total - (per - 1) * NX;
	const int x = pid % NX, l = pid / NX;
	const int q = (x < tall) ? (x * per + l) : (tall * per + (x - tall) * (per - 1) + l);
	const int gsz = GT * gx;
	const int g = q / gsz, fy = g * GT, h = (gy - fy < GT) ? (gy - fy) : GT;
	const int r = q - g * gsz;
	ty = fy + r % h;
	tx = r / h;
}

__global__ __launch_bounds__(256, 2) void k_zgemm(int M, int N, int K, const double2 *__restrict__ A, long lda,
                                                const double2 *__restrict__ B, long ldb, double2 *__restrict__ C,
                                                long ldc)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	int tx, ty;
	xcd_tile(tx, ty);
	const int row0 = ty * 64 + (wave >> 1) * 32;
	const int col0 = tx * 128 + (wave & 1) * 64;
	const int li = lane & 15, lk = lane >> 4;
	const double2 *arow[2];
	const double2 *bcol[4];
	bool aok[2], bok[4];
#pragma unroll
	for (int rb = 0; rb < 2; rb++) {
		const int r = row0 + rb * 16 + li;
		aok[rb] = r < M;
		arow[rb] = A + (long)(aok[rb] ? r : 0) * lda;
	}
#pragma unroll
	for (int cb = 0; cb < 4; cb++) {
		const int c = col0 + cb * 16 + li;
		bok[cb] = c < N;
		bcol[cb] = B + (bok[cb] ? c : 0);
	}
	d4 ar[2][4], ai[2][4];
#pragma unroll
	for (int a = 0; a < 2; a++)
#pragma unroll
		for (int b = 0; b < 4; b++) {
			ar[a][b] = d4{0, 0, 0, 0};
			ai[a][b] = d4{0, 0, 0, 0};
		}
	double2 fa[2], fb[4];
	{
		const int kk = lk;
#pragma unroll
		for (int rb = 0; rb < 2; rb++) fa[rb] = (aok[rb] && (kk < K)) ? arow[rb][kk] : make_double2(0, 0);
#pragma unroll
		for (int cb = 0; cb < 4; cb++) fb[cb] = (bok[cb] && (kk < K)) ? bcol[cb][(long)kk * ldb] : make_double2(0, 0);
	}
#pragma unroll 1
	for (int k = 0; k < K; k += 4) {
		double2 na[2], nb4[4];
		const int kk = k + 4 + lk;
#pragma unroll
		for (int rb = 0; rb < 2; rb++) na[rb] = (aok[rb] && (kk < K)) ? arow[rb][kk] : make_double2(0, 0);
#pragma unroll
		for (int cb = 0; cb < 4; cb++) nb4[cb] = (bok[cb] && (kk < K)) ? bcol[cb][(long)kk * ldb] : make_double2(0, 0);
		// four passes over the 8 blocks so that the two MFMAs feeding one accumulator are 8
		// instructions apart (no back-to-back dependency)
#pragma unroll
		for (int rb = 0; rb < 2; rb++)
#pragma unroll
			for (int cb = 0; cb < 4; cb++) ar[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[rb].x, fb[cb].x, ar[rb][cb], 0, 0, 0);
#pragma unroll
		for (int rb = 0; rb < 2; rb++)
#pragma unroll
			for (int cb = 0; cb < 4; cb++) ai[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[rb].x, fb[cb].y, ai[rb][cb], 0, 0, 0);
#pragma unroll
		for (int rb = 0; rb < 2; rb++)
#pragma unroll
			for (int cb = 0; cb < 4; cb++) ar[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(-fa[rb].y, fb[cb].y, ar[rb][cb], 0, 0, 0);
#pragma unroll
		for (int rb = 0; rb < 2; rb++)
#pragma unroll
			for (int cb = 0; cb < 4; cb++) ai[rb][cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[rb].y, fb[cb].x, ai[rb][cb], 0, 0, 0);
#pragma unroll
		for (int q = 0; q < 2; q++) fa[q] = na[q];
#pragma unroll
		for (int q = 0; q < 4; q++) fb[q] = nb4[q];
	}
#pragma unroll
	for (int rb = 0; rb < 2; rb++) {
#pragma unroll
		for (int cb = 0; cb < 4; cb++) {
			const int c = col0 + cb * 16 + li;
			if (c >= N) continue;
#pragma unroll
			for (int v = 0; v < 4; v++) {
				const int r = row0 + rb * 16 + lk + 4 * v;
				if (r >= M) continue;
				double2 x = C[(long)r * ldc + c];
				x.x -= ar[rb][cb][v];
				x.y -= ai[rb][cb][v];
				C[(long)r * ldc + c] = x;
			}
		}
	}
}

// Few right-hand sides (nrhs <= 4): triangular solve of one NB x NB diagonal block by a single
// wavefront, lane r holding row r (its NB triangle entries in registers) and x[r]; the column
// sweep broadcasts x[q] with a lane shuffle.  Unit lower when LOWER, else non-unit upper.
template <bool LOWER, typename TM>
__global__ __launch_bounds__(64) void k_trsv(const TM *__restrict__ T, long ldt, double2 *X, long ldx, int nb, int nrhs)
{
	const int r = threadIdx.x;
	double2 t[NB];
#pragma unroll
	for (int q = 0; q < NB; q++) t[q] = ((r < nb) && (q < nb)) ? ld2(T[(long)r * ldt + q]) : make_double2(0, 0);
	const double2 dinv = (r < nb) ? cinv(ld2(T[(long)r * ldt + r])) : make_double2(0, 0);
	for (int j = 0; j < nrhs; j++) {
		double2 x = (r < nb) ? X[(long)r * ldx + j] : make_double2(0, 0);
		if (LOWER) {
#pragma unroll
			for (int q = 0; q < NB; q++) {
				const double2 xq = make_double2(__shfl(x.x, q), __shfl(x.y, q));
				if (r > q) {
					const double2 p = cmul(t[q], xq);
					x.x -= p.x;
					x.y -= p.y;
				}
			}
		}
		else {
#pragma unroll
			for (int q = NB - 1; q >= 0; q--) {
				if (r == q) x = cmul(x, dinv);
				const double2 xq = make_double2(__shfl(x.x, q), __shfl(x.y, q));
				if (r < q) {
					const double2 p = cmul(t[q], xq);
					x.x -= p.x;
					x.y -= p.y;
				}
			}
		}
		if (r < nb) X[(long)r * ldx + j] = x;
	}
}

// Y[M x nrhs] -= A[M x K] X[K x nrhs] for K <= NB and nrhs <= 4: 16 lanes per row, 16 rows per
// 256-thread workgroup, partial dot products reduced by shuffles.
template <typename TM>
__global__ __launch_bounds__(256) void k_gemv(int M, int K, int nrhs, const TM *__restrict__ A, long lda,
                                               const double2 *__restrict__ Xv, long ldx, double2 *Y, long ldy)
{
	__shared__ double2 xs[NB * 4];
	for (int e = threadIdx.x; e < K * nrhs; e += 256) {
		const int k = e / nrhs, j = e - k * nrhs;
		xs[k * 4 + j] = Xv[(long)k * ldx + j];
	}
	__syncthreads();
	const int row = blockIdx.x * 16 + (threadIdx.x >> 4);
	const int sub = threadIdx.x & 15;
	double2 acc[4];
	for (int j = 0; j < 4; j++) acc[j] = make_double2(0, 0);
	if (row < M) {
		for (int k = sub; k < K; k += 16) {
			const double2 a = ld2(A[(long)row * lda + k]);
			for (int j = 0; j < nrhs; j++) {
				const double2 p = cmul(a, xs[k * 4 + j]);
				acc[j].x += p.x;
				acc[j].y += p.y;
			}
		}
	}
	for (int j = 0; j < nrhs; j++) {
		for (int off = 8; off > 0; off >>= 1) {
			acc[j].x += __shfl_xor(acc[j].x, off);
			acc[j].y += __shfl_xor(acc[j].y, off);
		}
		if ((row < M) && (sub == 0)) {
			double2 y = Y[(long)row * ldy + j];
			y.x -= acc[j].x;
			y.y -= acc[j].y;
			Y[(long)row * ldy + j] = y;
		}
	}
}

__global__ void k_iota(int *p, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = i;
}

// dst[k][:] = src[perm[k]][:]
__global__ void k_gather_rows(const double2 *src, long lds, double2 *dst, long ldd, const int *perm, int n, int ncol)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * ncol) return;
	const int k = (int)(e / ncol), j = (int)(e - (long)k * ncol);
	dst[(long)k * ldd + j] = src[(long)perm[k] * lds + j];
}

__global__ void k_copy_rows(const double2 *src, long lds, double2 *dst, long ldd, int n, int ncol)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * ncol) return;
	const int k = (int)(e / ncol), j = (int)(e - (long)k * ncol);
	dst[(long)k * ldd + j] = src[(long)k * lds + j];
}

struct lu_work {
	size_t cap = 0;
	char *base = nullptr;
	panel_ws ws{};
	int *perm = nullptr;
	int nperm = 0;
	int ncu = 0;
	bool attr = false;
};
lu_work g_lw;

void lu_release_at_exit();

// the panel's exchange buffers, laid out in one block of PANEL_WS_BYTES
constexpr size_t PANEL_WS_BYTES = 2 * (size_t)PANEL_PMAX * NB * sizeof(double2) + 2 * NB * sizeof(double2) + SLOT_BYTES +
                                  2 * PANEL_PMAX * sizeof(int) + 256;

panel_ws carve_panel_ws(char *p)
{
	panel_ws w;
	w.pubrow = (double2 *)p;
	p += 2 * (size_t)PANEL_PMAX * NB * sizeof(double2);
	w.diagrow = (double2 *)p;
	p += 2 * NB * sizeof(double2);
	w.slot = (double *)p;
	p += SLOT_BYTES;
	w.pubi = (int *)p;
	p += 2 * PANEL_PMAX * sizeof(int);
	w.info = (int *)p;
	return w;
}

// per device: the panel's dynamic-LDS attribute and the CU count (no allocation)
void ensure_attr()
{
	if (!g_lw.attr) {
		HIPCHK(hipFuncSetAttribute((const void *)k_panel<double2>, hipFuncAttributeMaxDynamicSharedMemorySize,
		                           PANEL_RMAX * SROW * (int)sizeof(double2)));
		HIPCHK(hipFuncSetAttribute((const void *)k_panel<float2>, hipFuncAttributeMaxDynamicSharedMemorySize,
		                           PANEL_RMAX * SROW * (int)sizeof(double2)));
		int dev = 0;
		HIPCHK(hipGetDevice(&dev));
		HIPCHK(hipDeviceGetAttribute(&g_lw.ncu, hipDeviceAttributeMultiprocessorCount, dev));
		g_lw.attr = true;
	}
}

void ensure_ws(int n)
{
	ensure_attr();
	if (g_lw.base == nullptr) {
		pgpu_at_exit(lu_release_at_exit);
		HIPCHK(hipMalloc((void **)&g_lw.base, PANEL_WS_BYTES));
		g_lw.ws = carve_panel_ws(g_lw.base);
	}
	if (g_lw.nperm < n) {
		if (g_lw.perm) HIPCHK(hipFree(g_lw.perm));
		HIPCHK(hipMalloc((void **)&g_lw.perm, (size_t)n * sizeof(int)));
		g_lw.nperm = n;
	}
}

// The products C -= A B of the factorisation and of the triangular inversion are plain GEMMs:
// rocBLAS (zgemm / cgemm, Tensile kernels tuned for gfx950) runs them by default; PEEC_GEMM=hip
// selects the hand-written MFMA kernels k_zgemm / k_cgemm32.  Row-major C -= A B is issued as the
// column-major C^T = -B^T A^T + C^T.
bool gemm_hip()
{
	const char *e = getenv("PEEC_GEMM");   // read per call (a few hundred calls per factorisation)
	return e && (strcmp(e, "hip") == 0);
}

// one rocBLAS handle per stream (a handle's workspace must not serve two streams at once), bound
// to that stream; destroyed with the LU workspace at process exit
constexpr int BLAS_MAX = 32;
std::pair<hipStream_t, rocblas_handle> g_blas[BLAS_MAX];
int g_nblas = 0;

void lu_release_at_exit()
{
	for (int k = 0; k < g_nblas; k++) rocblas_destroy_handle(g_blas[k].second);
	g_nblas = 0;
	if (g_lw.base) HIPCHK(hipFree(g_lw.base));
	if (g_lw.perm) HIPCHK(hipFree(g_lw.perm));
	g_lw.base = nullptr;
	g_lw.perm = nullptr;
	g_lw.nperm = 0;
}

rocblas_handle blas(hipStream_t s)
{
	for (int k = 0; k < g_nblas; k++)
		if (g_blas[k].first == s) return g_blas[k].second;
	if (g_nblas == BLAS_MAX) {
		// streams come and go with device mirrors: recycle the oldest handle's slot
		rocblas_destroy_handle(g_blas[0].second);
		for (int k = 1; k < BLAS_MAX; k++) g_blas[k - 1] = g_blas[k];
		g_nblas--;
	}
	rocblas_handle h = nullptr;
	if ((rocblas_create_handle(&h) != rocblas_status_success) || (rocblas_set_stream(h, s) != rocblas_status_success)) {
		fprintf(stderr, "*** rocblas_create_handle failed\n");
		abort();
	}
	pgpu_at_exit(lu_release_at_exit);
	g_blas[g_nblas++] = std::make_pair(s, h);
	return h;
}

void blas_check(rocblas_status st, const char *what)
{
	if (st != rocblas_status_success) {
		fprintf(stderr, "*** %s failed: %s\n", what, rocblas_status_to_string(st));
		abort();
	}
}

void launch_gemm(hipStream_t s, int M, int N, int K, const double2 *A, long lda, const double2 *B, long ldb,
                 double2 *C, long ldc)
{
	if ((M <= 0) || (N <= 0) || (K <= 0)) return;
	if (!gemm_hip()) {
		rocblas_handle h = blas(s);
		const rocblas_double_complex alpha{-1.0, 0.0}, beta{1.0, 0.0};
		blas_check(rocblas_zgemm(h, rocblas_operation_none, rocblas_operation_none, N, M, K, &alpha,
		                         (const rocblas_double_complex *)B, (rocblas_int)ldb, (const rocblas_double_complex *)A,
		                         (rocblas_int)lda, &beta, (rocblas_double_complex *)C, (rocblas_int)ldc),
		           "rocblas_zgemm");
		return;
	}
	dim3 grid((N + 127) / 128, (M + 63) / 64);
	hipLaunchKernelGGL(k_zgemm, grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc);
}

// C[M x N] -= A[M x K] B[K x N] with complex FP32 storage, FP32 MFMA products accumulated in FP32.
// Workgroup tile 128 x 128 (four waves of 64 x 64), K in steps of 8 staged through LDS in two
// stages: the next step's global loads are issued before the current step's products.  A's rows
// and B's columns are stored in LDS as real and imaginary planes, k-contiguous with a 12-float
// stride (16-byte reads without bank conflicts).  Products on v_mfma_f32_32x32x2f32: lane l feeds
// row (A) / column (B) l % 32 and, in step t of a K step, k = 4 (l / 32) + t, so one 16-byte LDS
// read serves four products; D register v of lane l holds row 8 (v / 4) + 4 (l / 32) + v % 4,
// column l % 32.  Complex = four real products into (re, im) accumulators.  The epilogue reads a
// 32 x 32 block's C entries before writing any of them.
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4v __attribute__((ext_vector_type(4)));
constexpr int CG_BM = 128, CG_BN = 128, CG_KT = 8, CG_LS = 12;

__global__ __launch_bounds__(256, 2) void k_cgemm32(int M, int N, int K, const float2 *__restrict__ A, long lda,
                                                    const float2 *__restrict__ B, long ldb, float2 *__restrict__ C, long ldc)
{
	__shared__ __attribute__((aligned(16))) float sA[2][2][CG_BM * CG_LS];
	__shared__ __attribute__((aligned(16))) float sB[2][2][CG_BN * CG_LS];
	const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
	int tx, ty;
	xcd_tile(tx, ty);
	const int bm0 = ty * CG_BM, bn0 = tx * CG_BN;
	const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
	// global loads: thread t takes row t / 2 of A's tile (k half t % 2) and column t % 128 of B's
	// (k half t / 128), four k each
	const int ar = tid >> 1, akh = (tid & 1) * 4;
	const int bj = tid & 127, bkh = (tid >> 7) * 4;
	const int agr = bm0 + ar, bgc = bn0 + bj;
	const bool aok = agr < M, bok = bgc < N;
	const float2 *ap = A + (long)(aok ? agr : 0) * lda;
	const float2 *bp = B + (bok ? bgc : 0);
	float2 ra[4], rb[4];
	auto gload = [&](int k0) {
#pragma unroll
		for (int u = 0; u < 4; u++) {
			const int ka = k0 + akh + u, kb = k0 + bkh + u;
			ra[u] = (aok && (ka < K)) ? ap[ka] : make_float2(0, 0);
			rb[u] = (bok && (kb < K)) ? bp[(long)kb * ldb] : make_float2(0, 0);
		}
	};
	auto sstore = [&](int st) {
		*(f32x4v *)&sA[st][0][ar * CG_LS + akh] = f32x4v{ra[0].x, ra[1].x, ra[2].x, ra[3].x};
		*(f32x4v *)&sA[st][1][ar * CG_LS + akh] = f32x4v{ra[0].y, ra[1].y, ra[2].y, ra[3].y};
		*(f32x4v *)&sB[st][0][bj * CG_LS + bkh] = f32x4v{rb[0].x, rb[1].x, rb[2].x, rb[3].x};
		*(f32x4v *)&sB[st][1][bj * CG_LS + bkh] = f32x4v{rb[0].y, rb[1].y, rb[2].y, rb[3].y};
	};
	f32x16 cr[2][2], ci[2][2];
#pragma unroll
	for (int a = 0; a < 2; a++)
#pragma unroll
		for (int b = 0; b < 2; b++)
#pragma unroll
			for (int v = 0; v < 16; v++) cr[a][b][v] = ci[a][b][v] = 0;
	const int li = lane & 31, lh = (lane >> 5) * 4;
	const int nkt = (K + CG_KT - 1) / CG_KT;
	gload(0);
	sstore(0);
	__syncthreads();
	for (int kt = 0; kt < nkt; kt++) {
		const int st = kt & 1;
		if (kt + 1 < nkt) gload((kt + 1) * CG_KT);
		f32x4v xr[2], xi[2], xn[2], yr[2], yi[2];
#pragma unroll
		for (int b = 0; b < 2; b++) {
			xr[b] = *(const f32x4v *)&sA[st][0][(wm + b * 32 + li) * CG_LS + lh];
			xi[b] = *(const f32x4v *)&sA[st][1][(wm + b * 32 + li) * CG_LS + lh];
			xn[b] = -xi[b];
			yr[b] = *(const f32x4v *)&sB[st][0][(wn + b * 32 + li) * CG_LS + lh];
			yi[b] = *(const f32x4v *)&sB[st][1][(wn + b * 32 + li) * CG_LS + lh];
		}
#pragma unroll
		for (int t = 0; t < 4; t++)
#pragma unroll
			for (int a = 0; a < 2; a++)
#pragma unroll
				for (int b = 0; b < 2; b++) {
					cr[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(xr[a][t], yr[b][t], cr[a][b], 0, 0, 0);
					cr[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(xn[a][t], yi[b][t], cr[a][b], 0, 0, 0);
					ci[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(xr[a][t], yi[b][t], ci[a][b], 0, 0, 0);
					ci[a][b] = __builtin_amdgcn_mfma_f32_32x32x2f32(xi[a][t], yr[b][t], ci[a][b], 0, 0, 0);
				}
		if (kt + 1 < nkt) sstore(st ^ 1);
		__syncthreads();
	}
#pragma unroll
	for (int a = 0; a < 2; a++)
#pragma unroll
		for (int b = 0; b < 2; b++) {
			const int c = bn0 + wn + b * 32 + li;
			float2 cv[16];
#pragma unroll
			for (int v = 0; v < 16; v++) {
				const int r = bm0 + wm + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
				cv[v] = ((r < M) && (c < N)) ? C[(long)r * ldc + c] : make_float2(0, 0);
			}
#pragma unroll
			for (int v = 0; v < 16; v++) {
				const int r = bm0 + wm + a * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
				if ((r < M) && (c < N)) C[(long)r * ldc + c] = make_float2(cv[v].x - cr[a][b][v], cv[v].y - ci[a][b][v]);
			}
		}
}

void launch_update(hipStream_t s, int M, int N, int K, const double2 *A, long lda, const double2 *B, long ldb, double2 *C,
                   long ldc, bool own)
{
	if (own) {
		if ((M <= 0) || (N <= 0) || (K <= 0)) return;
		dim3 grid((N + 127) / 128, (M + 63) / 64);
		hipLaunchKernelGGL(k_zgemm, grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc);
		return;
	}
	launch_gemm(s, M, N, K, A, lda, B, ldb, C, ldc);
}

void launch_update(hipStream_t s, int M, int N, int K, const float2 *A, long lda, const float2 *B, long ldb, float2 *C,
                   long ldc, bool own)
{
	if ((M <= 0) || (N <= 0) || (K <= 0)) return;
	if (!own && !gemm_hip()) {
		rocblas_handle h = blas(s);
		const rocblas_float_complex alpha{-1.0f, 0.0f}, beta{1.0f, 0.0f};
		blas_check(rocblas_cgemm(h, rocblas_operation_none, rocblas_operation_none, N, M, K, &alpha,
		                         (const rocblas_float_complex *)B, (rocblas_int)ldb, (const rocblas_float_complex *)A,
		                         (rocblas_int)lda, &beta, (rocblas_float_complex *)C, (rocblas_int)ldc),
		           "rocblas_cgemm");
		return;
	}
	dim3 grid((N + CG_BN - 1) / CG_BN, (M + CG_BM - 1) / CG_BM);
	hipLaunchKernelGGL(k_cgemm32, grid, dim3(256), 0, s, M, N, K, A, lda, B, ldb, C, ldc);
}

// ---------------------------------------------------------------- triangular inversion
// Explicit inverses of the triangular factors, for preconditioners applied many times.
constexpr int TI = 64;

__device__ __forceinline__ double2 ti_mul(double2 a, double2 b)
{
	return make_double2((a.x * b.x) - (a.y * b.y), (a.x * b.y) + (a.y * b.x));
}

__device__ __forceinline__ double2 ti_inv(double2 z)
{
	const double n2 = (z.x * z.x) + (z.y * z.y);
	if (n2 <= 0) return make_double2(0, 0);
	return make_double2(z.x / n2, -z.y / n2);
}

template <typename TS>
__global__ void k_ti_zero(TS *X, long cnt)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e < cnt) st2(X[e], make_double2(0, 0));
}

unsigned tiblocks(long n)
{
	return (unsigned)((n + 255) / 256);
}

// Inverses of the TI x TI diagonal blocks of the packed LU factors (unit lower when LOWER,
// else upper), FP64 arithmetic on either storage.  One workgroup per block; thread c solves
// T y = e_c in LDS.  Rows past the end of a short last block act as identity.
template <bool LOWER, typename TS, typename TD = TS>
__global__ __launch_bounds__(64) void k_diaginv(const TS *A, long lda, int n, TD *dinv)
{
	__shared__ double2 T[TI * TI];
	__shared__ double2 Y[TI * TI];
	const int b = blockIdx.x, c = threadIdx.x;
	const int r0 = b * TI, nb = min(TI, n - r0);
	for (int r = 0; r < TI; r++) {
		double2 v = make_double2((r == c) ? 1.0 : 0.0, 0);
		if ((r < nb) && (c < nb)) v = ld2(A[(long)(r0 + r) * lda + r0 + c]);
		T[r * TI + c] = v;
	}
	__syncthreads();
	if (LOWER) {
		for (int r = 0; r < TI; r++) {
			double2 acc = make_double2((r == c) ? 1.0 : 0.0, 0);
			for (int q = 0; q < r; q++) {
				const double2 p = ti_mul(T[r * TI + q], Y[q * TI + c]);
				acc.x -= p.x;
				acc.y -= p.y;
			}
			Y[r * TI + c] = acc;
		}
	}
	else {
		for (int r = TI - 1; r >= 0; r--) {
			double2 acc = make_double2((r == c) ? 1.0 : 0.0, 0);
			for (int q = r + 1; q < TI; q++) {
				const double2 p = ti_mul(T[r * TI + q], Y[q * TI + c]);
				acc.x -= p.x;
				acc.y -= p.y;
			}
			Y[r * TI + c] = ti_mul(acc, ti_inv(T[r * TI + r]));
		}
	}
	__syncthreads();
	for (int r = 0; r < TI; r++) st2(dinv[((long)b * TI + r) * TI + c], Y[r * TI + c]);
}

// X[r][c] *= s over a rows x cols block (s = 0 clears, s = -1 negates)
template <typename TS>
__global__ void k_scale_block(TS *X, long ld, int rows, int cols, double sc)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)rows * cols) return;
	const int r = (int)(e / cols), c = (int)(e - (long)r * cols);
	TS &x = X[(long)r * ld + c];
	const double2 v = ld2(x);
	st2(x, (sc == 0) ? make_double2(0, 0) : make_double2(sc * v.x, sc * v.y));
}

// every diagonal TI x TI block of the n x n inverse at once (blockIdx.y: the block; a short last
// block has n - TI y rows and columns)
template <typename TS>
__global__ void k_put_blocks(const TS *D, TS *Out, long ld, int n)
{
	const int b = blockIdx.y, i0 = b * TI, nb = min(TI, n - i0);
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nb * nb) return;
	const int r = e / nb, c = e - r * nb;
	Out[(long)(i0 + r) * ld + i0 + c] = D[(long)b * TI * TI + r * TI + c];
}

template <typename TS>
void scale_block(hipStream_t s, TS *X, long ld, int rows, int cols, double sc)
{
	if ((rows <= 0) || (cols <= 0)) return;
	hipLaunchKernelGGL((k_scale_block<TS>), dim3(tiblocks((long)rows * cols)), dim3(256), 0, s, X, ld, rows, cols, sc);
}

// C = alpha * B * A (right) or alpha * A * B (left) for a triangular A, all column-major as
// rocBLAS sees them (rocblas_?trmm, out of place); upper = A's fill in that view
void trmm(hipStream_t s, bool right, bool upper, int m, int n, double alpha, const float2 *A, long lda, const float2 *B,
          long ldb, float2 *C, long ldc)
{
	rocblas_handle h = blas(s);
	const rocblas_float_complex a{(float)alpha, 0.0f};
	blas_check(rocblas_ctrmm(h, right ? rocblas_side_right : rocblas_side_left, upper ? rocblas_fill_upper : rocblas_fill_lower,
	                         rocblas_operation_none, rocblas_diagonal_non_unit, m, n, &a, (const rocblas_float_complex *)A,
	                         (rocblas_int)lda, (const rocblas_float_complex *)B, (rocblas_int)ldb, (rocblas_float_complex *)C,
	                         (rocblas_int)ldc),
	           "rocblas_ctrmm");
}

void trmm(hipStream_t s, bool right, bool upper, int m, int n, double alpha, const double2 *A, long lda, const double2 *B,
          long ldb, double2 *C, long ldc)
{
	rocblas_handle h = blas(s);
	const rocblas_double_complex a{alpha, 0.0};
	blas_check(rocblas_ztrmm(h, right ? rocblas_side_right : rocblas_side_left, upper ? rocblas_fill_upper : rocblas_fill_lower,
	                         rocblas_operation_none, rocblas_diagonal_non_unit, m, n, &a, (const rocblas_double_complex *)A,
	                         (rocblas_int)lda, (const rocblas_double_complex *)B, (rocblas_int)ldb, (rocblas_double_complex *)C,
	                         (rocblas_int)ldc),
	           "rocblas_ztrmm");
}

// Out[i0.., i0..] (n x n, leading dimension ld, zero on entry) = inverse of the unit-lower (lower)
// or upper part of the packed factor T over the same rows/columns; scr holds at least
// ceil(n/2)^2 + TI n entries.  Recursive: inv([T11 0; T21 T22]) = [O11 0; -O22 T21 O11  O22]
// (lower; the upper case mirrors it), halves split on TI boundaries, the diagonal 64 x 64 blocks
// inverted directly (k_diaginv).  With rocBLAS both
// products of a level are triangular multiplies (O11 and O22 are triangular), half the flops of
// full GEMMs; row-major X is the column-major X^T, so a row-major lower triangle is an upper
// one to rocBLAS and a right-hand product becomes a left-hand one.  PEEC_GEMM=hip runs them as
// full products on the hand-written GEMM (the zeros above the diagonal included).
template <typename TS>
void tri_inv(hipStream_t s, bool lower, const TS *T, long ld, TS *Out, int i0, int n, const TS *dinv, TS *scr)
{
	if (n <= TI) return;   // the diagonal blocks are in place already (k_put_blocks)
	const int n1 = ((n / 2 + TI - 1) / TI) * TI, n2 = n - n1;
	tri_inv(s, lower, T, ld, Out, i0, n1, dinv, scr);
	tri_inv(s, lower, T, ld, Out, i0 + n1, n2, dinv, scr);
	const TS *O11 = Out + (long)i0 * ld + i0;
	const TS *O22 = Out + (long)(i0 + n1) * ld + (i0 + n1);
	if (!gemm_hip()) {
		if (lower) {
			// scr = -T21 O11 (n2 x n1, ld n1); O21 = O22 scr
			TS *O21 = Out + (long)(i0 + n1) * ld + i0;
			trmm(s, false, true, n1, n2, -1.0, O11, ld, T + (long)(i0 + n1) * ld + i0, ld, scr, n1);
			trmm(s, true, true, n1, n2, 1.0, O22, ld, (const TS *)scr, n1, O21, ld);
		}
		else {
			// scr = -T12 O22 (n1 x n2, ld n2); O12 = O11 scr
			TS *O12 = Out + (long)i0 * ld + (i0 + n1);
			trmm(s, false, false, n2, n1, -1.0, O22, ld, T + (long)i0 * ld + (i0 + n1), ld, scr, n2);
			trmm(s, true, false, n2, n1, 1.0, O11, ld, (const TS *)scr, n2, O12, ld);
		}
		return;
	}
	if (lower) {
		// C = -T21 inv(T11) (n2 x n1); Out21 = -(inv(T22) C), then negated
		scale_block(s, scr, n1, n2, n1, 0);
		launch_update(s, n2, n1, n1, T + (long)(i0 + n1) * ld + i0, ld, O11, ld, scr, n1, false);
		TS *O21 = Out + (long)(i0 + n1) * ld + i0;
		launch_update(s, n2, n1, n2, O22, ld, (const TS *)scr, n1, O21, ld, false);
		scale_block(s, O21, ld, n2, n1, -1);
	}
	else {
		// C = -T12 inv(T22) (n1 x n2); Out12 = -(inv(T11) C), then negated
		scale_block(s, scr, n2, n1, n2, 0);
		launch_update(s, n1, n2, n2, T + (long)i0 * ld + (i0 + n1), ld, O22, ld, scr, n2, false);
		TS *O12 = Out + (long)i0 * ld + (i0 + n1);
		launch_update(s, n1, n2, n1, O11, ld, (const TS *)scr, n2, O12, ld, false);
		scale_block(s, O12, ld, n1, n2, -1);
	}
}

template <typename TS>
void tri_inverse_t(hipStream_t s, int n, const TS *LU, int lda, int lower, TS *Out, TS *scr, TS *dinv)
{
	const int nblk = (n + TI - 1) / TI;
	if (lower) hipLaunchKernelGGL((k_diaginv<true, TS>), dim3(nblk), dim3(64), 0, s, LU, (long)lda, n, dinv);
	else hipLaunchKernelGGL((k_diaginv<false, TS>), dim3(nblk), dim3(64), 0, s, LU, (long)lda, n, dinv);
	hipLaunchKernelGGL((k_ti_zero<TS>), dim3(tiblocks((long)n * n)), dim3(256), 0, s, Out, (long)n * n);
	// the diagonal blocks in one launch; the recursion's leaves then have nothing to do
	hipLaunchKernelGGL((k_put_blocks<TS>), dim3((TI * TI + 255) / 256, nblk), dim3(256), 0, s, (const TS *)dinv, Out, (long)lda, n);
	tri_inv(s, lower != 0, LU, (long)lda, Out, 0, n, (const TS *)dinv, scr);
}

}  // namespace

extern "C" void dlu_zgemm_sub(void *stream, int M, int N, int K, const d_complex_t *A, long lda,
                              const d_complex_t *B, long ldb, d_complex_t *C, long ldc)
{
	launch_gemm((hipStream_t)stream, M, N, K, (const double2 *)A, lda, (const double2 *)B, ldb, (double2 *)C, ldc);
}


namespace {
// fills n 32-bit words with 0xFFFFFFFF: -1 as int, NaN as either half of a double (a kernel
// rather than hipMemsetAsync, so the sequence stays valid under stream capture)
__global__ void k_fill_ones(unsigned *p, int n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) p[i] = 0xFFFFFFFFu;
}

// The blocked factorisation on a given workspace: panel exchange buffers ws, permutation
// scratch perm_ws (n ints, used when dperm is NULL); own_gemm selects the hand-written trailing
// update (no library call, no allocation) whatever PEEC_GEMM says.
// Columns per trailing update of the FP32 factorisation (PEEC_LU_KB: a multiple of 64 up to 512;
// 256 by default): a block of that many columns is factored panel by panel, each panel updating
// only the rest of the block, and the columns beyond it take one update with K = the block's
// width, where a library GEMM runs nearer its peak than at K = 64.
int lu_kb()
{
	static int kb = -1;
	if (kb < 0) {
		const char *e = getenv("PEEC_LU_KB");
		const int v = e ? atoi(e) : 256;
		kb = ((v >= NB) && (v <= 8 * NB) && (v % NB == 0)) ? v : 256;
	}
	return kb;
}

template <typename TS>
int lu_factor_ws(hipStream_t s, int n, TS *A, int lda, int *dpiv, int *dperm, pgpu_stats_t *st, const panel_ws &wsp,
                 int *perm_ws, bool own_gemm)
{
	if (n <= 0) return -1;
	int *perm = (dperm != nullptr) ? dperm : perm_ws;
	hipLaunchKernelGGL(k_fill_ones, dim3(1), dim3(64), 0, s, (unsigned *)wsp.info, 2);   // -1, -1
	hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, perm, n);
	hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, dpiv, n);
	hipEvent_t e0 = nullptr, e1 = nullptr;
	if (st) {   // timing-only events: no system-scope fence around the timed GEMM
		HIPCHK(hipEventCreateWithFlags(&e0, hipEventDisableSystemFence));
		HIPCHK(hipEventCreateWithFlags(&e1, hipEventDisableSystemFence));
	}
	const int pmax = (g_lw.ncu < PANEL_PMAX) ? g_lw.ncu : PANEL_PMAX;
	const long ldal = lda;
	// columns per trailing update: FP32 storage takes blocks of lu_kb() columns (panels of NB
	// inside), FP64 one panel
	const int kb = (sizeof(TS) == sizeof(float2)) ? lu_kb() : NB;
	for (int k0 = 0; k0 < n; k0 += NB) {
		const int nb = (n - k0 < NB) ? (n - k0) : NB;
		const int m = n - k0;
		int P = (m + PANEL_RMAX - 1) / PANEL_RMAX;
		if (P < 1) P = 1;
		if (P > pmax) {
			// a size limit, not a property of the matrix: reported as the step n, which no
			// zero pivot can have (those are 0 .. n-1), so callers that test info >= 0 stop
			fprintf(stderr, "*** LU factorisation: n = %d exceeds the panel capacity (%d rows)\n", n, pmax * PANEL_RMAX);
			return n;
		}
		int R = (m + P - 1) / P;
		int *pivp = dpiv;
		int nn = n, kk0 = k0, nbb = nb;
		panel_ws ws = wsp;
		TS *Ap = A;
		long ldp = ldal;
		void *kargs[] = {(void *)&Ap, (void *)&ldp, (void *)&nn, (void *)&kk0, (void *)&nbb,
		                 (void *)&R, (void *)&pivp, (void *)&ws};
		hipLaunchKernelGGL(k_fill_ones, dim3((SLOT_BYTES / 4 + 255) / 256), dim3(256), 0, s, (unsigned *)wsp.slot,
		                   SLOT_BYTES / 4);   // every slot NaN
		HIPCHK(hipLaunchCooperativeKernel((const void *)k_panel<TS>, dim3(P), dim3(PANEL_T), kargs,
		                                  (unsigned)(R * SROW * sizeof(double2)), s));
		const int nout = n - nb;
		hipLaunchKernelGGL(k_laswp<TS>, dim3((nout + 1 + 255) / 256), dim3(256), 0, s, A, ldal, n, k0, nb, dpiv, perm,
		                   (const int *)wsp.info);
		// the block's columns right of this panel: their U rows, and their rows below updated
		const int b0 = (k0 / kb) * kb;                 // the block of this panel: [b0, be)
		const int be = (b0 + kb < n) ? (b0 + kb) : n;
		const int nbr = be - k0 - nb;
		if (nbr > 0) {
			TS *A11 = A + (long)k0 * lda + k0;
			TS *A12 = A + (long)k0 * lda + k0 + nb;
			hipLaunchKernelGGL((k_trsm<true, TS>), dim3((nbr + 63) / 64), dim3(64), 0, s, A11, ldal, A12, ldal, nb, nbr);
			launch_update(s, n - k0 - nb, nbr, nb, (const TS *)(A + (long)(k0 + nb) * lda + k0), ldal, (const TS *)A12, ldal,
			              A + (long)(k0 + nb) * lda + k0 + nb, ldal, own_gemm);
		}
		if (k0 + nb < be) continue;
		// the block is factored: its U rows of the columns beyond it (a 64-row triangular solve
		// per panel, each followed by the update of the block's later rows), then the trailing
		// update with K = the block's width
		const int nr = n - be;
		if (nr > 0) {
			for (int j0 = b0; j0 < be; j0 += NB) {
				const int jb = (be - j0 < NB) ? (be - j0) : NB;
				TS *X = A + (long)j0 * lda + be;
				hipLaunchKernelGGL((k_trsm<true, TS>), dim3((nr + 63) / 64), dim3(64), 0, s, A + (long)j0 * lda + j0, ldal, X,
				                   ldal, jb, nr);
				const int mr = be - j0 - jb;
				if (mr > 0)
					launch_update(s, mr, nr, jb, (const TS *)(A + (long)(j0 + jb) * lda + j0), ldal, (const TS *)X, ldal,
					              A + (long)(j0 + jb) * lda + be, ldal, own_gemm);
			}
			const int kw = be - b0;
			TS *A21 = A + (long)be * lda + b0;
			TS *A12 = A + (long)b0 * lda + be;
			TS *A22 = A + (long)be * lda + be;
			if (st) HIPCHK(hipEventRecord(e0, s));
			launch_update(s, nr, nr, kw, (const TS *)A21, ldal, (const TS *)A12, ldal, A22, ldal, own_gemm);
			if (st) {
				HIPCHK(hipEventRecord(e1, s));
				HIPCHK(hipEventSynchronize(e1));
				float ms = 0;
				HIPCHK(hipEventElapsedTime(&ms, e0, e1));
				st->lu_gemm_ms += ms;
				st->lu_gemm_flop += 8.0 * nr * (double)nr * kw;
				st->lu_gemm_launches++;
			}
		}
	}
	if (st) {
		HIPCHK(hipEventDestroy(e0));
		HIPCHK(hipEventDestroy(e1));
	}
	int info[2] = {-1, -1};
	HIPCHK(hipMemcpyAsync(info, wsp.info, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	// a pivot exchange that gave up means the cooperative grid was not co-resident: a runtime
	// failure like a failed HIP call, never reported as a singular matrix
	if (info[1] >= 0) pgpu_fatal("LU panel: pivot exchange between workgroups timed out");
	return info[0];
}

template <typename TS>
int lu_factor_t(hipStream_t s, int n, TS *A, int lda, int *dpiv, int *dperm, pgpu_stats_t *st)
{
	if (n <= 0) return -1;
	ensure_ws(n);
	return lu_factor_ws<TS>(s, n, A, lda, dpiv, dperm, st, g_lw.ws, g_lw.perm, false);
}

}  // namespace

extern "C" int dlu_factor(void *stream, int n, d_complex_t *dA_, int lda, int *dpiv, int *dperm, pgpu_stats_t *st)
{
	return lu_factor_t<double2>((hipStream_t)stream, n, (double2 *)dA_, lda, dpiv, dperm, st);
}

// The same factorisation with the matrix stored in FP32: panels, pivot search and TRSM in FP64
// arithmetic on FP32 storage, trailing updates as FP32 complex GEMMs on MFMA.
extern "C" int slu_factor(void *stream, int n, void *dA, int lda, int *dpiv, int *dperm, pgpu_stats_t *st)
{
	return lu_factor_t<float2>((hipStream_t)stream, n, (float2 *)dA, lda, dpiv, dperm, st);
}

// X[0..nb) <- D X[0..nb) for one 64 x 64 block D (row-major, FP64) and nrhs <= 4 columns of X:
// the block is staged in LDS, thread t forms row t % 64 of column t / 64
__global__ __launch_bounds__(256) void k_dmv(const double2 *__restrict__ D, double2 *X, long ldx, int nb, int nrhs)
{
	__shared__ double2 xs[TI * 4];
	const int t = threadIdx.x, r = t & 63, j = t >> 6;
	if ((r < nb) && (j < nrhs)) xs[r * 4 + j] = X[(long)r * ldx + j];
	__syncthreads();
	if ((r < nb) && (j < nrhs)) {
		double ar = 0, ai = 0;
		const double2 *d = D + (long)r * TI;
		for (int c = 0; c < nb; c++) {
			const double2 a = d[c], x = xs[c * 4 + j];
			ar = fma(a.x, x.x, ar);
			ar = fma(-a.y, x.y, ar);
			ai = fma(a.x, x.y, ai);
			ai = fma(a.y, x.x, ai);
		}
		X[(long)r * ldx + j] = make_double2(ar, ai);
	}
}

// B <- A^-1 B (FP64 right-hand sides) with FP32-stored factors from slu_factor, nrhs <= 4 per
// pass; dwork holds n * nrhs entries.
extern "C" void slu_solve_perm(void *stream, int n, const void *dA_, int lda, const int *dperm, d_complex_t *dB_, int ldb,
                               int nrhs, d_complex_t *dwork)
{
	hipStream_t s = (hipStream_t)stream;
	const float2 *A = (const float2 *)dA_;
	double2 *B0 = (double2 *)dB_;
	double2 *W = (double2 *)dwork;
	for (int j0 = 0; j0 < nrhs; j0 += 4) {
		const int nr = (nrhs - j0 < 4) ? (nrhs - j0) : 4;
		double2 *B = B0 + j0;
		const long tot = (long)n * nr;
		hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, (const double2 *)B, (long)ldb, W,
		                   (long)nr, dperm, n, nr);
		hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, (const double2 *)W, (long)nr, B,
		                   (long)ldb, n, nr);
		for (int r0 = 0; r0 < n; r0 += NB) {
			const int nb = (n - r0 < NB) ? (n - r0) : NB;
			hipLaunchKernelGGL((k_trsv<true, float2>), dim3(1), dim3(64), 0, s, A + (long)r0 * lda + r0, (long)lda,
			                   B + (long)r0 * ldb, (long)ldb, nb, nr);
			const int rest = n - r0 - nb;
			if (rest > 0)
				hipLaunchKernelGGL(k_gemv<float2>, dim3((rest + 15) / 16), dim3(256), 0, s, rest, nb, nr,
				                   A + (long)(r0 + nb) * lda + r0, (long)lda, (const double2 *)(B + (long)r0 * ldb), (long)ldb,
				                   B + (long)(r0 + nb) * ldb, (long)ldb);
		}
		const int nblk = (n + NB - 1) / NB;
		for (int b = nblk - 1; b >= 0; b--) {
			const int r0 = b * NB;
			const int nb = (n - r0 < NB) ? (n - r0) : NB;
			hipLaunchKernelGGL((k_trsv<false, float2>), dim3(1), dim3(64), 0, s, A + (long)r0 * lda + r0, (long)lda,
			                   B + (long)r0 * ldb, (long)ldb, nb, nr);
			if (r0 > 0)
				hipLaunchKernelGGL(k_gemv<float2>, dim3((r0 + 15) / 16), dim3(256), 0, s, r0, nb, nr, A + r0, (long)lda,
				                   (const double2 *)(B + (long)r0 * ldb), (long)ldb, B, (long)ldb);
		}
	}
}

// The inverses of the 64 x 64 diagonal blocks of slu_factor's factors (unit lower, upper), in
// FP64: dinvL, dinvU hold ceil(n / 64) blocks each.
extern "C" void slu_diag_inverse(void *stream, int n, const void *dA_, int lda, d_complex_t *dinvL, d_complex_t *dinvU)
{
	hipStream_t s = (hipStream_t)stream;
	const float2 *A = (const float2 *)dA_;
	const int nblk = (n + TI - 1) / TI;
	hipLaunchKernelGGL((k_diaginv<true, float2, double2>), dim3(nblk), dim3(64), 0, s, A, (long)lda, n, (double2 *)dinvL);
	hipLaunchKernelGGL((k_diaginv<false, float2, double2>), dim3(nblk), dim3(64), 0, s, A, (long)lda, n, (double2 *)dinvU);
}

// slu_solve_perm with each 64-row step's triangle solved by its block inverse (slu_diag_inverse)
// instead of a substitution: a step is then two parallel products, where the substitution's
// 64 dependent rows made it the chain's longest link.  nrhs <= 4.
extern "C" void slu_solve_perm_di(void *stream, int n, const void *dA_, int lda, const int *dperm, d_complex_t *dB_,
                                  int ldb, int nrhs, d_complex_t *dwork, const d_complex_t *dinvL, const d_complex_t *dinvU)
{
	hipStream_t s = (hipStream_t)stream;
	const float2 *A = (const float2 *)dA_;
	double2 *B = (double2 *)dB_;
	double2 *W = (double2 *)dwork;
	const double2 *DL = (const double2 *)dinvL, *DU = (const double2 *)dinvU;
	const int nr = nrhs;
	const long tot = (long)n * nr;
	hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, (const double2 *)B, (long)ldb, W,
	                   (long)nr, dperm, n, nr);
	hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, (const double2 *)W, (long)nr, B,
	                   (long)ldb, n, nr);
	for (int r0 = 0; r0 < n; r0 += NB) {
		const int nb = (n - r0 < NB) ? (n - r0) : NB;
		hipLaunchKernelGGL(k_dmv, dim3(1), dim3(256), 0, s, DL + (long)(r0 / TI) * TI * TI, B + (long)r0 * ldb, (long)ldb, nb, nr);
		const int rest = n - r0 - nb;
		if (rest > 0)
			hipLaunchKernelGGL(k_gemv<float2>, dim3((rest + 15) / 16), dim3(256), 0, s, rest, nb, nr, A + (long)(r0 + nb) * lda + r0,
			                   (long)lda, (const double2 *)(B + (long)r0 * ldb), (long)ldb, B + (long)(r0 + nb) * ldb, (long)ldb);
	}
	const int nblk = (n + NB - 1) / NB;
	for (int b = nblk - 1; b >= 0; b--) {
		const int r0 = b * NB;
		const int nb = (n - r0 < NB) ? (n - r0) : NB;
		hipLaunchKernelGGL(k_dmv, dim3(1), dim3(256), 0, s, DU + (long)b * TI * TI, B + (long)r0 * ldb, (long)ldb, nb, nr);
		if (r0 > 0)
			hipLaunchKernelGGL(k_gemv<float2>, dim3((r0 + 15) / 16), dim3(256), 0, s, r0, nb, nr, A + r0, (long)lda,
			                   (const double2 *)(B + (long)r0 * ldb), (long)ldb, B, (long)ldb);
	}
}

// B <- A^-1 B using the factors and the row permutation recorded by the last dlu_factor on
// this matrix (perm must be the one returned with these factors: dperm).
extern "C" void dlu_solve_perm(void *stream, int n, const d_complex_t *dA_, int lda, const int *dperm,
                               d_complex_t *dB_, int ldb, int nrhs, d_complex_t *dwork)
{
	hipStream_t s = (hipStream_t)stream;
	const double2 *A = (const double2 *)dA_;
	double2 *B = (double2 *)dB_;
	double2 *W = (double2 *)dwork;
	const long tot = (long)n * nrhs;
	hipLaunchKernelGGL(k_gather_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, B, (long)ldb, W, (long)nrhs,
	                   dperm, n, nrhs);
	hipLaunchKernelGGL(k_copy_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, W, (long)nrhs, B, (long)ldb, n,
	                   nrhs);
	const bool few = (nrhs <= 4);
	// forward: unit lower, block rows of NB
	for (int r0 = 0; r0 < n; r0 += NB) {
		const int nb = (n - r0 < NB) ? (n - r0) : NB;
		const double2 *T = A + (long)r0 * lda + r0;
		if (few) hipLaunchKernelGGL((k_trsv<true, double2>), dim3(1), dim3(64), 0, s, T, (long)lda, B + (long)r0 * ldb, (long)ldb, nb, nrhs);
		else hipLaunchKernelGGL((k_trsm<true, double2>), dim3((nrhs + 63) / 64), dim3(64), 0, s, T, (long)lda, B + (long)r0 * ldb, (long)ldb, nb, nrhs);
		const int rest = n - r0 - nb;
		if (rest <= 0) continue;
		if (few)
			hipLaunchKernelGGL(k_gemv<double2>, dim3((rest + 15) / 16), dim3(256), 0, s, rest, nb, nrhs, A + (long)(r0 + nb) * lda + r0,
			                   (long)lda, B + (long)r0 * ldb, (long)ldb, B + (long)(r0 + nb) * ldb, (long)ldb);
		else
			launch_gemm(s, rest, nrhs, nb, A + (long)(r0 + nb) * lda + r0, lda, B + (long)r0 * ldb, ldb,
			            B + (long)(r0 + nb) * ldb, ldb);
	}
	// backward: upper, from the last block row
	const int nblk = (n + NB - 1) / NB;
	for (int b = nblk - 1; b >= 0; b--) {
		const int r0 = b * NB;
		const int nb = (n - r0 < NB) ? (n - r0) : NB;
		const double2 *T = A + (long)r0 * lda + r0;
		if (few) hipLaunchKernelGGL((k_trsv<false, double2>), dim3(1), dim3(64), 0, s, T, (long)lda, B + (long)r0 * ldb, (long)ldb, nb, nrhs);
		else hipLaunchKernelGGL((k_trsm<false, double2>), dim3((nrhs + 63) / 64), dim3(64), 0, s, T, (long)lda, B + (long)r0 * ldb, (long)ldb, nb, nrhs);
		if (r0 <= 0) continue;
		if (few)
			hipLaunchKernelGGL(k_gemv<double2>, dim3((r0 + 15) / 16), dim3(256), 0, s, r0, nb, nrhs, A + r0, (long)lda,
			                   B + (long)r0 * ldb, (long)ldb, B, (long)ldb);
		else
			launch_gemm(s, r0, nrhs, nb, A + r0, lda, B + (long)r0 * ldb, ldb, B, ldb);
	}
}

// Out (n x n, leading dimension n) = inverse of the unit-lower (lower = 1) or upper part of the
// packed LU factors in LU (leading dimension lda).  dinv receives the 64 x 64 diagonal-block
// inverses (ceil(n/64) blocks), scr is GEMM scratch of at least (ceil(n/2) + 64)^2 entries.
extern "C" void dlu_tri_inverse(void *stream, int n, const d_complex_t *LU, int lda, int lower, d_complex_t *Out,
                                d_complex_t *scr, d_complex_t *dinv)
{
	tri_inverse_t<double2>((hipStream_t)stream, n, (const double2 *)LU, lda, lower, (double2 *)Out, (double2 *)scr,
	                       (double2 *)dinv);
}

// the same on FP32 storage (factors from slu_factor): diagonal blocks inverted in FP64 arithmetic,
// off-diagonal blocks by FP32 MFMA GEMMs; Out, scr and dinv are float2 arrays
extern "C" void slu_tri_inverse(void *stream, int n, const void *LU, int lda, int lower, void *Out, void *scr, void *dinv)
{
	tri_inverse_t<float2>((hipStream_t)stream, n, (const float2 *)LU, lda, lower, (float2 *)Out, (float2 *)scr,
	                      (float2 *)dinv);
}

// ---------------------------------------------------------------- no-allocation entry point
// lu_decomp on a device-resident matrix for callers that cannot allocate (code under stream
// capture, serving loops): the panel's exchange buffers and the permutation scratch live in a
// file-scope device array, one copy per GPU, whose address is looked up once per device.  The
// trailing update is the hand-written k_zgemm (a library GEMM would allocate its handle).
namespace {
constexpr int NOALLOC_NMAX = 32768;   // largest order served from the static array
constexpr size_t NOALLOC_BYTES = PANEL_WS_BYTES + (size_t)NOALLOC_NMAX * sizeof(int);   // 668 KiB
constexpr int NOALLOC_MAXDEV = 16;
}  // namespace

// In host code `scratch` names a host-side stand-in: only the address from hipGetSymbolAddress
// (per device, below) is passed to kernels.
static __device__ __align__(256) char scratch[NOALLOC_BYTES];

namespace {
char *g_scratch_dev[NOALLOC_MAXDEV] = {};   // per device: the array's device address
std::mutex g_scratch_mu;

char *scratch_for_current_device()
{
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess) return nullptr;
	if ((dev < 0) || (dev >= NOALLOC_MAXDEV)) return nullptr;
	std::lock_guard<std::mutex> lk(g_scratch_mu);
	if (g_scratch_dev[dev] == nullptr) {
		void *ptr = nullptr;
		if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
		g_scratch_dev[dev] = (char *)ptr;
	}
	return g_scratch_dev[dev];
}
}  // namespace

extern "C" int lu_decomp_device_noalloc(int n, d_complex_t *dA, int lda, int *dpiv, void *stream)
{
	if (n <= 0) return -1;
	if ((n > NOALLOC_NMAX) || (lda < n) || (dA == nullptr) || (dpiv == nullptr)) {
		fprintf(stderr, "*** lu_decomp_device_noalloc: n = %d (at most %d), lda = %d: invalid arguments\n", n, NOALLOC_NMAX,
		        lda);
		return -2;
	}
	char *base = scratch_for_current_device();
	if (base == nullptr) {
		fprintf(stderr, "*** lu_decomp_device_noalloc: the scratch array's device address is unavailable\n");
		return -2;
	}
	ensure_attr();
	const panel_ws ws = carve_panel_ws(base);
	int *perm = (int *)(base + PANEL_WS_BYTES);
	return lu_factor_ws<double2>((hipStream_t)stream, n, (double2 *)dA, lda, dpiv, nullptr, nullptr, ws, perm, true);
}
