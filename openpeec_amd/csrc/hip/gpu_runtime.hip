// gpu_runtime.hip — device mirror of a peec_t and the device-resident frequency sweep.
//
// A peec_t keeps its host arrays (malloc memory owned by the caller, freed by peec_free); the
// device copies live here in a side table keyed by the peec_t address.  Per frequency point
// the sweep runs entirely on the GPU:
//   Lp fill -> charge-cell fill (P fused or aggregated) -> LU(P), C = P^-1 -> MNA assembly
//   (-jw Lp block, jw C node block, sparse stamps) -> LU(A) -> all port right-hand sides ->
//   port voltages (Z column), branch currents, cell charges.
// Only the Z entries (and optional distributions) come back over PCIe.
// Reference: solve.c:239-363 (sweep), mna.c:175-308 (assembly), potential.c:258-281 (C = P^-1).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>
#include <chrono>
#include "../peec_internal.h"
#include "../../../include/peec_mi355x.h"
#include "hip_check.hpp"
#include "fill_kernels.hpp"

struct pset {   // one packed cell set on the device (cells and their images)
	int n = 0, km = 0;
	dcell_t *c = nullptr, *cm = nullptr;
	dedge_t *e = nullptr, *em = nullptr;
	dpnorm_t *pn = nullptr, *pnm = nullptr;
};

struct pgpu {
	const peec_t *p = nullptr;
	hipStream_t s = nullptr;
	pset seg, chg;
	int ncell = 0;
	int *csoff = nullptr, *csidx = nullptr;
	double *carea = nullptr;
	bool cells_identity = false;
	double2 *lp = nullptr;       // nseg x nseg
	double2 *G = nullptr;        // nchg x nchg (non-identity aggregation only)
	double2 *P = nullptr;        // ncell x ncell, factored in place by pgpu_cmat
	double2 *C = nullptr;        // ncell x ncell
	int *ppiv = nullptr, *pperm = nullptr;
	double2 *cwork = nullptr;    // ncell x ncell solve workspace
	// MNA
	int n = 0;
	double2 *A = nullptr;
	int *apiv = nullptr, *aperm = nullptr;
	double2 *B = nullptr, *Bw = nullptr;   // n x nport right-hand sides and workspace
	float2 *A32 = nullptr;                 // precision form: FP32 copy that is factored
	double2 *dinv32 = nullptr;             // its factors' 64 x 64 diagonal-block inverses (L, then U)
	double2 *Brhs = nullptr, *R = nullptr; // precision form: right-hand sides, residuals
	unsigned long long *cberr = nullptr;  // precision form: componentwise backward error per port
	int *cellrow = nullptr;      // [n] capacitive cell of node row r, or -1
	int *d_st_ri = nullptr, *d_st_ci = nullptr;
	double2 *d_st_v = nullptr;
	int st_cap = 0;
	double2 *qbuf = nullptr;     // ncell x nport cell charges
	// k-factored fill over a wavenumber band (kfill.hip)
	bool kf_ready = false, kf_pending = false;
	double kf_lo = 0, kf_hi = 0;
	double *kf_rec_lp = nullptr, *kf_rec_ig = nullptr;
	int kf_T = 0;
	int2 *kf_list_lp = nullptr, *kf_list_ig = nullptr;
	unsigned *kf_cnt = nullptr;
	int kf_nd_lp = 0, kf_nd_ig = 0;
	double *kf_rho_seg = nullptr, *kf_rho_chg = nullptr;
	// solve() points: right-hand sides, and acceleration = 1's kept LU and GMRES workspace
	double2 *xr = nullptr;       // n x xr_cap solution vectors (one per right-hand side)
	int xr_cap = 0;
	int *rowofcell = nullptr;    // [ncell] MNA row of each capacitive cell's node, or -1
	double2 *alu = nullptr;      // n x n factors of the kept LU
	int *alu_piv = nullptr, *alu_perm = nullptr;
	bool have_lu = false;
	dgmres_ws *gw = nullptr;
};

namespace {
std::mutex g_mu;
std::vector<pgpu *> g_tab;
bool g_dev_init = false;

// Process-exit teardown.  Device buffers, streams and library handles held by the library's
// side tables are released by one atexit handler, registered after the first HIP call of the
// process: handlers run in reverse order of registration, so this one runs before the HIP
// runtime's and a profiler's own exit handlers, while the runtime is still whole.  Modules
// with process-lifetime device state (LU workspace and rocBLAS handles, the accelerated form's
// cache) add a hook; the hooks run last-registered first, then every device mirror is dropped.
constexpr int EXIT_HOOKS_MAX = 16;
void (*g_exit_hooks[EXIT_HOOKS_MAX])(void) = {};
int g_nexit_hooks = 0;
bool g_exit_registered = false, g_exiting = false;

void release_mirror(pgpu *g);

void teardown_at_exit()
{
	g_exiting = true;
	if (hipDeviceSynchronize() != hipSuccess) return;   // a failed device: leave it to the runtime
	for (int k = g_nexit_hooks - 1; k >= 0; k--) g_exit_hooks[k]();
	std::lock_guard<std::mutex> lk(g_mu);
	for (pgpu *g : g_tab) release_mirror(g);
	g_tab.clear();
}

void register_teardown()
{
	if (g_exit_registered) return;
	g_exit_registered = true;
	atexit(teardown_at_exit);
}

template <typename T>
T *dalloc(size_t n)
{
	T *p = nullptr;
	if (n == 0) n = 1;
	HIPCHK(hipMalloc((void **)&p, n * sizeof(T)));
	return p;
}

void dfree(void *p)
{
	if (p) HIPCHK(hipFree(p));
}

void init_device()
{
	if (g_dev_init) return;
	int ndev = 0;
	if ((hipGetDeviceCount(&ndev) != hipSuccess) || (ndev <= 0)) {
		fprintf(stderr, "*** libpeec_mi355x: no HIP device available (this library has no CPU path)\n");
		fflush(stderr);
		abort();
	}
	const char *env = getenv("PEEC_DEVICE");
	const int dev = env ? atoi(env) : 0;
	HIPCHK(hipSetDevice(dev));
	g_dev_init = true;
	register_teardown();
}

void upload_set(pset &ps, const seg_t *src, int n, int gp, double gpz)
{
	ps.n = n;
	ps.km = 0;
	for (int i = 0; i < n; i++) {
		const int k = (src[i].npv > 0) ? CELL_POLY : src[i].vol ? CELL_BAR : (src[i].wid > 0) ? CELL_RIBBON : CELL_WIRE;
		ps.km |= 1 << k;
	}
	if (n <= 0) return;
	for (int m = 0; m < (gp ? 2 : 1); m++) {
		dcell_t *hc = nullptr;
		dedge_t *he = nullptr;
		dpnorm_t *hp = nullptr;
		int ne = 0;
		if (cells_pack(src, n, m, gpz, &hc, &he, &hp, &ne)) pgpu_fatal("cell packing (host allocation)");
		dcell_t *dc = dalloc<dcell_t>(n);
		dedge_t *de = dalloc<dedge_t>(ne);
		dpnorm_t *dp = dalloc<dpnorm_t>(n);
		HIPCHK(hipMemcpy(dc, hc, (size_t)n * sizeof(dcell_t), hipMemcpyHostToDevice));
		if (ne > 0) HIPCHK(hipMemcpy(de, he, (size_t)ne * sizeof(dedge_t), hipMemcpyHostToDevice));
		HIPCHK(hipMemcpy(dp, hp, (size_t)n * sizeof(dpnorm_t), hipMemcpyHostToDevice));
		free(hc);
		free(he);
		free(hp);
		if (m == 0) {
			ps.c = dc;
			ps.e = de;
			ps.pn = dp;
		}
		else {
			ps.cm = dc;
			ps.em = de;
			ps.pnm = dp;
		}
	}
}

void free_set(pset &ps)
{
	dfree(ps.c);
	dfree(ps.cm);
	dfree(ps.e);
	dfree(ps.em);
	dfree(ps.pn);
	dfree(ps.pnm);
	ps = pset();
}

pe::fill_args base_args(const pset &ps, int n, double kw, int gp, void *out, long ld)
{
	pe::fill_args a;
	memset(&a, 0, sizeof(a));
	a.S = ps.c;
	a.pS.edge = ps.e;
	a.pS.pn = ps.pn;
	a.SM = ps.cm;
	a.pSM.edge = ps.em;
	a.pSM.pn = ps.pnm;
	a.n = n;
	a.nt = (n + 15) / 16;
	a.kw = kw;
	a.coef = MU0 / (4 * PI);
	a.gp = gp;
	a.out = out;
	a.ld = ld;
	a.k4pe = 4 * PI * EPS0;
	return a;
}

// capacitive cells need pot_cells() on the host first
void ensure_cells(pgpu *g)
{
	const peec_t *p = g->p;
	if ((g->csoff != nullptr) || (p->ncell <= 0) || (p->csoff == nullptr)) return;
	g->ncell = p->ncell;
	g->csoff = dalloc<int>(p->ncell + 1);
	g->csidx = dalloc<int>(p->nchg);
	g->carea = dalloc<double>(p->ncell);
	HIPCHK(hipMemcpy(g->csoff, p->csoff, (size_t)(p->ncell + 1) * sizeof(int), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(g->csidx, p->csidx, (size_t)p->nchg * sizeof(int), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(g->carea, p->carea, (size_t)p->ncell * sizeof(double), hipMemcpyHostToDevice));
	bool ident = (p->ncell == p->nchg);
	for (int h = 0; ident && (h < p->nchg); h++) ident = (p->cellof[h] == h);
	g->cells_identity = ident;
}

double ms_since(std::chrono::steady_clock::time_point t0)
{
	return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ------------------------------------------------------------------------------------------
// MNA assembly kernels

// dense part: -jw Lp on the branch block, jw C on the node block, 0 elsewhere
__global__ void k_assemble_dense(double2 *A, long lda, int n, int offS, int nseg, const double2 *lp, double omega,
                                 const int *cellrow, const double2 *C, int ncell)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * n) return;
	const int r = (int)(e / n), c = (int)(e - (long)r * n);
	double2 v = make_double2(0, 0);
	if ((r >= offS) && (r < offS + nseg) && (c >= offS) && (c < offS + nseg)) {
		const double2 x = lp[(long)(r - offS) * nseg + (c - offS)];
		// d_mul((0, -w), x) = (0 x.r - (-w) x.i, 0 x.i + (-w) x.r)
		v = make_double2((0 * x.x) - ((-omega) * x.y), (0 * x.y) + ((-omega) * x.x));
	}
	else if (cellrow != nullptr) {
		const int cr = cellrow[r], cc = cellrow[c];
		if ((cr >= 0) && (cc >= 0)) {
			const double2 x = C[(long)cr * ncell + cc];
			// d_mul((0, w), x)
			v = make_double2((0 * x.x) - (omega * x.y), (0 * x.y) + (omega * x.x));
		}
	}
	A[(long)r * lda + c] = make_double2(0 + v.x, 0 + v.y);
}

// sparse stamps (coalesced per entry in the reference's order): A += v
__global__ void k_assemble_sparse(double2 *A, long lda, const int *ri, const int *ci, const double2 *val, int nst)
{
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= nst) return;
	double2 *a = &A[(long)ri[e] * lda + ci[e]];
	const double2 x = *a, v = val[e];
	*a = make_double2(v.x + x.x, v.y + x.y);
}

__global__ void k_fill_identity(double2 *B, long ldb, int n)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * n) return;
	const int r = (int)(e / n), c = (int)(e - (long)r * n);
	B[(long)r * ldb + c] = make_double2((r == c) ? 1.0 : 0.0, 0);
}

// port right-hand sides: column j has +1 at node row i1(j), -1 at i2(j)
__global__ void k_port_rhs(double2 *B, int n, int nport, const int *i1, const int *i2)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)n * nport) return;
	const int r = (int)(e / nport), j = (int)(e - (long)r * nport);
	double x = 0;
	if (r == i1[j]) x = 0 + 1.0;
	if (r == i2[j]) x = x - 1.0;
	B[e] = make_double2(x, 0);
}

// q_m = sum_l C[m][l] v_l with v_l the node voltage of cell l (reference solve.c:352-358)
__global__ void k_cell_charge(const double2 *C, int ncell, const int *rowofcell, const double2 *B, int nport,
                              double2 *q)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (long)ncell * nport) return;
	const int m = (int)(e / nport), j = (int)(e - (long)m * nport);
	double qr = 0, qi = 0;
	for (int l = 0; l < ncell; l++) {
		const int im = rowofcell[l];
		const double2 v = (im < 0) ? make_double2(0, 0) : B[(long)im * nport + j];
		const double2 c = C[(long)m * ncell + l];
		qr = qr + ((c.x * v.x) - (c.y * v.y));
		qi = qi + ((c.x * v.y) + (c.y * v.x));
	}
	q[e] = make_double2(qr, qi);
}

// ---- precision form: FP32-stored factors + FP64 iterative refinement

__global__ void k_demote64(const double2 *src, float2 *dst, long cnt)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e < cnt) dst[e] = make_float2((float)src[e].x, (float)src[e].y);
}

// R[:, j] = B[:, j] - A X[:, j] in FP64 (one wave per row), and the componentwise (Oettli-Prager)
// backward error of each column, max_i |r_i| / (|A| |x| + |b|)_i, folded into cb[j] with an
// integer max (non-negative doubles order as their bit patterns)
__global__ __launch_bounds__(256) void k_resid(const double2 *A, int n, const double2 *X, const double2 *B, double2 *R, int nrhs,
                                               unsigned long long *cb)
{
	const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
	if (row >= n) return;
	const double2 *a = A + (long)row * n;
	for (int j = 0; j < nrhs; j++) {
		double sr = 0, si = 0, sa = 0;
		for (int c = lane; c < n; c += 64) {
			const double2 u = a[c], x = X[(long)c * nrhs + j];
			sr += (u.x * x.x) - (u.y * x.y);
			si += (u.x * x.y) + (u.y * x.x);
			sa += sqrt((u.x * u.x) + (u.y * u.y)) * sqrt((x.x * x.x) + (x.y * x.y));
		}
		for (int off = 32; off > 0; off >>= 1) {
			sr += __shfl_xor(sr, off);
			si += __shfl_xor(si, off);
			sa += __shfl_xor(sa, off);
		}
		if (lane == 0) {
			const double2 b = B[(long)row * nrhs + j];
			const double rr = b.x - sr, ri = b.y - si;
			R[(long)row * nrhs + j] = make_double2(rr, ri);
			const double den = sa + sqrt((b.x * b.x) + (b.y * b.y));
			const double num = sqrt((rr * rr) + (ri * ri));
			const double be = (den > 0) ? (num / den) : ((num > 0) ? 1.0 : 0.0);
			atomicMax(&cb[j], (unsigned long long)__double_as_longlong(be));
		}
	}
}

__global__ void k_zero_u64(unsigned long long *p, int cnt)
{
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e < cnt) p[e] = 0ull;
}

__global__ void k_addto(double2 *X, const double2 *D, long cnt)
{
	const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
	if (e < cnt) X[e] = make_double2(X[e].x + D[e].x, X[e].y + D[e].y);
}

void ensure_mna(pgpu *g, const peec_t *p)
{
	if (g->A != nullptr) return;
	const int n = p->nunknown;
	g->n = n;
	g->A = dalloc<double2>((size_t)n * n);
	g->apiv = dalloc<int>(n);
	g->aperm = dalloc<int>(n);
	const int np = (p->nport > 0) ? p->nport : 1;
	g->B = dalloc<double2>((size_t)n * np);
	g->Bw = dalloc<double2>((size_t)n * np);
	if (p->capacitance && (p->ncell > 0) && (p->offQ >= p->nunknown)) {
		std::vector<int> cr(n, -1);
		for (int i = 0; i < p->ncell; i++) {
			const int m = p->nodemap[p->cellid[i]];
			if (m >= 0) cr[m] = i;
		}
		g->cellrow = dalloc<int>(n);
		HIPCHK(hipMemcpy(g->cellrow, cr.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
		g->qbuf = dalloc<double2>((size_t)p->ncell * np);
		std::vector<int> rc(p->ncell);
		for (int l = 0; l < p->ncell; l++) rc[l] = p->nodemap[p->cellid[l]];
		g->rowofcell = dalloc<int>(p->ncell);
		HIPCHK(hipMemcpy(g->rowofcell, rc.data(), (size_t)p->ncell * sizeof(int), hipMemcpyHostToDevice));
	}
}

}  // namespace

// fill arguments of one cell set (0: current cells for Lp, 1: charge cells) for other modules
pe::fill_args pgpu_set_args(pgpu_t *g, int which, double kw, int *km, const int **csoff, const int **csidx)
{
	const pset &ps = (which == 0) ? g->seg : g->chg;
	if (which == 1) ensure_cells(g);
	pe::fill_args a = base_args(ps, ps.n, kw, g->p->gp, nullptr, 0);
	if (which == 1) a.carea = g->carea;
	if (km) *km = ps.km;
	if (csoff) *csoff = g->csoff;
	if (csidx) *csidx = g->csidx;
	return a;
}

hipStream_t pgpu_hip_stream(pgpu_t *g)
{
	return g->s;
}

// ------------------------------------------------------------------------------------------
extern "C" {

void *pgpu_stream(pgpu_t *g) { return (void *)g->s; }
const double2 *pgpu_lp_dev(pgpu_t *g) { return g->lp; }
const double2 *pgpu_p_dev(pgpu_t *g) { return g->P; }

void pgpu_fatal(const char *what)
{
	fprintf(stderr, "*** libpeec_mi355x: %s\n", what);
	fflush(stderr);
	abort();
}

int pgpu_device_ok(void)
{
	int ndev = 0;
	return (hipGetDeviceCount(&ndev) == hipSuccess) && (ndev > 0);
}

pgpu_t *pgpu_get(const peec_t *p)
{
	std::lock_guard<std::mutex> lk(g_mu);
	for (pgpu *g : g_tab)
		if (g->p == p) {
			ensure_cells(g);
			return g;
		}
	init_device();
	pgpu *g = new pgpu();
	g->p = p;
	HIPCHK(hipStreamCreateWithFlags(&g->s, hipStreamNonBlocking));
	upload_set(g->seg, p->seg, p->nseg, p->gp, p->gpz);
	upload_set(g->chg, p->chg, p->nchg, p->gp, p->gpz);
	ensure_cells(g);
	g_tab.push_back(g);
	return g;
}

namespace {
void release_mirror(pgpu *g)
{
		HIPCHK(hipStreamSynchronize(g->s));
		free_set(g->seg);
		free_set(g->chg);
		dfree(g->csoff);
		dfree(g->csidx);
		dfree(g->carea);
		dfree(g->lp);
		dfree(g->kf_rec_lp);
		dfree(g->kf_rec_ig);
		dfree(g->kf_list_lp);
		dfree(g->kf_list_ig);
		dfree(g->kf_cnt);
		dfree(g->kf_rho_seg);
		dfree(g->kf_rho_chg);
		dfree(g->G);
		dfree(g->P);
		dfree(g->C);
		dfree(g->ppiv);
		dfree(g->pperm);
		dfree(g->cwork);
		dfree(g->A);
		dfree(g->apiv);
		dfree(g->aperm);
		dfree(g->B);
		dfree(g->Bw);
		dfree(g->A32);
		dfree(g->dinv32);
		dfree(g->Brhs);
		dfree(g->R);
		dfree(g->cberr);
		dfree(g->cellrow);
		dfree(g->d_st_ri);
		dfree(g->d_st_ci);
		dfree(g->d_st_v);
		dfree(g->qbuf);
		dfree(g->xr);
		dfree(g->rowofcell);
		dfree(g->alu);
		dfree(g->alu_piv);
		dfree(g->alu_perm);
		dgmres_ws_free(g->gw);
		HIPCHK(hipStreamDestroy(g->s));
		delete g;
}
}  // namespace

void pgpu_at_exit(void (*fn)(void))
{
	std::lock_guard<std::mutex> lk(g_mu);
	register_teardown();
	for (int k = 0; k < g_nexit_hooks; k++)
		if (g_exit_hooks[k] == fn) return;
	if (g_nexit_hooks == EXIT_HOOKS_MAX) pgpu_fatal("too many exit hooks");
	g_exit_hooks[g_nexit_hooks++] = fn;
}

void peec_gpu_release(peec_t *p)
{
	std::lock_guard<std::mutex> lk(g_mu);
	if (g_exiting) return;   // the exit handler has released every mirror
	for (size_t k = 0; k < g_tab.size(); k++) {
		pgpu *g = g_tab[k];
		if (g->p != p) continue;
		release_mirror(g);
		g_tab.erase(g_tab.begin() + (long)k);
		return;
	}
}

int pgpu_fill_lp(pgpu_t *g, double kw)
{
	const int n = g->seg.n;
	if (n <= 0) return 0;
	if (g->lp == nullptr) g->lp = dalloc<double2>((size_t)n * n);
	pe::fill_args a = base_args(g->seg, n, kw, g->p->gp, g->lp, n);
	fill_launch(g->s, 0, g->seg.km, a);
	return 0;
}

int pgpu_fill_pot(pgpu_t *g, double kw)
{
	const peec_t *p = g->p;
	const int nh = g->chg.n, nc = p->ncell;
	if ((nh <= 0) || (nc <= 0)) return 0;
	ensure_cells(g);
	if (g->P == nullptr) g->P = dalloc<double2>((size_t)nc * nc);
	if (g->cells_identity) {
		pe::fill_args a = base_args(g->chg, nh, kw, p->gp, g->P, nc);
		a.carea = g->carea;
		fill_launch(g->s, 1, g->chg.km, a);
	}
	else {
		if (g->G == nullptr) g->G = dalloc<double2>((size_t)nh * nh);
		pe::fill_args a = base_args(g->chg, nh, kw, p->gp, g->G, nh);
		fill_launch(g->s, 1, g->chg.km, a);
		aggregate_launch(g->s, g->G, nh, g->csoff, g->csidx, g->carea, nc, 4 * PI * EPS0, g->P, nc);
	}
	return 0;
}

// Prepare the k-factored fill for wavenumbers in [klo, khi] (reused while a later band lies
// inside it).  Returns 0 when ready, 1 when the records would not fit the memory budget (the
// caller then fills directly).
// launches the record setup on the mirror's stream; pgpu_kf_setup_finish reads the direct-pair
// counts (so that many structures' setups can run concurrently)
int pgpu_kf_setup_launch(pgpu_t *g, double klo, double khi)
{
	const peec_t *p = g->p;
	if (g->kf_ready && (klo >= g->kf_lo) && (khi <= g->kf_hi)) return 0;
	const int ns = g->seg.n, nh = g->chg.n;
	const int T = p->gp ? 2 : 1;
	const long sl = kf_slots(ns), sh = (nh > 0) ? kf_slots(nh) : 0;
	const double bytes = (double)T * kf_record_doubles() * 8.0 * (double)(sl + sh) + 8.0 * (double)(sl + sh);
	// records of an earlier band go first; the new ones may take up to 70 % of the free device
	// memory (the sweep's systems and factors are allocated before them; the rest is for the
	// plain and precision forms' buffers); beyond that the caller fills element by element
	dfree(g->kf_rec_lp);
	dfree(g->kf_rec_ig);
	dfree(g->kf_list_lp);
	dfree(g->kf_list_ig);
	g->kf_rec_lp = g->kf_rec_ig = nullptr;
	g->kf_list_lp = g->kf_list_ig = nullptr;
	g->kf_ready = false;
	size_t mfree = 0, mtotal = 0;
	HIPCHK(hipMemGetInfo(&mfree, &mtotal));
	if (bytes > 0.7 * (double)mfree) return 1;
	if (g->kf_cnt == nullptr) g->kf_cnt = dalloc<unsigned>(2);
	if (g->kf_rho_seg == nullptr) g->kf_rho_seg = dalloc<double>(ns);
	HIPCHK(hipMemsetAsync(g->kf_cnt, 0, 2 * sizeof(unsigned), g->s));
	g->kf_T = T;
	if (ns > 0) {
		g->kf_rec_lp = dalloc<double>((size_t)T * kf_record_doubles() * sl);
		g->kf_list_lp = dalloc<int2>(sl);
		kf_rho_launch(g->s, g->seg.c, g->seg.e, ns, g->kf_rho_seg);
		pe::fill_args a = base_args(g->seg, ns, khi, p->gp, nullptr, ns);
		kf_setup_launch(g->s, 0, g->seg.km, a, klo, khi, g->kf_rho_seg, g->kf_rec_lp, T, g->kf_list_lp, g->kf_cnt);
	}
	if (nh > 0 && p->ncell > 0) {
		ensure_cells(g);
		if (g->kf_rho_chg == nullptr) g->kf_rho_chg = dalloc<double>(nh);
		g->kf_rec_ig = dalloc<double>((size_t)T * kf_record_doubles() * sh);
		g->kf_list_ig = dalloc<int2>(sh);
		kf_rho_launch(g->s, g->chg.c, g->chg.e, nh, g->kf_rho_chg);
		pe::fill_args a = base_args(g->chg, nh, khi, p->gp, nullptr, nh);
		kf_setup_launch(g->s, 1, g->chg.km, a, klo, khi, g->kf_rho_chg, g->kf_rec_ig, T, g->kf_list_ig, g->kf_cnt + 1);
	}
	g->kf_lo = klo;
	g->kf_hi = khi;
	g->kf_pending = true;
	return 0;
}

void pgpu_kf_setup_finish(pgpu_t *g)
{
	if (!g->kf_pending) return;
	unsigned cnt[2] = {0, 0};
	HIPCHK(hipMemcpyAsync(cnt, g->kf_cnt, sizeof(cnt), hipMemcpyDeviceToHost, g->s));
	HIPCHK(hipStreamSynchronize(g->s));
	g->kf_nd_lp = (int)cnt[0];
	g->kf_nd_ig = (int)cnt[1];
	g->kf_pending = false;
	g->kf_ready = true;
}

int pgpu_kf_setup(pgpu_t *g, double klo, double khi)
{
	if (pgpu_kf_setup_launch(g, klo, khi)) return 1;
	pgpu_kf_setup_finish(g);
	return 0;
}

// records of one cell set (0 current cells, 1 charge sub-cells) when a band covering kw is set up
const double *pgpu_kf_records(pgpu_t *g, int which, double kw, long *nslots, int *T)
{
	if (!g->kf_ready || (kw < g->kf_lo * (1 - 1e-12)) || (kw > g->kf_hi * (1 + 1e-12))) return nullptr;
	const double *r = (which == 0) ? g->kf_rec_lp : g->kf_rec_ig;
	if (r == nullptr) return nullptr;
	*nslots = kf_slots((which == 0) ? g->seg.n : g->chg.n);
	*T = g->kf_T;
	return r;
}

void pgpu_batch_view(pgpu_t *g, pgpu_bview *v)
{
	const peec_t *p = g->p;
	v->s = g->s;
	v->seg = pgpu_set_args(g, 0, 0, &v->km_seg, nullptr, nullptr);
	v->nseg = g->seg.n;
	v->nh = g->chg.n;
	v->ncell = (p->capacitance && (g->chg.n > 0)) ? p->ncell : 0;
	if (v->ncell > 0) v->chg = pgpu_set_args(g, 1, 0, &v->km_chg, nullptr, nullptr);
	v->ident = g->cells_identity;
	v->rec_lp = g->kf_rec_lp;
	v->rec_ig = g->kf_rec_ig;
	v->T = g->kf_T;
	v->nd_lp = g->kf_nd_lp;
	v->nd_ig = g->kf_nd_ig;
	v->list_lp = g->kf_list_lp;
	v->list_ig = g->kf_list_ig;
	v->csoff = g->csoff;
	v->csidx = g->csidx;
	v->carea = g->carea;
}

// Lp (nseg x nseg) and P (ncell x ncell) at kw from the records into caller buffers
void pgpu_kf_fill(pgpu_t *g, double kw, d_complex_t *dlp, d_complex_t *dP)
{
	double2 *lp_out = (double2 *)dlp, *p_out = (double2 *)dP;
	const peec_t *p = g->p;
	const int ns = g->seg.n, nh = g->chg.n, nc = p->ncell;
	if ((ns > 0) && lp_out) {
		pe::fill_args a = base_args(g->seg, ns, kw, p->gp, lp_out, ns);
		kf_eval_launch(g->s, 0, a, g->kf_rec_lp, g->kf_T);
		fill_list_launch(g->s, 0, g->seg.km, a, g->kf_list_lp, g->kf_nd_lp);
	}
	if ((nh > 0) && (nc > 0) && p_out) {
		if (g->cells_identity) {
			pe::fill_args a = base_args(g->chg, nh, kw, p->gp, p_out, nc);
			a.carea = g->carea;
			kf_eval_launch(g->s, 1, a, g->kf_rec_ig, g->kf_T);
			fill_list_launch(g->s, 1, g->chg.km, a, g->kf_list_ig, g->kf_nd_ig);
		}
		else {
			if (g->G == nullptr) g->G = dalloc<double2>((size_t)nh * nh);
			pe::fill_args a = base_args(g->chg, nh, kw, p->gp, g->G, nh);
			kf_eval_launch(g->s, 1, a, g->kf_rec_ig, g->kf_T);
			fill_list_launch(g->s, 1, g->chg.km, a, g->kf_list_ig, g->kf_nd_ig);
			aggregate_launch(g->s, g->G, nh, g->csoff, g->csidx, g->carea, nc, 4 * PI * EPS0, p_out, nc);
		}
	}
}

// Lp (which 0) or P (which 1) at kw from the records into out (original numbering, leading
// dimension the cell count); 1 when the records do not cover kw or the sub-cell matrix that P
// is aggregated from would not fit in a quarter of free memory (nothing launched then)
int pgpu_kf_fill_kind(pgpu_t *g, int which, double kw, d_complex_t *out)
{
	long ns = 0;
	int T = 0;
	if (pgpu_kf_records(g, which, kw, &ns, &T) == nullptr) return 1;
	if ((which == 1) && !g->cells_identity && (g->G == nullptr)) {
		size_t fr = 0, tot = 0;
		HIPCHK(hipMemGetInfo(&fr, &tot));
		if ((double)g->chg.n * g->chg.n * sizeof(double2) > 0.25 * (double)fr) return 1;
	}
	pgpu_kf_fill(g, kw, (which == 0) ? out : nullptr, (which == 1) ? out : nullptr);
	return 0;
}

// the same for nk wavenumbers (a batch of frequencies): the records are read once for all of them.
// With upper set, the lower triangles outside the diagonal 64 x 64 blocks are left as they were
// wherever the records serve (the sub-cell path for P still writes whole matrices).
void pgpu_kf_fill_multi(pgpu_t *g, int nk, const double *kw, d_complex_t *const *dlp, d_complex_t *const *dP, int upper)
{
	const peec_t *p = g->p;
	const int ns = g->seg.n, nh = g->chg.n, nc = p->ncell;
	if (nk > KF_FMAX) {
		for (int k0 = 0; k0 < nk; k0 += KF_FMAX)
			pgpu_kf_fill_multi(g, std::min(KF_FMAX, nk - k0), kw + k0, dlp + k0, dP ? dP + k0 : nullptr, upper);
		return;
	}
	if (ns > 0) {
		pe::fill_args a = base_args(g->seg, ns, kw[0], p->gp, dlp[0], ns);
		kf_eval_multi_launch(g->s, 0, a, g->kf_rec_lp, g->kf_T, nk, kw, (void *const *)dlp, upper != 0);
		for (int k = 0; k < nk; k++) {
			a.kw = kw[k];
			a.out = dlp[k];
			fill_list_launch(g->s, 0, g->seg.km, a, g->kf_list_lp, g->kf_nd_lp);
		}
	}
	if ((nh > 0) && (nc > 0) && dP) {
		if (g->cells_identity) {
			pe::fill_args a = base_args(g->chg, nh, kw[0], p->gp, dP[0], nc);
			a.carea = g->carea;
			kf_eval_multi_launch(g->s, 1, a, g->kf_rec_ig, g->kf_T, nk, kw, (void *const *)dP, upper != 0);
			for (int k = 0; k < nk; k++) {
				a.kw = kw[k];
				a.out = dP[k];
				fill_list_launch(g->s, 1, g->chg.km, a, g->kf_list_ig, g->kf_nd_ig);
			}
		}
		else {
			// sub-cell integrals aggregated per frequency through the one G buffer
			for (int k = 0; k < nk; k++) pgpu_kf_fill(g, kw[k], nullptr, dP[k]);
		}
	}
}

void pgpu_kf_reset_all(void)
{
	std::lock_guard<std::mutex> lk(g_mu);
	for (pgpu *g : g_tab) g->kf_ready = false;
}

int pgpu_kf_direct_pairs(pgpu_t *g)
{
	return g->kf_nd_lp + g->kf_nd_ig;
}

void pgpu_fill_to_host(pgpu_t *g, double kw, int mode, d_complex_t *lp, d_complex_t *P)
{
	const peec_t *p = g->p;
	const int ns = g->seg.n, nc = p->ncell;
	const bool cap = p->capacitance && (g->chg.n > 0) && (nc > 0);
	if (ns > 0 && g->lp == nullptr) g->lp = dalloc<double2>((size_t)ns * ns);
	if (cap && g->P == nullptr) g->P = dalloc<double2>((size_t)nc * nc);
	if (mode == 1) {
		pgpu_kf_fill(g, kw, (d_complex_t *)g->lp, cap ? (d_complex_t *)g->P : nullptr);
	}
	else {
		pgpu_fill_lp(g, kw);
		if (cap) pgpu_fill_pot(g, kw);
	}
	if (ns > 0) HIPCHK(hipMemcpyAsync(lp, g->lp, (size_t)ns * ns * sizeof(double2), hipMemcpyDeviceToHost, g->s));
	if (cap && P) HIPCHK(hipMemcpyAsync(P, g->P, (size_t)nc * nc * sizeof(double2), hipMemcpyDeviceToHost, g->s));
	HIPCHK(hipStreamSynchronize(g->s));
}

// element-by-element Lp and P at kw into the mirror's own device buffers (synchronised);
// *lp / *P receive them (P null without capacitive cells)
void pgpu_fill_direct_dev(pgpu_t *g, double kw, const void **lp, const void **P)
{
	const peec_t *p = g->p;
	const int ns = g->seg.n, nc = p->ncell;
	const bool cap = p->capacitance && (g->chg.n > 0) && (nc > 0);
	if (ns > 0 && g->lp == nullptr) g->lp = dalloc<double2>((size_t)ns * ns);
	if (cap && g->P == nullptr) g->P = dalloc<double2>((size_t)nc * nc);
	pgpu_fill_lp(g, kw);
	if (cap) pgpu_fill_pot(g, kw);
	HIPCHK(hipStreamSynchronize(g->s));
	*lp = g->lp;
	*P = cap ? g->P : nullptr;
}

int pgpu_cmat(pgpu_t *g, int *ising)
{
	const int nc = g->p->ncell;
	*ising = -1;
	if (nc <= 0) return 0;
	if (g->C == nullptr) {
		g->C = dalloc<double2>((size_t)nc * nc);
		g->ppiv = dalloc<int>(nc);
		g->pperm = dalloc<int>(nc);
		g->cwork = dalloc<double2>((size_t)nc * nc);
	}
	const int info = dlu_factor(g->s, nc, (d_complex_t *)g->P, nc, g->ppiv, g->pperm, nullptr);
	if (info >= 0) {
		*ising = info;
		return 1;
	}
	const long tot = (long)nc * nc;
	hipLaunchKernelGGL(k_fill_identity, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, g->s, g->C, (long)nc, nc);
	dlu_solve_perm(g->s, nc, (const d_complex_t *)g->P, nc, g->pperm, (d_complex_t *)g->C, nc, nc,
	               (d_complex_t *)g->cwork);
	return 0;
}

void pgpu_lp_to_host(pgpu_t *g, d_complex_t *lp)
{
	const size_t n = (size_t)g->seg.n;
	HIPCHK(hipMemcpyAsync(lp, g->lp, n * n * sizeof(double2), hipMemcpyDeviceToHost, g->s));
	HIPCHK(hipStreamSynchronize(g->s));
}

void pgpu_cmat_to_host(pgpu_t *g, d_complex_t *cmat)
{
	const size_t n = (size_t)g->p->ncell;
	HIPCHK(hipMemcpyAsync(cmat, g->C, n * n * sizeof(double2), hipMemcpyDeviceToHost, g->s));
	HIPCHK(hipStreamSynchronize(g->s));
}

void pgpu_lp_from_host(pgpu_t *g, const d_complex_t *lp)
{
	const size_t n = (size_t)g->seg.n;
	if (n == 0) return;
	if (g->lp == nullptr) g->lp = dalloc<double2>(n * n);
	HIPCHK(hipMemcpyAsync(g->lp, lp, n * n * sizeof(double2), hipMemcpyHostToDevice, g->s));
	HIPCHK(hipStreamSynchronize(g->s));
}

void pgpu_cmat_from_host(pgpu_t *g, const d_complex_t *cmat)
{
	const size_t n = (size_t)g->p->ncell;
	if (n == 0) return;
	if (g->C == nullptr) {
		g->C = dalloc<double2>(n * n);
		g->ppiv = dalloc<int>(n);
		g->pperm = dalloc<int>(n);
		g->cwork = dalloc<double2>(n * n);
	}
	HIPCHK(hipMemcpyAsync(g->C, cmat, n * n * sizeof(double2), hipMemcpyHostToDevice, g->s));
	HIPCHK(hipStreamSynchronize(g->s));
}

d_complex_t pgpu_lp_entry(pgpu_t *g, int i, int j, double kw)
{
	if (i > j) {
		const int t = i;
		i = j;
		j = t;
	}
	double2 *res = dalloc<double2>(1);
	pe::fill_args a = base_args(g->seg, g->seg.n, kw, g->p->gp, nullptr, 0);
	fill_one(g->s, 0, g->seg.km, a, i, j, nullptr, nullptr, res);
	double2 h;
	HIPCHK(hipMemcpyAsync(&h, res, sizeof(h), hipMemcpyDeviceToHost, g->s));
	HIPCHK(hipStreamSynchronize(g->s));
	dfree(res);
	return d_complex(h.x, h.y);
}

d_complex_t pgpu_pot_entry(pgpu_t *g, int i, int j, double kw)
{
	ensure_cells(g);
	double2 *res = dalloc<double2>(1);
	pe::fill_args a = base_args(g->chg, g->chg.n, kw, g->p->gp, nullptr, 0);
	a.carea = g->carea;
	fill_one(g->s, 1, g->chg.km, a, i, j, g->csoff, g->csidx, res);
	double2 h;
	HIPCHK(hipMemcpyAsync(&h, res, sizeof(h), hipMemcpyDeviceToHost, g->s));
	HIPCHK(hipStreamSynchronize(g->s));
	dfree(res);
	return d_complex(h.x, h.y);
}

// ------------------------------------------------------------------------------------------
namespace {
// The MNA matrix of one frequency point in g->A (solve.c:239-299 up to the factorisation):
// Lp and, with capacitive cells, P -> C = P^-1 when retarded (the static setup's matrices
// otherwise), then the dense blocks and the sparse stamps.  ev (may be NULL) receives events
// 0..4 around the stages.  Returns 1 on a singular P.
int point_system(pgpu *g, const peec_t *p, double freq, bool usecap, hipEvent_t *ev)
{
	const int n = p->nunknown;
	const double kw = p->retardation ? (2 * PI * freq / C0) : 0;
	const double omega = 2 * PI * freq;
	auto mark = [&](int k) {
		if (ev) HIPCHK(hipEventRecord(ev[k], g->s));
	};
	mark(0);
	if (p->retardation) {
		pgpu_fill_lp(g, kw);
		mark(1);
		if (usecap) {
			pgpu_fill_pot(g, kw);
			mark(2);
			int ising = -1;
			if (pgpu_cmat(g, &ising)) {
				printf("*** singular potential coefficient matrix (row %d)\n", ising);
				return 1;
			}
		}
		else {
			mark(2);
		}
	}
	else {
		mark(1);
		mark(2);
	}
	mark(3);
	stamps_t sp;
	mna_stamps(p, freq, &sp);
	if (sp.n > g->st_cap) {
		dfree(g->d_st_ri);
		dfree(g->d_st_ci);
		dfree(g->d_st_v);
		g->st_cap = sp.n;
		g->d_st_ri = dalloc<int>(sp.n);
		g->d_st_ci = dalloc<int>(sp.n);
		g->d_st_v = dalloc<double2>(sp.n);
	}
	if (sp.n > 0) {
		HIPCHK(hipMemcpyAsync(g->d_st_ri, sp.ri, (size_t)sp.n * sizeof(int), hipMemcpyHostToDevice, g->s));
		HIPCHK(hipMemcpyAsync(g->d_st_ci, sp.ci, (size_t)sp.n * sizeof(int), hipMemcpyHostToDevice, g->s));
		HIPCHK(hipMemcpyAsync(g->d_st_v, sp.val, (size_t)sp.n * sizeof(double2), hipMemcpyHostToDevice, g->s));
	}
	const long tot = (long)n * n;
	hipLaunchKernelGGL(k_assemble_dense, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, g->s, g->A, (long)n, n,
	                   p->offS, p->nseg, g->lp, omega, usecap ? g->cellrow : nullptr, g->C, p->ncell);
	if (sp.n > 0)
		hipLaunchKernelGGL(k_assemble_sparse, dim3((sp.n + 255) / 256), dim3(256), 0, g->s, g->A, (long)n, g->d_st_ri,
		                   g->d_st_ci, g->d_st_v, sp.n);
	mark(4);
	HIPCHK(hipStreamSynchronize(g->s));   // the stamps' host buffers are released next
	mna_stamps_free(&sp);
	return 0;
}

}  // namespace

int pgpu_sweep(pgpu_t *g, const peec_t *p, int ifreq0, int nfreq, d_complex_t *zmat, d_complex_t *segi,
               d_complex_t *cellq, pgpu_stats_t *st, int precision)
{
	const int n = p->nunknown, np = p->nport;
	ensure_cells(g);
	ensure_mna(g, p);
	const bool usecap = p->capacitance && (p->ncell > 0) && (g->cellrow != nullptr);
	std::vector<int> i1(np), i2(np);
	for (int j = 0; j < np; j++) {
		i1[j] = p->nodemap[p->port[j].n1];
		i2[j] = p->nodemap[p->port[j].n2];
	}
	int *d_i1 = dalloc<int>(np), *d_i2 = dalloc<int>(np);
	HIPCHK(hipMemcpy(d_i1, i1.data(), (size_t)np * sizeof(int), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(d_i2, i2.data(), (size_t)np * sizeof(int), hipMemcpyHostToDevice));
	int *d_rowofcell = nullptr;
	if (usecap) {
		std::vector<int> rc(p->ncell);
		for (int l = 0; l < p->ncell; l++) rc[l] = p->nodemap[p->cellid[l]];
		d_rowofcell = dalloc<int>(p->ncell);
		HIPCHK(hipMemcpy(d_rowofcell, rc.data(), (size_t)p->ncell * sizeof(int), hipMemcpyHostToDevice));
	}
	std::vector<double2> hB((size_t)n * np);
	std::vector<double2> hq(usecap ? (size_t)p->ncell * np : 1);
	hipEvent_t ev[8];
	for (int k = 0; k < 8; k++) HIPCHK(hipEventCreate(&ev[k]));
	int rc = 0;
	// static sweeps reuse the setup fill (host p->lp / p->cmat) uploaded once
	if (!p->retardation) {
		if (p->nseg > 0) pgpu_lp_from_host(g, p->lp);
		if (usecap) pgpu_cmat_from_host(g, p->cmat);
	}
	for (int f = 0; f < nfreq; f++) {
		const int ifreq = ifreq0 + f;
		const double freq = freq_at(p, ifreq);
		const auto t0 = std::chrono::steady_clock::now();
		if (point_system(g, p, freq, usecap, ev)) {
			rc = 1;
			break;
		}
		// factor: in place in FP64, or an FP32 copy (the FP64 matrix is kept for the refinement)
		int info;
		bool fp64 = !precision;
		if (precision) {
			if (g->A32 == nullptr) {
				g->A32 = dalloc<float2>((size_t)n * n);
				g->dinv32 = dalloc<double2>((size_t)2 * ((n + 63) / 64) * 64 * 64);
				const int npa = (np > 0) ? np : 1;
				g->Brhs = dalloc<double2>((size_t)n * npa);
				g->R = dalloc<double2>((size_t)n * npa);
				g->cberr = dalloc<unsigned long long>(npa);
			}
			hipLaunchKernelGGL(k_demote64, dim3((unsigned)(((long)n * n + 255) / 256)), dim3(256), 0, g->s, (const double2 *)g->A,
			                   g->A32, (long)n * n);
			info = slu_factor(g->s, n, g->A32, n, g->apiv, g->aperm, st);
			if (info >= 0) {
				// a pivot that rounds to zero in FP32 storage: this point is solved in FP64
				if (st) st->fp64_fallbacks++;
				info = dlu_factor(g->s, n, (d_complex_t *)g->A, n, g->apiv, g->aperm, st);
				fp64 = true;
			}
		}
		else {
			info = dlu_factor(g->s, n, (d_complex_t *)g->A, n, g->apiv, g->aperm, st);
		}
		HIPCHK(hipEventRecord(ev[5], g->s));
		if (info >= 0) {
			printf("*** singular matrix at f = %.5e Hz (row %d : floating node or ideal voltage source loop?)\n", freq, info);
			if (st) st->ising = info;
			rc = 2;
			break;
		}
		// all ports at once
		const long tb = (long)n * np;
		hipLaunchKernelGGL(k_port_rhs, dim3((unsigned)((tb + 255) / 256)), dim3(256), 0, g->s, g->B, n, np, d_i1, d_i2);
		if (fp64) {
			dlu_solve_perm(g->s, n, (const d_complex_t *)g->A, n, g->aperm, (d_complex_t *)g->B, np, np,
			               (d_complex_t *)g->Bw);
		}
		else {
			// x = LU32^-1 b, then r = b - A x in FP64 and x += LU32^-1 r until the componentwise
			// backward error max_i |r_i| / (|A| |x| + |b|)_i is at most RF_STOP, or stops falling
			// by half per sweep, or reaches 2 eps64 (LAPACK zgerfs's rule; at most RF_MAX sweeps);
			// a point whose error stays above RF_TARGET (the form's guarantee) is factored again
			// in FP64.  One sweep usually takes the FP32 solve's 1e-7 to about 1e-14.
			constexpr int RF_MAX = 10;
			constexpr double RF_TARGET = 1e-10, RF_STOP = 1e-12;
			// the solves with the factors: each 64-row step through its diagonal block's inverse
			// (up to four ports at once), else by substitution
			d_complex_t *dL = (d_complex_t *)g->dinv32, *dU = dL + (size_t)((n + 63) / 64) * 64 * 64;
			if (np <= 4) slu_diag_inverse(g->s, n, g->A32, n, dL, dU);
			auto solve32 = [&](double2 *X) {
				if (np <= 4) slu_solve_perm_di(g->s, n, g->A32, n, g->aperm, (d_complex_t *)X, np, np, (d_complex_t *)g->Bw, dL, dU);
				else slu_solve_perm(g->s, n, g->A32, n, g->aperm, (d_complex_t *)X, np, np, (d_complex_t *)g->Bw);
			};
			HIPCHK(hipMemcpyAsync(g->Brhs, g->B, (size_t)n * np * sizeof(double2), hipMemcpyDeviceToDevice, g->s));
			solve32(g->B);
			std::vector<double> h(np);
			double berr = 1, last = 2;
			int its = 0;
			for (;;) {
				hipLaunchKernelGGL(k_zero_u64, dim3((np + 63) / 64), dim3(64), 0, g->s, g->cberr, np);
				hipLaunchKernelGGL(k_resid, dim3((n + 3) / 4), dim3(256), 0, g->s, (const double2 *)g->A, n, (const double2 *)g->B,
				                   (const double2 *)g->Brhs, g->R, np, g->cberr);
				HIPCHK(hipMemcpyAsync(h.data(), g->cberr, (size_t)np * sizeof(double), hipMemcpyDeviceToHost, g->s));
				HIPCHK(hipStreamSynchronize(g->s));
				berr = 0;
				for (int j = 0; j < np; j++)
					if (h[j] > berr) berr = h[j];
				if ((berr <= RF_STOP) || (berr <= 4.5e-16) || (berr > 0.5 * last) || (its == RF_MAX)) break;
				last = berr;
				solve32(g->R);
				hipLaunchKernelGGL(k_addto, dim3((unsigned)(((long)n * np + 255) / 256)), dim3(256), 0, g->s, g->B,
				                   (const double2 *)g->R, (long)n * np);
				its++;
			}
			if (berr > RF_TARGET) {
				if (st) st->fp64_fallbacks++;
				HIPCHK(hipMemcpyAsync(g->B, g->Brhs, (size_t)n * np * sizeof(double2), hipMemcpyDeviceToDevice, g->s));
				if (dlu_factor(g->s, n, (d_complex_t *)g->A, n, g->apiv, g->aperm, st) >= 0) {
					rc = 2;
					break;
				}
				dlu_solve_perm(g->s, n, (const d_complex_t *)g->A, n, g->aperm, (d_complex_t *)g->B, np, np,
				               (d_complex_t *)g->Bw);
				berr = 0;
			}
			if (st) {
				st->refine_its += its;
				if (berr > st->berr) st->berr = berr;
			}
		}
		if (usecap && (cellq != nullptr)) {
			const long tq = (long)p->ncell * np;
			hipLaunchKernelGGL(k_cell_charge, dim3((unsigned)((tq + 255) / 256)), dim3(256), 0, g->s, g->C, p->ncell,
			                   d_rowofcell, g->B, np, g->qbuf);
		}
		HIPCHK(hipEventRecord(ev[6], g->s));
		HIPCHK(hipMemcpyAsync(hB.data(), g->B, (size_t)n * np * sizeof(double2), hipMemcpyDeviceToHost, g->s));
		if (usecap && (cellq != nullptr))
			HIPCHK(hipMemcpyAsync(hq.data(), g->qbuf, (size_t)p->ncell * np * sizeof(double2), hipMemcpyDeviceToHost, g->s));
		HIPCHK(hipStreamSynchronize(g->s));
		for (int j = 0; j < np; j++) {
			for (int i = 0; i < np; i++) {
				const double2 v1 = (i1[i] < 0) ? make_double2(0, 0) : hB[(size_t)i1[i] * np + j];
				const double2 v2 = (i2[i] < 0) ? make_double2(0, 0) : hB[(size_t)i2[i] * np + j];
				zmat[((size_t)f * np + i) * np + j] = d_complex(v1.x - v2.x, v1.y - v2.y);
			}
			if (segi != nullptr)
				for (int m = 0; m < p->nseg; m++) {
					const double2 x = hB[(size_t)(p->offS + m) * np + j];
					segi[((size_t)f * np + j) * p->nseg + m] = d_complex(x.x, x.y);
				}
			if (usecap && (cellq != nullptr))
				for (int m = 0; m < p->ncell; m++) {
					const double2 x = hq[(size_t)m * np + j];
					cellq[((size_t)f * np + j) * p->ncell + m] = d_complex(x.x, x.y);
				}
		}
		if (st) {
			float a = 0;
			HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
			st->fill_ms += a;
			HIPCHK(hipEventElapsedTime(&a, ev[1], ev[2]));
			st->pot_ms += a;
			HIPCHK(hipEventElapsedTime(&a, ev[2], ev[3]));
			st->cinv_ms += a;
			HIPCHK(hipEventElapsedTime(&a, ev[3], ev[4]));
			st->asm_ms += a;
			HIPCHK(hipEventElapsedTime(&a, ev[4], ev[5]));
			st->lu_ms += a;
			HIPCHK(hipEventElapsedTime(&a, ev[5], ev[6]));
			st->solve_ms += a;
			st->total_ms += ms_since(t0);
		}
	}
	for (int k = 0; k < 8; k++) HIPCHK(hipEventDestroy(ev[k]));
	dfree(d_i1);
	dfree(d_i2);
	dfree(d_rowofcell);
	return rc;
}

void pgpu_accel_reset(pgpu_t *g)
{
	g->have_lu = false;
}

// One point of the reference's solve() on the dense path (solve.c:239-405): the system of this
// frequency, then nr right-hand sides (host, one n-vector each) solved on the device and copied
// back into x (nr x n).  accel 0: a fresh LU of A, as lu_decomp / lu_solve.  accel 1 (accel_rhs,
// solve.c:47-64): the LU kept from an earlier point preconditions one GMRES(60) cycle per
// right-hand side to a residual of 1e-10 |b| (gmres_solve); the first cycle that does not
// converge refactors at this point, and the fresh LU then solves that right-hand side and the
// rest.  q (may be NULL): the cell charges q = C v of the first nq solutions (ncell each, dense
// capacitive decks).  Returns 0, 1 (singular P) or 2 (singular MNA matrix).
int pgpu_solve_point(pgpu_t *g, const peec_t *p, int ifreq, int nr, const d_complex_t *rhs, d_complex_t *x, int nq,
                     d_complex_t *q, int accel, pgpu_accel_stats_t *as)
{
	const int n = p->nunknown;
	const double freq = freq_at(p, ifreq);
	ensure_cells(g);
	ensure_mna(g, p);
	const bool usecap = p->capacitance && (p->ncell > 0) && (g->cellrow != nullptr);
	if (!p->retardation) {
		if (p->nseg > 0) pgpu_lp_from_host(g, p->lp);
		if (usecap) pgpu_cmat_from_host(g, p->cmat);
	}
	if (point_system(g, p, freq, usecap, nullptr)) return 1;
	if (nr > g->xr_cap) {
		dfree(g->xr);
		g->xr = dalloc<double2>((size_t)n * nr);
		g->xr_cap = nr;
	}
	HIPCHK(hipMemcpyAsync(g->xr, rhs, (size_t)n * nr * sizeof(double2), hipMemcpyHostToDevice, g->s));
	auto vec = [&](int r) { return (d_complex_t *)(g->xr + (size_t)r * n); };
	if (!accel) {
		const int info = dlu_factor(g->s, n, (d_complex_t *)g->A, n, g->apiv, g->aperm, nullptr);
		if (info >= 0) {
			printf("*** singular matrix at f = %.5e Hz (row %d : floating node or ideal voltage source loop?)\n", freq, info);
			return 2;
		}
		for (int r = 0; r < nr; r++)
			dlu_solve_perm(g->s, n, (const d_complex_t *)g->A, n, g->aperm, vec(r), 1, 1, (d_complex_t *)g->Bw);
	}
	else {
		if (g->alu == nullptr) {
			g->alu = dalloc<double2>((size_t)n * n);
			g->alu_piv = dalloc<int>(n);
			g->alu_perm = dalloc<int>(n);
			g->gw = dgmres_ws_create(n);
		}
		auto refactor = [&]() -> int {
			HIPCHK(hipMemcpyAsync(g->alu, g->A, (size_t)n * n * sizeof(double2), hipMemcpyDeviceToDevice, g->s));
			const int info = dlu_factor(g->s, n, (d_complex_t *)g->alu, n, g->alu_piv, g->alu_perm, nullptr);
			if (info >= 0) {
				g->have_lu = false;
				printf("*** singular matrix at f = %.5e Hz (row %d : floating node or ideal voltage source loop?)\n", freq, info);
				return 2;
			}
			g->have_lu = true;
			if (as) as->nlu++;
			return 0;
		};
		bool fresh = false;
		if (!g->have_lu) {
			if (refactor()) return 2;
			fresh = true;
		}
		for (int r = 0; r < nr; r++) {
			if (!fresh) {
				const int it = dgmres_cycle(g->s, g->gw, n, (const d_complex_t *)g->A, n, (const d_complex_t *)g->alu, g->alu_perm,
				                            vec(r), 1e-10);
				if (it >= 0) {
					if (as) {
						as->ngm++;
						as->itsum += it;
					}
					continue;
				}
				if (refactor()) return 2;   // the preconditioner is stale: LU at this point
				fresh = true;
			}
			dlu_solve_perm(g->s, n, (const d_complex_t *)g->alu, n, g->alu_perm, vec(r), 1, 1, (d_complex_t *)g->Bw);
		}
	}
	if (usecap && (q != nullptr))
		for (int j = 0; j < nq; j++)
			hipLaunchKernelGGL(k_cell_charge, dim3((unsigned)((p->ncell + 255) / 256)), dim3(256), 0, g->s, g->C, p->ncell,
			                   g->rowofcell, (const double2 *)vec(j), 1, g->qbuf + (size_t)j * p->ncell);
	HIPCHK(hipMemcpyAsync(x, g->xr, (size_t)n * nr * sizeof(double2), hipMemcpyDeviceToHost, g->s));
	if (usecap && (q != nullptr) && (nq > 0))
		HIPCHK(hipMemcpyAsync(q, g->qbuf, (size_t)p->ncell * nq * sizeof(double2), hipMemcpyDeviceToHost, g->s));
	HIPCHK(hipStreamSynchronize(g->s));
	return 0;
}

}  // extern "C"
