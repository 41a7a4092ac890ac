// fill_kernels.hpp — launch interface of the dense fills (fill.hip) for the runtime layer.
#pragma once
#include <hip/hip_runtime.h>
#include "../peec_internal.h"

namespace pe {

struct polys {            // polygon pools of one packed cell set
	const dedge_t *edge;
	const dpnorm_t *pn;
};

// Arguments of one fill launch over an n x n symmetric matrix cut into nt x nt tiles of 16.
struct fill_args {
	const dcell_t *S;          // cells (rows and columns)
	polys pS;
	const dcell_t *SM;         // ground-plane images of S (nullptr when gp == 0)
	polys pSM;
	int n, nt;
	long tile0;                // first tile of this launch (launches are chunked)
	double kw;                 // wavenumber 2 pi f / c0 (0: static)
	double coef;               // mu0 / 4 pi for Lp
	int gp;                    // ground plane present
	void *out;                 // n x n output, row-major, leading dimension ld
	long ld;
	const double *carea;       // fused P: per-cell length/area sums, else nullptr
	double k4pe;               // 4 pi eps0 (fused P)
};

}  // namespace pe

void fill_launch(hipStream_t s, int which, int km, const pe::fill_args &a);   // which: 0 Lp, 1 G/P
void fill_one(hipStream_t s, int which, int km, const pe::fill_args &a, int i, int j, const int *csoff,
              const int *csidx, double2 *res);
void aggregate_launch(hipStream_t s, const double2 *G, long ldg, const int *csoff, const int *csidx,
                      const double *carea, int nc, double k4pe, double2 *P, long ldp);
void fill_list_launch(hipStream_t s, int which, int km, const pe::fill_args &a, const int2 *list, int cnt);

// k-factored fill (kfill.hip)
long kf_slots(int n);
int kf_record_doubles(void);
void kf_rho_launch(hipStream_t s, const dcell_t *C, const dedge_t *edge, int n, double *rho);
void kf_setup_launch(hipStream_t s, int which, int km, const pe::fill_args &f, double klo, double khi, const double *rho,
                     double *rec, int T, int2 *list, unsigned *nlist);
void kf_eval_launch(hipStream_t s, int which, const pe::fill_args &f, const double *rec, int T);
// the same for nk <= KF_FMAX wavenumbers in one pass over the records (out[k]: matrix at kw[k],
// leading dimension f.ld; f.kw and f.out unused)
constexpr int KF_FMAX = 8;
void kf_eval_multi_launch(hipStream_t s, int which, const pe::fill_args &f, const double *rec, int T, int nk,
                          const double *kw, void *const *out, bool upper);

// cell set of a device mirror as fill arguments (gpu_runtime.hip); which 0 current cells, 1
// charge cells (carea set, sub-cell CSR returned)
pe::fill_args pgpu_set_args(pgpu_t *g, int which, double kw, int *km, const int **csoff, const int **csidx);
hipStream_t pgpu_hip_stream(pgpu_t *g);

// one structure's device mirror as the batched sweep sees it (batch.hip)
struct pgpu_bview {
	hipStream_t s;
	pe::fill_args seg, chg;     // cell sets (kw, out unset)
	int km_seg, km_chg;
	int nseg, nh, ncell;        // current cells, charge sub-cells, capacitive cells (0: none)
	bool ident;                 // one sub-cell per capacitive cell
	const double *rec_lp, *rec_ig;
	int T;                      // record terms per pair
	int nd_lp, nd_ig;           // pairs evaluated directly at every frequency
	const int2 *list_lp, *list_ig;
	const int *csoff, *csidx;
	const double *carea;
};
extern "C" {
int pgpu_kf_setup_launch(pgpu_t *g, double klo, double khi);
void pgpu_kf_setup_finish(pgpu_t *g);
void pgpu_batch_view(pgpu_t *g, pgpu_bview *v);
void pgpu_fill_direct_dev(pgpu_t *g, double kw, const void **lp, const void **P);
const double *pgpu_kf_records(pgpu_t *g, int which, double kw, long *nslots, int *T);
}
