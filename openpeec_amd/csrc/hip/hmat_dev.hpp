// hmat_dev.hpp — device representation of a compressed kernel matrix (hmat.hip) for the
// modules that apply it (the compressed solver, hsolve.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdio>
#include <vector>
#include "fill_kernels.hpp"

namespace pe {

// entry context: which kernel (0 Lp, 1 P aggregated over sub-cells) over which cells
struct hent {
	fill_args f;
	int kind;
	const int *csoff, *csidx;
	const double *rec;        // k-factored records of the cell set covering f.kw, or null
	long nslots;
	int T;
	// every entry at this wavenumber in original numbering, full[i * nfull + j] = A(i, j),
	// evaluated in bulk before the build, or null
	const double2 *full;
	int nfull;
};

// far block: row cluster [r0, r0+m), column cluster [c0, c0+nn) of the permutation; rank > 0:
// U (rank x m) at uo, V (rank x nn) at vo, V^T x at fo; rank 0: dense m x nn at uo in the D pool
struct hfar {
	int r0, m, c0, nn;
	int rank, pad;
	long uo, vo, fo;
};

struct hdense {
	int r0, m, c0, nn;
	long off;
};

struct hleafs {
	const int *lo, *hi;            // leaf row ranges (permuted positions)
	const int *rfoff, *rfidx;      // far blocks covering each leaf, registered order
	const int *rnoff, *rnidx;      // near blocks
};

struct aca_args {
	hent E;
	const int *perm;
	const hfar *far;
	const int *order;              // far blocks in launch order (largest first)
	const long *soff, *moff;       // scratch offsets: factors (double2); row then column flags
	double2 *scr;
	double *urep;                  // per block: max |u_l[i]| over rows, then max |v_l[j]| over columns
	int *rused;                    // per block: rows visited, then columns pivoted or checked
	int *rank;
	double tol;
};

}  // namespace pe

struct hmat_t {
	int n = 0, kind = 0, km = 0;
	double kw = 0, tol = 0;
	pgpu_t *g = nullptr;
	hipStream_t s = nullptr;
	bool own_stream = false;     // built from caller entries (hmat_build_gen): its own stream
	hm_tree_t t{};
	int *d_perm = nullptr;
	pe::hfar *d_far = nullptr;
	pe::hdense *d_near = nullptr;
	double2 *d_U = nullptr, *d_V = nullptr, *d_D = nullptr;
	long nU = 0, nV = 0, nD = 0, ftlen = 0;
	pe::hleafs L{};
	double2 *d_ft = nullptr, *d_xp = nullptr, *d_x = nullptr, *d_y = nullptr;
	double2 *d_part = nullptr;   // per-row partial sums of the leaf products (HM_SPLIT x n)
	int *d_item = nullptr;       // far block of each (block, rank) pair of V^T x, by ft index
	std::vector<int> rank;
	std::vector<pe::hfar> far;
	std::vector<pe::hdense> near;
	std::vector<double2> leafblk;   // host copies of the leaf diagonal blocks (hmat_leaf_block)
	std::vector<long> leafoff;
	std::vector<int> leafhas;
};

/* g: the device mirror whose cells give the entries (kind 0 Lp, 1 P); or NULL with host_full,
 * the n x n entries in original numbering evaluated by the caller (hmat_build_gen) */
struct hmat_t *hm_build_dev(pgpu_t *g, int kind, int n, const double *pts, double kw, double tol, const char *label,
                            FILE *fp_log, const double2 *host_full = nullptr);
void hm_matvec_dev(const hmat_t *h, const double2 *x, double2 *y);
void hm_near_off_dev(const hmat_t *h, const double2 *x, double2 *y);
void hm_near_norms(const hmat_t *h, double *off, double *diag);
