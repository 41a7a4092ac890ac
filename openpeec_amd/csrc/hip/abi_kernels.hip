// abi_kernels.hip — the scalar element entry points and the dense LU of peec.h, evaluated on
// the GPU through the same device code as the fills.
//
// neumann_* (partial.c:55-328), rect_potential / ribbon_* (surface.c:76-178), bar_* (volume.c:
// 92-155), poly_potential / poly_static / poly_corr (polygon.c:68-334) take host seg_t records:
// each call packs its cells, runs one device thread and returns the value.  lu_decomp /
// lu_solve (lu.c:16-96) stage the host matrix through device memory around the blocked device
// LU.  These exist for ABI completeness and parity tests; the sweep never goes through them.
#include <hip/hip_runtime.h>
#include <cstring>
#include <vector>
#include "elements.hpp"
#include "hip_check.hpp"

namespace {

// one packed cell set of up to 2 cells on the device
struct tiny {
	dcell_t *c = nullptr;
	dedge_t *e = nullptr;
	dpnorm_t *pn = nullptr;
	pe::polys pp() const { return pe::polys{e, pn}; }
	~tiny()
	{
		if (c) HIPCHK(hipFree(c));
		if (e) HIPCHK(hipFree(e));
		if (pn) HIPCHK(hipFree(pn));
	}
};

void put(tiny &t, const seg_t *a, const seg_t *b)
{
	int hip_dev_count = 0;
	if ((hipGetDeviceCount(&hip_dev_count) != hipSuccess) || (hip_dev_count <= 0)) pgpu_fatal("no HIP device");
	seg_t two[2];
	two[0] = *a;
	two[1] = (b != nullptr) ? *b : *a;
	dcell_t *hc;
	dedge_t *he;
	dpnorm_t *hp;
	int ne = 0;
	if (cells_pack(two, 2, 0, 0, &hc, &he, &hp, &ne)) pgpu_fatal("cell packing");
	HIPCHK(hipMalloc((void **)&t.c, 2 * sizeof(dcell_t)));
	HIPCHK(hipMalloc((void **)&t.e, (size_t)(ne > 0 ? ne : 1) * sizeof(dedge_t)));
	HIPCHK(hipMalloc((void **)&t.pn, 2 * sizeof(dpnorm_t)));
	HIPCHK(hipMemcpy(t.c, hc, 2 * sizeof(dcell_t), hipMemcpyHostToDevice));
	if (ne > 0) HIPCHK(hipMemcpy(t.e, he, (size_t)ne * sizeof(dedge_t), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(t.pn, hp, 2 * sizeof(dpnorm_t), hipMemcpyHostToDevice));
	free(hc);
	free(he);
	free(hp);
}

enum op_t { OP_SELF, OP_PAIR, OP_SELFK, OP_PAIRK, OP_RECT, OP_RSTAT, OP_RCORR, OP_HLUSE, OP_HL, OP_PPOT, OP_PSTAT, OP_PCORR };

struct opargs {
	int op;
	double l, a1, a2, kw;
	int nsub, nq, same;
	double pt[3];
};

__global__ void k_eval(const dcell_t *c, pe::polys pp, opargs o, double2 *res)
{
	const dcell_t &s1 = c[0];
	const dcell_t &s2 = c[1];
	pe::cplx v = pe::mk(0, 0);
	switch (o.op) {
	case OP_SELF: v = pe::mk(pe::wire_self_static(o.l, o.a1), 0); break;
	case OP_PAIR: v = pe::mk(pe::wire_pair_static(s1, s2, o.a1, o.a2), 0); break;
	case OP_SELFK: v = pe::self_k<15>(s1, pp, 0, o.a1, o.kw); break;
	case OP_PAIRK: v = pe::pair_k<15>(s1, pp, 0, s2, pp, 1, o.a1, o.a2, o.kw, o.same != 0); break;
	case OP_RECT: v = pe::mk(pe::rect_potential(s1, o.pt), 0); break;
	case OP_RSTAT: v = pe::mk(pe::ribbon_static(s1, s2, o.nsub, o.nq), 0); break;
	case OP_RCORR: v = pe::ribbon_corr(s1, s2, o.kw, o.nsub); break;
	case OP_HLUSE: v = pe::mk(pe::bar_use_hl(s1, s2), 0); break;
	case OP_HL: v = pe::mk(pe::bar_pair(s1, s2), 0); break;
	case OP_PPOT: v = pe::mk(pe::poly_potential(s1, pp, 0, o.pt), 0); break;
	case OP_PSTAT: v = pe::mk(pe::poly_static(s1, pp, 0, s2, pp, 1, o.nsub, o.nq), 0); break;
	case OP_PCORR: v = pe::poly_corr(s1, pp, s2, pp, o.kw); break;
	}
	*res = make_double2(v.r, v.i);
}

d_complex_t run(const seg_t *a, const seg_t *b, opargs o)
{
	tiny t;
	put(t, a, b);
	double2 *res = nullptr;
	HIPCHK(hipMalloc((void **)&res, sizeof(double2)));
	hipLaunchKernelGGL(k_eval, dim3(1), dim3(1), 0, 0, t.c, t.pp(), o, res);
	HIPCHK(hipGetLastError());
	double2 h;
	HIPCHK(hipMemcpy(&h, res, sizeof(h), hipMemcpyDeviceToHost));
	HIPCHK(hipFree(res));
	return d_complex(h.x, h.y);
}

opargs mkop(int op)
{
	opargs o;
	memset(&o, 0, sizeof(o));
	o.op = op;
	return o;
}

}  // namespace

extern "C" {

double neumann_self(double l, double a)
{
	seg_t s;
	memset(&s, 0, sizeof(s));
	s.len = 1;
	s.x2[0] = 1;
	opargs o = mkop(OP_SELF);
	o.l = l;
	o.a1 = a;
	return run(&s, nullptr, o).r;
}

double neumann_pair(const seg_t *s1, const seg_t *s2, double a1, double a2)
{
	opargs o = mkop(OP_PAIR);
	o.a1 = a1;
	o.a2 = a2;
	return run(s1, s2, o).r;
}

d_complex_t neumann_self_k(const seg_t *s, double a, double kw)
{
	opargs o = mkop(OP_SELFK);
	o.a1 = a;
	o.kw = kw;
	return run(s, nullptr, o);
}

d_complex_t neumann_pair_k(const seg_t *s1, const seg_t *s2, double a1, double a2, double kw)
{
	opargs o = mkop(OP_PAIRK);
	o.a1 = a1;
	o.a2 = a2;
	o.kw = kw;
	o.same = (s1 == s2);
	return run(s1, s2, o);
}

double rect_potential(const seg_t *s, const double *pt)
{
	opargs o = mkop(OP_RECT);
	for (int k = 0; k < 3; k++) o.pt[k] = pt[k];
	return run(s, nullptr, o).r;
}

double ribbon_static(const seg_t *s1, const seg_t *s2, int nsub, int nq)
{
	opargs o = mkop(OP_RSTAT);
	o.nsub = nsub;
	o.nq = nq;
	return run(s1, s2, o).r;
}

d_complex_t ribbon_corr(const seg_t *s1, const seg_t *s2, double kw, int nsub)
{
	opargs o = mkop(OP_RCORR);
	o.kw = kw;
	o.nsub = nsub;
	return run(s1, s2, o);
}

int bar_use_hl(const seg_t *s1, const seg_t *s2)
{
	return (int)run(s1, s2, mkop(OP_HLUSE)).r;
}

double bar_pair(const seg_t *s1, const seg_t *s2)
{
	return run(s1, s2, mkop(OP_HL)).r;
}

double poly_potential(const seg_t *s, const double *pt)
{
	opargs o = mkop(OP_PPOT);
	for (int k = 0; k < 3; k++) o.pt[k] = pt[k];
	return run(s, nullptr, o).r;
}

double poly_static(const seg_t *s1, const seg_t *s2, int nsub, int n7)
{
	opargs o = mkop(OP_PSTAT);
	o.nsub = nsub;
	o.nq = n7;
	return run(s1, s2, o).r;
}

d_complex_t poly_corr(const seg_t *s1, const seg_t *s2, double kw)
{
	opargs o = mkop(OP_PCORR);
	o.kw = kw;
	return run(s1, s2, o);
}

// ---- dense LU through host arrays ----

int lu_decomp(int n, d_complex_t *a, int *piv)
{
	if (n <= 0) return -1;
	int hip_dev_count = 0;
	if ((hipGetDeviceCount(&hip_dev_count) != hipSuccess) || (hip_dev_count <= 0)) pgpu_fatal("no HIP device");
	d_complex_t *dA = nullptr;
	int *dpiv = nullptr;
	const size_t bytes = (size_t)n * n * sizeof(d_complex_t);
	HIPCHK(hipMalloc((void **)&dA, bytes));
	HIPCHK(hipMalloc((void **)&dpiv, (size_t)n * sizeof(int)));
	HIPCHK(hipMemcpy(dA, a, bytes, hipMemcpyHostToDevice));
	const int info = dlu_factor(nullptr, n, dA, n, dpiv, nullptr, nullptr);
	HIPCHK(hipMemcpy(a, dA, bytes, hipMemcpyDeviceToHost));
	HIPCHK(hipMemcpy(piv, dpiv, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
	HIPCHK(hipFree(dA));
	HIPCHK(hipFree(dpiv));
	return info;
}

// precision form's factorisation: the matrix is rounded to FP32 storage and factored there
// (FP64 panel arithmetic and pivot search, FP32 MFMA trailing update); factors come back
// widened to FP64 in the layout lu_decomp writes
int lu_decomp32(int n, d_complex_t *a, int *piv)
{
	if (n <= 0) return -1;
	int hip_dev_count = 0;
	if ((hipGetDeviceCount(&hip_dev_count) != hipSuccess) || (hip_dev_count <= 0)) pgpu_fatal("no HIP device");
	const long cnt = (long)n * n;
	std::vector<float2> h((size_t)cnt);
	for (long e = 0; e < cnt; e++) h[e] = make_float2((float)a[e].r, (float)a[e].i);
	float2 *dA = nullptr;
	int *dpiv = nullptr;
	HIPCHK(hipMalloc((void **)&dA, (size_t)cnt * sizeof(float2)));
	HIPCHK(hipMalloc((void **)&dpiv, (size_t)n * sizeof(int)));
	HIPCHK(hipMemcpy(dA, h.data(), (size_t)cnt * sizeof(float2), hipMemcpyHostToDevice));
	const int info = slu_factor(nullptr, n, dA, n, dpiv, nullptr, nullptr);
	HIPCHK(hipMemcpy(h.data(), dA, (size_t)cnt * sizeof(float2), hipMemcpyDeviceToHost));
	HIPCHK(hipMemcpy(piv, dpiv, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
	HIPCHK(hipFree(dA));
	HIPCHK(hipFree(dpiv));
	for (long e = 0; e < cnt; e++) a[e] = d_complex(h[e].x, h[e].y);
	return info;
}

void lu_solve(int n, const d_complex_t *a, const int *piv, d_complex_t *b)
{
	if (n <= 0) return;
	int hip_dev_count = 0;
	if ((hipGetDeviceCount(&hip_dev_count) != hipSuccess) || (hip_dev_count <= 0)) pgpu_fatal("no HIP device");
	// the row permutation of the recorded interchanges (index bookkeeping only)
	std::vector<int> perm(n);
	for (int k = 0; k < n; k++) perm[k] = k;
	for (int k = 0; k < n; k++)
		if (piv[k] != k) std::swap(perm[k], perm[piv[k]]);
	d_complex_t *dA = nullptr, *dB = nullptr, *dW = nullptr;
	int *dperm = nullptr;
	const size_t bytes = (size_t)n * n * sizeof(d_complex_t);
	HIPCHK(hipMalloc((void **)&dA, bytes));
	HIPCHK(hipMalloc((void **)&dB, (size_t)n * sizeof(d_complex_t)));
	HIPCHK(hipMalloc((void **)&dW, (size_t)n * sizeof(d_complex_t)));
	HIPCHK(hipMalloc((void **)&dperm, (size_t)n * sizeof(int)));
	HIPCHK(hipMemcpy(dA, a, bytes, hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(dB, b, (size_t)n * sizeof(d_complex_t), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(dperm, perm.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice));
	dlu_solve_perm(nullptr, n, dA, n, dperm, dB, 1, 1, dW);
	HIPCHK(hipMemcpy(b, dB, (size_t)n * sizeof(d_complex_t), hipMemcpyDeviceToHost));
	HIPCHK(hipFree(dA));
	HIPCHK(hipFree(dB));
	HIPCHK(hipFree(dW));
	HIPCHK(hipFree(dperm));
}

}  // extern "C"
