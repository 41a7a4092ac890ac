// kf_eval.hpp — record layout of the k-factored fill and the evaluation of one pair at a
// wavenumber (shared by the per-structure fill and the batched sweep).
//
// Record of one term, structure-of-arrays over nslots slots: q = 0 base (static part minus the
// m = 0 moment), q = 1 centre distance Rc, q = 2 .. 2 + KF_M the scaled moments N'_m.  Slot of
// the pair (i <= j): tile * 256 + (i mod 16) * 16 + (j mod 16), tiles of 16 x 16 numbered row by
// row over the upper triangle of the tile grid.
#pragma once
#include <hip/hip_runtime.h>

namespace pe {

constexpr int KF_M = 14;
constexpr int KF_Q = KF_M + 3;

__device__ __forceinline__ long kf_slot(int i, int j, int nt)
{
	const long ti = i >> 4, tj = j >> 4;
	const long tile = ti * nt - (ti * (ti - 1)) / 2 + (tj - ti);
	return tile * 256 + ((i & 15) << 4) + (j & 15);
}

// false for a pair the records list for direct evaluation (Rc of its first term is -1)
__device__ __forceinline__ bool kf_has(const double *rec, long nslots, long slot)
{
	return rec[nslots + slot] >= 0;
}

// sum over the T terms of  base + exp(-jk Rc) sum_m (-jk)^m / m! N'_m  (Horner from the top)
__device__ __forceinline__ void kf_value(const double *rec, long nslots, long slot, int T, double kw, double &vr,
                                         double &vi)
{
	vr = 0;
	vi = 0;
	for (int t = 0; t < T; t++) {
		const double *r = rec + (long)t * KF_Q * nslots + slot;
		double zr = r[(long)(2 + KF_M) * nslots], zi = 0;
#pragma unroll
		for (int m = KF_M - 1; m >= 0; m--) {
			const double s = kw / (m + 1);
			const double nr = zi * s, ni = -zr * s;   // z * (-j s)
			zr = r[(long)(2 + m) * nslots] + nr;
			zi = ni;
		}
		const double Rc = r[nslots];
		double er = zr, ei = zi;
		if (Rc != 0) {
			double sn, cs;
			sincos(kw * Rc, &sn, &cs);
			er = (cs * zr) + (sn * zi);
			ei = (cs * zi) - (sn * zr);
		}
		vr += r[0] + er;
		vi += ei;
	}
}

// one term of a record held in registers (q[0 .. KF_Q)) at wavenumber kw, added to (vr, vi)
__device__ __forceinline__ void kf_term(const double (&q)[KF_Q], double kw, double &vr, double &vi)
{
	double zr = q[2 + KF_M], zi = 0;
#pragma unroll
	for (int m = KF_M - 1; m >= 0; m--) {
		const double s = kw / (m + 1);
		const double nr = zi * s, ni = -zr * s;
		zr = q[2 + m] + nr;
		zi = ni;
	}
	const double Rc = q[1];
	double er = zr, ei = zi;
	if (Rc != 0) {
		double sn, cs;
		sincos(kw * Rc, &sn, &cs);
		er = (cs * zr) + (sn * zi);
		ei = (cs * zi) - (sn * zr);
	}
	vr += q[0] + er;
	vi += ei;
}

}  // namespace pe
