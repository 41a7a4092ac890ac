// fill_entries.hpp — one matrix entry of Lp and of the charge-cell integral matrix, as the
// fills, the k-factored fill's direct pairs and the H-matrix ACA evaluate it (device code).
//
//   Lp(i,j) = mu0/4pi [ tdot I(si,sj) - [gp] tdm I(si, M sj) ]   lp_entry, partial.c:371-424
//   G(h1,h2) = I(h1,h2) - [gp] I(h1, M h2)                       pot_fill, potential.c:217-237
// Both expect i <= j (the reference's quadrature order follows the first argument).
#pragma once
#include "elements.hpp"

namespace pe {

template <int KM>
__device__ cplx lp_entry(const dcell_t *S, const polys &pS, const dcell_t *SM, const polys &pSM, int i, int j,
                         double kw, double coef, int gp)
{
	const dcell_t si = S[i];
	if (i == j) {
		cplx v = self_k<KM>(si, pS, i, si.aL, kw);
		if (gp) {
			const dcell_t mi = SM[i];
			const double tdm = (si.t[0] * si.t[0]) + (si.t[1] * si.t[1]) - (si.t[2] * si.t[2]);
			v = csub(v, cscale(tdm, pair_k<KM>(si, pS, i, mi, pSM, i, si.aL, si.aL, kw, false)));
		}
		return cscale(coef, v);
	}
	const dcell_t sj = S[j];
	const double tdot = dot3(si.t, sj.t);
	const double tdm = (si.t[0] * sj.t[0]) + (si.t[1] * sj.t[1]) - (si.t[2] * sj.t[2]);
	cplx v = mk(0, 0);
	if (tdot != 0) v = cscale(tdot, pair_k<KM>(si, pS, i, sj, pS, j, si.aL, sj.aL, kw, false));
	if (gp && (tdm != 0)) {
		const dcell_t mj = SM[j];
		v = csub(v, cscale(tdm, pair_k<KM>(si, pS, i, mj, pSM, j, si.aL, sj.aL, kw, false)));
	}
	return cscale(coef, v);
}

template <int KM>
__device__ cplx ig_entry(const dcell_t *C, const polys &pC, const dcell_t *CM, const polys &pCM, int h1, int h2,
                         double kw, int gp)
{
	const dcell_t c1 = C[h1];
	cplx v;
	if (h1 == h2) {
		v = self_k<KM>(c1, pC, h1, c1.aP, kw);
	}
	else {
		const dcell_t c2 = C[h2];
		v = pair_k<KM>(c1, pC, h1, c2, pC, h2, c1.aP, c2.aP, kw, false);
	}
	if (gp) {
		const dcell_t m2 = CM[h2];
		v = csub(v, pair_k<KM>(c1, pC, h1, m2, pCM, h2, c1.aP, m2.aP, kw, false));
	}
	return v;
}

}  // namespace pe
