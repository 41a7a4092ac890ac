/*
 * lu.c — oracle restatement of the dense complex LU (OpenPEEC src/lu.c:16-96).
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Right-looking unblocked Doolittle, row pivot =
 * first row of maximal |re| + |im| (lu.c:19-28), failure below 1e-30 (lu.c:30); rows of the
 * trailing update are independent, so the OpenMP split keeps results thread-count invariant.
 */
#include "oracle.h"

static double l1(d_complex_t z)
{
	return fabs(z.r) + fabs(z.i);
}

int orc_lu_decomp(int n, d_complex_t *a, int *piv)
{
	for (int k = 0; k < n; k++) {
		int p = k;
		double best = l1(a[(size_t)k * n + k]);
		for (int i = k + 1; i < n; i++) {
			const double m = l1(a[(size_t)i * n + k]);
			if (m > best) {
				best = m;
				p = i;
			}
		}
		piv[k] = p;
		if (best < 1e-30) return k;
		if (p != k)
			for (int j = 0; j < n; j++) {
				const d_complex_t t = a[(size_t)k * n + j];
				a[(size_t)k * n + j] = a[(size_t)p * n + j];
				a[(size_t)p * n + j] = t;
			}
		const d_complex_t inv = d_inv(a[(size_t)k * n + k]);
		int i;
#pragma omp parallel for if ((n - k) > 64)
		for (i = k + 1; i < n; i++) {
			const d_complex_t l = d_mul(a[(size_t)i * n + k], inv);
			a[(size_t)i * n + k] = l;
			for (int j = k + 1; j < n; j++) a[(size_t)i * n + j] = d_sub(a[(size_t)i * n + j], d_mul(l, a[(size_t)k * n + j]));
		}
	}
	return -1;
}

void orc_lu_solve(int n, const d_complex_t *a, const int *piv, d_complex_t *b)
{
	for (int k = 0; k < n; k++)
		if (piv[k] != k) {
			const d_complex_t t = b[k];
			b[k] = b[piv[k]];
			b[piv[k]] = t;
		}
	for (int i = 1; i < n; i++) {
		d_complex_t s = b[i];
		for (int j = 0; j < i; j++) s = d_sub(s, d_mul(a[(size_t)i * n + j], b[j]));
		b[i] = s;
	}
	for (int i = n - 1; i >= 0; i--) {
		d_complex_t s = b[i];
		for (int j = i + 1; j < n; j++) s = d_sub(s, d_mul(a[(size_t)i * n + j], b[j]));
		b[i] = d_mul(s, d_inv(a[(size_t)i * n + i]));
	}
}
