/*
 * hmat_host.c — cluster tree, block partition and leaf-row grouping of the H-matrix
 * (compression = 1).  Geometry only, so it is built once per cell set on the host and reused
 * at every frequency; the entries (ACA and near blocks) are computed on the device
 * (hmat.hip).
 *
 * Same tree and blocks as the reference (hmatrix.c:43-227, 483-586): bisection of the
 * bounding box along its longest axis at the box centre (median index when one side would be
 * empty), leaves of at most HM_LEAF cells, admissibility min(diam) <= HM_ETA * dist between
 * bounding spheres, far blocks low-rank, near blocks (a leaf on either side) dense; the leaf
 * clusters tile [0, n) of the permutation, and each leaf lists, in registration order, the far
 * and near blocks whose row range covers it (the matvec accumulates in that order).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "peec.h"
#include "peec_internal.h"

static void ct_grow(hm_tree_t *t)
{
	if (t->nc < t->cap) return;
	t->cap = t->cap ? 2 * t->cap : 64;
	t->lo = (int *)realloc(t->lo, (size_t)t->cap * sizeof(int));
	t->hi = (int *)realloc(t->hi, (size_t)t->cap * sizeof(int));
	t->left = (int *)realloc(t->left, (size_t)t->cap * sizeof(int));
	t->right = (int *)realloc(t->right, (size_t)t->cap * sizeof(int));
	t->cen = (double *)realloc(t->cen, (size_t)t->cap * 3 * sizeof(double));
	t->rad = (double *)realloc(t->rad, (size_t)t->cap * sizeof(double));
}

/* hmatrix.c:86-148 */
static int ct_node(hm_tree_t *t, const double *pts, int lo, int hi)
{
	ct_grow(t);
	const int id = t->nc++;
	int *perm = t->perm;
	t->lo[id] = lo;
	t->hi[id] = hi;
	t->left[id] = t->right[id] = -1;
	double bmin[3], bmax[3];
	for (int c = 0; c < 3; c++) bmin[c] = bmax[c] = pts[3 * perm[lo] + c];
	for (int i = lo + 1; i < hi; i++) {
		const double *x = &pts[3 * perm[i]];
		for (int c = 0; c < 3; c++) {
			if (x[c] < bmin[c]) bmin[c] = x[c];
			if (x[c] > bmax[c]) bmax[c] = x[c];
		}
	}
	double d2 = 0;
	for (int c = 0; c < 3; c++) {
		t->cen[3 * id + c] = 0.5 * (bmin[c] + bmax[c]);
		const double d = bmax[c] - bmin[c];
		d2 += d * d;
	}
	t->rad[id] = 0.5 * sqrt(d2);
	if ((hi - lo) <= HM_LEAF) return id;
	int ax = 0;
	for (int c = 1; c < 3; c++)
		if ((bmax[c] - bmin[c]) > (bmax[ax] - bmin[ax])) ax = c;
	if ((bmax[ax] - bmin[ax]) <= 0) return id;
	const double cut = 0.5 * (bmin[ax] + bmax[ax]);
	int i = lo, j = hi - 1;
	while (i <= j) {
		if (pts[3 * perm[i] + ax] < cut) {
			i++;
		}
		else {
			const int tmp = perm[i];
			perm[i] = perm[j];
			perm[j] = tmp;
			j--;
		}
	}
	int mid = i;
	if ((mid == lo) || (mid == hi)) mid = (lo + hi) / 2;
	const int l = ct_node(t, pts, lo, mid);
	const int r = ct_node(t, pts, mid, hi);
	t->left[id] = l;
	t->right[id] = r;
	return id;
}

/* hmatrix.c:151-169 */
static int admissible(const hm_tree_t *t, int a, int b)
{
	double d2 = 0;
	for (int c = 0; c < 3; c++) {
		const double d = t->cen[3 * a + c] - t->cen[3 * b + c];
		d2 += d * d;
	}
	double dist = sqrt(d2) - t->rad[a] - t->rad[b];
	if (!(dist > 0)) dist = 0;
	const double dmin = 2 * ((t->rad[a] < t->rad[b]) ? t->rad[a] : t->rad[b]);
	return (dist > 0) && (dmin <= HM_ETA * dist);
}

static void push(int **r, int **c, int *n, int *cap, int a, int b)
{
	if (*n >= *cap) {
		*cap = *cap ? 2 * *cap : 256;
		*r = (int *)realloc(*r, (size_t)*cap * sizeof(int));
		*c = (int *)realloc(*c, (size_t)*cap * sizeof(int));
	}
	(*r)[*n] = a;
	(*c)[*n] = b;
	(*n)++;
}

/* hmatrix.c:211-227 */
static void split(hm_tree_t *t, int a, int b, int *fcap, int *ncap)
{
	if (admissible(t, a, b)) {
		push(&t->fr, &t->fc, &t->nfar, fcap, a, b);
		return;
	}
	if ((t->left[a] < 0) || (t->left[b] < 0)) {
		push(&t->nr, &t->nc2, &t->nnear, ncap, a, b);
		return;
	}
	split(t, t->left[a], t->left[b], fcap, ncap);
	split(t, t->left[a], t->right[b], fcap, ncap);
	split(t, t->right[a], t->left[b], fcap, ncap);
	split(t, t->right[a], t->right[b], fcap, ncap);
}

/* leaf-row CSR of the blocks covering each leaf (hmatrix.c:483-586) */
static int leaf_lists(hm_tree_t *t)
{
	const int n = t->n;
	int *leafat = (int *)malloc((size_t)n * sizeof(int));
	int *lidx = (int *)malloc((size_t)n * sizeof(int));
	if (!leafat || !lidx) {
		free(leafat);
		free(lidx);
		return 1;
	}
	for (int i = 0; i < n; i++) leafat[i] = -1;
	for (int i = 0; i < t->nc; i++)
		if (t->left[i] < 0) leafat[t->lo[i]] = i;
	t->nleaf = 0;
	for (int i = 0; i < n; i++)
		if (leafat[i] >= 0) t->nleaf++;
	t->leafid = (int *)malloc((size_t)t->nleaf * sizeof(int));
	t->rfoff = (int *)calloc((size_t)t->nleaf + 1, sizeof(int));
	t->rnoff = (int *)calloc((size_t)t->nleaf + 1, sizeof(int));
	int e = 0;
	for (int i = 0; i < n; i++)
		if (leafat[i] >= 0) t->leafid[e++] = leafat[i];
	e = -1;
	for (int i = 0; i < n; i++) {
		if (leafat[i] >= 0) e++;
		lidx[i] = e;
	}
	free(leafat);
	int bad = (t->nleaf <= 0) || (t->lo[t->leafid[0]] != 0) || (t->hi[t->leafid[t->nleaf - 1]] != n);
	for (int q = 0; !bad && (q + 1 < t->nleaf); q++)
		if (t->hi[t->leafid[q]] != t->lo[t->leafid[q + 1]]) bad = 1;
	if (bad) {
		free(lidx);
		return 1;
	}
	for (int b = 0; b < t->nfar; b++)
		for (int q = lidx[t->lo[t->fr[b]]]; q <= lidx[t->hi[t->fr[b]] - 1]; q++) t->rfoff[q + 1]++;
	for (int b = 0; b < t->nnear; b++)
		for (int q = lidx[t->lo[t->nr[b]]]; q <= lidx[t->hi[t->nr[b]] - 1]; q++) t->rnoff[q + 1]++;
	for (int q = 0; q < t->nleaf; q++) {
		t->rfoff[q + 1] += t->rfoff[q];
		t->rnoff[q + 1] += t->rnoff[q];
	}
	t->rfidx = (int *)malloc((size_t)(t->rfoff[t->nleaf] > 0 ? t->rfoff[t->nleaf] : 1) * sizeof(int));
	t->rnidx = (int *)malloc((size_t)(t->rnoff[t->nleaf] > 0 ? t->rnoff[t->nleaf] : 1) * sizeof(int));
	int *fp = (int *)malloc((size_t)(t->nleaf + 1) * sizeof(int));
	int *np = (int *)malloc((size_t)(t->nleaf + 1) * sizeof(int));
	memcpy(fp, t->rfoff, (size_t)(t->nleaf + 1) * sizeof(int));
	memcpy(np, t->rnoff, (size_t)(t->nleaf + 1) * sizeof(int));
	for (int b = 0; b < t->nfar; b++)
		for (int q = lidx[t->lo[t->fr[b]]]; q <= lidx[t->hi[t->fr[b]] - 1]; q++) t->rfidx[fp[q]++] = b;
	for (int b = 0; b < t->nnear; b++)
		for (int q = lidx[t->lo[t->nr[b]]]; q <= lidx[t->hi[t->nr[b]] - 1]; q++) t->rnidx[np[q]++] = b;
	free(fp);
	free(np);
	free(lidx);
	return 0;
}

int hm_tree_build(hm_tree_t *t, int n, const double *pts)
{
	memset(t, 0, sizeof(*t));
	if (n <= 0) return 1;
	t->n = n;
	t->perm = (int *)malloc((size_t)n * sizeof(int));
	for (int i = 0; i < n; i++) t->perm[i] = i;
	ct_node(t, pts, 0, n);
	int fcap = 0, ncap = 0;
	split(t, 0, 0, &fcap, &ncap);
	if (leaf_lists(t)) {
		hm_tree_free(t);
		return 1;
	}
	return 0;
}

void hm_tree_free(hm_tree_t *t)
{
	free(t->perm);
	free(t->lo);
	free(t->hi);
	free(t->left);
	free(t->right);
	free(t->cen);
	free(t->rad);
	free(t->fr);
	free(t->fc);
	free(t->nr);
	free(t->nc2);
	free(t->leafid);
	free(t->rfoff);
	free(t->rfidx);
	free(t->rnoff);
	free(t->rnidx);
	memset(t, 0, sizeof(*t));
}

/* block partition for tests: counts = {far, near, leaves, n}; perm (n), far / near as
 * (r0, m, c0, nn) per block when the arrays are given */
int hmat_tree_info(int n, const double *pts, int *counts, int *perm, int *far, int *near)
{
	hm_tree_t t;
	if (hm_tree_build(&t, n, pts)) return 1;
	counts[0] = t.nfar;
	counts[1] = t.nnear;
	counts[2] = t.nleaf;
	counts[3] = n;
	if (perm) memcpy(perm, t.perm, (size_t)n * sizeof(int));
	for (int b = 0; far && (b < t.nfar); b++) {
		far[4 * b] = t.lo[t.fr[b]];
		far[4 * b + 1] = t.hi[t.fr[b]] - t.lo[t.fr[b]];
		far[4 * b + 2] = t.lo[t.fc[b]];
		far[4 * b + 3] = t.hi[t.fc[b]] - t.lo[t.fc[b]];
	}
	for (int b = 0; near && (b < t.nnear); b++) {
		near[4 * b] = t.lo[t.nr[b]];
		near[4 * b + 1] = t.hi[t.nr[b]] - t.lo[t.nr[b]];
		near[4 * b + 2] = t.lo[t.nc2[b]];
		near[4 * b + 3] = t.hi[t.nc2[b]] - t.lo[t.nc2[b]];
	}
	hm_tree_free(&t);
	return 0;
}
