/* A caller-side kernel for hmat_build_gen tests (test infrastructure): entries
 * A(i, j) = exp(-j kw r) / (r + 0.01) between points ctx[3 i .. 3 i + 2] and ctx[3 j ..]. */
#include <math.h>

typedef struct {
	double r, i;
} zc_t;

zc_t helix_kernel(void *ctx, int i, int j, double kw)
{
	const double *p = (const double *)ctx;
	const double dx = p[3 * i] - p[3 * j], dy = p[3 * i + 1] - p[3 * j + 1], dz = p[3 * i + 2] - p[3 * j + 2];
	const double r = sqrt((dx * dx) + (dy * dy) + (dz * dz));
	const double a = 1 / (r + 0.01);
	zc_t v = {a * cos(kw * r), -a * sin(kw * r)};
	return v;
}
