/*
 * mlp.c — CPU statement of the learned preconditioner's network and its loss gradient.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The network has no counterpart in the reference
 * (OpenPEEC's preconditioner, precond.c, is not learned): this is the checker for the device
 * training kernels, written independently in plain C from the definition:
 *
 *   8 features -> 16 tanh -> 16 tanh -> 2 linear (Re theta, Im theta); parameters row-major
 *   W1[16][8] b1[16] W2[16][16] b2[16] W3[2][16] b3[2];
 *   loss = (1/S) sum_s (d_s + 2 Re(theta_s p_s) + |theta_s|^2 c_s) / d_s,
 *   sample coefficients coef[4 s ..] = (Re p, Im p, c, d).
 *
 * Pinned by tests/test_learned.py: its gradient against central finite differences of its own
 * loss (which leaves no room for a dropped term or a transposed weight: every parameter is
 * checked), and its loss against the closed-form minimum of a single sample when the hidden
 * layers are zero (theta = b3, optimum b3 = -conj(p) / c, loss 1 - |p|^2 / (c d)).
 */
#include <math.h>
#include <string.h>
#include "oracle.h"

#define NF 8
#define NH 16
#define OW1 0
#define OB1 (OW1 + NH * NF)
#define OW2 (OB1 + NH)
#define OB2 (OW2 + NH * NH)
#define OW3 (OB2 + NH)
#define OB3 (OW3 + 2 * NH)
#define NP (OB3 + 2)

int orc_mlp_nparams(void)
{
	return NP;
}

static void fwd(const double *W, const double *x, double *h1, double *h2, double *o)
{
	for (int i = 0; i < NH; i++) {
		double z = W[OB1 + i];
		for (int k = 0; k < NF; k++) z += W[OW1 + i * NF + k] * x[k];
		h1[i] = tanh(z);
	}
	for (int i = 0; i < NH; i++) {
		double z = W[OB2 + i];
		for (int k = 0; k < NH; k++) z += W[OW2 + i * NH + k] * h1[k];
		h2[i] = tanh(z);
	}
	for (int i = 0; i < 2; i++) {
		double z = W[OB3 + i];
		for (int k = 0; k < NH; k++) z += W[OW3 + i * NH + k] * h2[k];
		o[i] = z;
	}
}

void orc_mlp_forward(const double *W, const double *X, int S, double *theta)
{
	double h1[NH], h2[NH];
	for (int s = 0; s < S; s++) fwd(W, X + (long)s * NF, h1, h2, theta + 2 * s);
}

double orc_mlp_loss_grad(const double *W, const double *X, const double *coef, int S, double *grad)
{
	double loss = 0;
	if (grad) memset(grad, 0, NP * sizeof(double));
	for (int s = 0; s < S; s++) {
		const double *x = X + (long)s * NF;
		double h1[NH], h2[NH], o[2];
		fwd(W, x, h1, h2, o);
		const double pr = coef[4 * s], pi = coef[4 * s + 1], c = coef[4 * s + 2], d = coef[4 * s + 3];
		loss += (d + 2 * (o[0] * pr - o[1] * pi) + (o[0] * o[0] + o[1] * o[1]) * c) / d / S;
		if (!grad) continue;
		const double go[2] = {(2 * pr + 2 * o[0] * c) / d / S, (-2 * pi + 2 * o[1] * c) / d / S};
		double gz2[NH], gz1[NH];
		for (int k = 0; k < NH; k++) {
			double gh = 0;
			for (int i = 0; i < 2; i++) gh += W[OW3 + i * NH + k] * go[i];
			gz2[k] = gh * (1 - h2[k] * h2[k]);
		}
		for (int k = 0; k < NH; k++) {
			double gh = 0;
			for (int i = 0; i < NH; i++) gh += W[OW2 + i * NH + k] * gz2[i];
			gz1[k] = gh * (1 - h1[k] * h1[k]);
		}
		for (int i = 0; i < 2; i++) {
			grad[OB3 + i] += go[i];
			for (int k = 0; k < NH; k++) grad[OW3 + i * NH + k] += go[i] * h2[k];
		}
		for (int i = 0; i < NH; i++) {
			grad[OB2 + i] += gz2[i];
			for (int k = 0; k < NH; k++) grad[OW2 + i * NH + k] += gz2[i] * h1[k];
			grad[OB1 + i] += gz1[i];
			for (int k = 0; k < NF; k++) grad[OW1 + i * NF + k] += gz1[i] * x[k];
		}
	}
	return loss;
}
