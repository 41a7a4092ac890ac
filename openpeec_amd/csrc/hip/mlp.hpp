// mlp.hpp — sizes and device entry points of the learned preconditioner's network (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

#define MLP_NF 8      // features
#define MLP_NH 16     // hidden width (two hidden layers)
#define MLP_NP (MLP_NH * MLP_NF + MLP_NH + MLP_NH * MLP_NH + MLP_NH + 2 * MLP_NH + 2)

void mlp_forward_dev(hipStream_t s, const double *dW, const double *dX, int S, double *dtheta);
void mlp_grad_dev(hipStream_t s, const double *dW, const double *dX, const double *dcoef, int S, double *dG, double *dL,
                  double *dgrad, double *dloss);
void mlp_init(double *W, unsigned long long seed);
void mlp_fit_dev(hipStream_t s, const double *X, const double *coef, int S, int steps, double lr, double *W,
                 double *loss_hist);
