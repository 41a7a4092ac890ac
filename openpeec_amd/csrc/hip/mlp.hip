// mlp.hip — the small network of the learned preconditioner correction, with its training
// (forward, backward, Adam) and inference kernels.
//
// The network maps 8 features of a compressed system (wavenumbers times sizes, near-field
// coupling ratios; see the compressed solver) to the complex weight theta of the near-field
// correction  P^-1 = M^-1 - theta M^-1 N M^-1.  Layers: 8 -> 16 (tanh) -> 16 (tanh) -> 2 (linear:
// Re theta, Im theta).  Parameters, row-major, in one array of MLP_NP doubles:
//   W1[16][8] b1[16] W2[16][16] b2[16] W3[2][16] b3[2]
//
// Training sample s carries its features x_s and the coefficients (p, c, d) of a quadratic loss
// in theta, (d + 2 Re(theta p) + |theta|^2 c) / d.  The compressed solver writes its samples with
// c = 1, p = -conj(theta*), d = 1 + |theta*|^2, theta* the weight with which GMRES needed the
// fewest iterations on that system (measured over a grid), so the loss is
// (1 + |theta - theta*|^2) / (1 + |theta*|^2).  The batch loss is the mean over samples.
// Per-sample gradients are written out and summed in sample order by one workgroup
// (deterministic).
#include <hip/hip_runtime.h>
#include <cmath>
#include <vector>
#include "hip_check.hpp"
#include "mlp.hpp"

namespace {

constexpr int F = MLP_NF, H = MLP_NH;
constexpr int OW1 = 0, OB1 = OW1 + H * F, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + 2 * H;
static_assert(OB3 + 2 == MLP_NP, "parameter layout");

struct fwd {
	double h1[H], h2[H];
	double o[2];
};

__device__ void forward(const double *W, const double *x, fwd &a)
{
	for (int i = 0; i < H; i++) {
		double z = W[OB1 + i];
		for (int k = 0; k < F; k++) z += W[OW1 + i * F + k] * x[k];
		a.h1[i] = tanh(z);
	}
	for (int i = 0; i < H; i++) {
		double z = W[OB2 + i];
		for (int k = 0; k < H; k++) z += W[OW2 + i * H + k] * a.h1[k];
		a.h2[i] = tanh(z);
	}
	for (int i = 0; i < 2; i++) {
		double z = W[OB3 + i];
		for (int k = 0; k < H; k++) z += W[OW3 + i * H + k] * a.h2[k];
		a.o[i] = z;
	}
}

__global__ __launch_bounds__(256) void k_mlp_forward(const double *W, const double *X, int S, double *theta)
{
	__shared__ double w[MLP_NP];
	for (int e = threadIdx.x; e < MLP_NP; e += blockDim.x) w[e] = W[e];
	__syncthreads();
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= S) return;
	fwd a;
	forward(w, X + (long)s * F, a);
	theta[2 * s] = a.o[0];
	theta[2 * s + 1] = a.o[1];
}

// per-sample loss and gradient of (loss / S)
__global__ __launch_bounds__(256) void k_mlp_grad(const double *W, const double *X, const double *coef, int S, double *G,
                                                  double *L)
{
	__shared__ double w[MLP_NP];
	for (int e = threadIdx.x; e < MLP_NP; e += blockDim.x) w[e] = W[e];
	__syncthreads();
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= S) return;
	const double *x = X + (long)s * F;
	const double pr = coef[4 * s], pi = coef[4 * s + 1], c = coef[4 * s + 2], d = coef[4 * s + 3];
	fwd a;
	forward(w, x, a);
	const double tr = a.o[0], ti = a.o[1];
	const double inv = 1.0 / (d * S);
	L[s] = (d + 2 * ((tr * pr) - (ti * pi)) + ((tr * tr) + (ti * ti)) * c) * inv;
	const double go[2] = {(2 * pr + 2 * tr * c) * inv, (-2 * pi + 2 * ti * c) * inv};
	double *g = G + (long)s * MLP_NP;
	double gh2[H];
	for (int k = 0; k < H; k++) gh2[k] = 0;
	for (int i = 0; i < 2; i++) {
		g[OB3 + i] = go[i];
		for (int k = 0; k < H; k++) {
			g[OW3 + i * H + k] = go[i] * a.h2[k];
			gh2[k] += w[OW3 + i * H + k] * go[i];
		}
	}
	double gz2[H], gh1[H];
	for (int i = 0; i < H; i++) gz2[i] = gh2[i] * (1 - a.h2[i] * a.h2[i]);
	for (int k = 0; k < H; k++) gh1[k] = 0;
	for (int i = 0; i < H; i++) {
		g[OB2 + i] = gz2[i];
		for (int k = 0; k < H; k++) {
			g[OW2 + i * H + k] = gz2[i] * a.h1[k];
			gh1[k] += w[OW2 + i * H + k] * gz2[i];
		}
	}
	for (int i = 0; i < H; i++) {
		const double gz1 = gh1[i] * (1 - a.h1[i] * a.h1[i]);
		g[OB1 + i] = gz1;
		for (int k = 0; k < F; k++) g[OW1 + i * F + k] = gz1 * x[k];
	}
}

// grad[e] = sum_s G[s][e], loss = sum_s L[s], in sample order
__global__ __launch_bounds__(512) void k_mlp_reduce(const double *G, const double *L, int S, double *grad, double *loss)
{
	for (int e = threadIdx.x; e < MLP_NP; e += blockDim.x) {
		double acc = 0;
		for (int s = 0; s < S; s++) acc += G[(long)s * MLP_NP + e];
		grad[e] = acc;
	}
	if (threadIdx.x == 0) {
		double acc = 0;
		for (int s = 0; s < S; s++) acc += L[s];
		*loss = acc;
	}
}

__global__ void k_mlp_adam(double *W, double *m, double *v, const double *grad, double lr, double b1, double b2, double eps,
                           double c1, double c2)
{
	const int e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= MLP_NP) return;
	const double g = grad[e];
	m[e] = b1 * m[e] + (1 - b1) * g;
	v[e] = b2 * v[e] + (1 - b2) * g * g;
	const double mh = m[e] / c1, vh = v[e] / c2;
	W[e] -= lr * mh / (sqrt(vh) + eps);
}

template <typename T>
T *dal(size_t n)
{
	T *p = nullptr;
	HIPCHK(hipMalloc((void **)&p, (n ? n : 1) * sizeof(T)));
	return p;
}

}  // namespace

void mlp_forward_dev(hipStream_t s, const double *dW, const double *dX, int S, double *dtheta)
{
	if (S <= 0) return;
	hipLaunchKernelGGL(k_mlp_forward, dim3((S + 255) / 256), dim3(256), 0, s, dW, dX, S, dtheta);
}

void mlp_grad_dev(hipStream_t s, const double *dW, const double *dX, const double *dcoef, int S, double *dG, double *dL,
                  double *dgrad, double *dloss)
{
	hipLaunchKernelGGL(k_mlp_grad, dim3((S + 255) / 256), dim3(256), 0, s, dW, dX, dcoef, S, dG, dL);
	hipLaunchKernelGGL(k_mlp_reduce, dim3(1), dim3(512), 0, s, (const double *)dG, (const double *)dL, S, dgrad, dloss);
}

// initial parameters: uniform in +-sqrt(6 / (fan_in + fan_out)) from a splitmix64 stream, biases 0,
// the output layer scaled by 0.1 (a start near theta = 0, the method without the correction)
void mlp_init(double *W, unsigned long long seed)
{
	unsigned long long st = seed;
	auto next = [&st]() {
		st += 0x9E3779B97F4A7C15ull;
		unsigned long long z = st;
		z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
		z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
		return z ^ (z >> 31);
	};
	auto uni = [&]() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0) * 2 - 1; };
	for (int e = 0; e < MLP_NP; e++) W[e] = 0;
	const double a1 = std::sqrt(6.0 / (F + H)), a2 = std::sqrt(6.0 / (H + H)), a3 = std::sqrt(6.0 / (H + 2));
	for (int e = 0; e < H * F; e++) W[OW1 + e] = a1 * uni();
	for (int e = 0; e < H * H; e++) W[OW2 + e] = a2 * uni();
	for (int e = 0; e < 2 * H; e++) W[OW3 + e] = 0.1 * a3 * uni();
}

// Adam on the device over the whole sample set (full batch); loss_hist[t] the loss before step t
void mlp_fit_dev(hipStream_t s, const double *X, const double *coef, int S, int steps, double lr, double *W,
                 double *loss_hist)
{
	double *dW = dal<double>(MLP_NP), *dm = dal<double>(MLP_NP), *dv = dal<double>(MLP_NP), *dg = dal<double>(MLP_NP);
	double *dX = dal<double>((size_t)S * F), *dc = dal<double>((size_t)S * 4), *dG = dal<double>((size_t)S * MLP_NP);
	double *dL = dal<double>(S), *dloss = dal<double>((size_t)steps + 1);
	HIPCHK(hipMemcpyAsync(dW, W, MLP_NP * sizeof(double), hipMemcpyHostToDevice, s));
	HIPCHK(hipMemcpyAsync(dX, X, (size_t)S * F * sizeof(double), hipMemcpyHostToDevice, s));
	HIPCHK(hipMemcpyAsync(dc, coef, (size_t)S * 4 * sizeof(double), hipMemcpyHostToDevice, s));
	std::vector<double> zero(MLP_NP, 0.0);
	HIPCHK(hipMemcpyAsync(dm, zero.data(), MLP_NP * sizeof(double), hipMemcpyHostToDevice, s));
	HIPCHK(hipMemcpyAsync(dv, zero.data(), MLP_NP * sizeof(double), hipMemcpyHostToDevice, s));
	const double b1 = 0.9, b2 = 0.999;
	double p1 = 1, p2 = 1;
	for (int t = 0; t < steps; t++) {
		mlp_grad_dev(s, dW, dX, dc, S, dG, dL, dg, dloss + t);
		p1 *= b1;
		p2 *= b2;
		hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NP + 255) / 256), dim3(256), 0, s, dW, dm, dv, (const double *)dg, lr, b1, b2,
		                   1e-12, 1 - p1, 1 - p2);
	}
	mlp_grad_dev(s, dW, dX, dc, S, dG, dL, dg, dloss + steps);
	HIPCHK(hipMemcpyAsync(W, dW, MLP_NP * sizeof(double), hipMemcpyDeviceToHost, s));
	if (loss_hist) HIPCHK(hipMemcpyAsync(loss_hist, dloss, (size_t)(steps + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
	HIPCHK(hipStreamSynchronize(s));
	for (double *p : {dW, dm, dv, dg, dX, dc, dG, dL, dloss}) HIPCHK(hipFree(p));
}

// ------------------------------------------------------------------------------------------
extern "C" {

int peec_lpc_nparams(void)
{
	return MLP_NP;
}

int peec_lpc_nfeatures(void)
{
	return MLP_NF;
}

int peec_lpc_mlp_forward(const double *W, const double *X, int S, double *theta)
{
	if (S <= 0) return 1;
	double *dW = dal<double>(MLP_NP), *dX = dal<double>((size_t)S * F), *dt = dal<double>((size_t)2 * S);
	HIPCHK(hipMemcpy(dW, W, MLP_NP * sizeof(double), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(dX, X, (size_t)S * F * sizeof(double), hipMemcpyHostToDevice));
	mlp_forward_dev(nullptr, dW, dX, S, dt);
	HIPCHK(hipMemcpy(theta, dt, (size_t)2 * S * sizeof(double), hipMemcpyDeviceToHost));
	for (double *p : {dW, dX, dt}) HIPCHK(hipFree(p));
	return 0;
}

int peec_lpc_mlp_grad(const double *W, const double *X, const double *coef, int S, double *grad, double *loss)
{
	if (S <= 0) return 1;
	double *dW = dal<double>(MLP_NP), *dX = dal<double>((size_t)S * F), *dc = dal<double>((size_t)S * 4);
	double *dG = dal<double>((size_t)S * MLP_NP), *dL = dal<double>(S), *dg = dal<double>(MLP_NP), *dl = dal<double>(1);
	HIPCHK(hipMemcpy(dW, W, MLP_NP * sizeof(double), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(dX, X, (size_t)S * F * sizeof(double), hipMemcpyHostToDevice));
	HIPCHK(hipMemcpy(dc, coef, (size_t)S * 4 * sizeof(double), hipMemcpyHostToDevice));
	mlp_grad_dev(nullptr, dW, dX, dc, S, dG, dL, dg, dl);
	HIPCHK(hipMemcpy(grad, dg, MLP_NP * sizeof(double), hipMemcpyDeviceToHost));
	HIPCHK(hipMemcpy(loss, dl, sizeof(double), hipMemcpyDeviceToHost));
	for (double *p : {dW, dX, dc, dG, dL, dg, dl}) HIPCHK(hipFree(p));
	return 0;
}

int peec_lpc_fit(const double *X, const double *coef, int S, int steps, double lr, unsigned long long seed, double *W,
                 double *loss_hist)
{
	if ((S <= 0) || (steps < 0)) return 1;
	mlp_init(W, seed);
	mlp_fit_dev(nullptr, X, coef, S, steps, lr, W, loss_hist);
	return 0;
}

}  // extern "C"
