// postproc.hip — the sweep's post-processing on the device: the far field from port #1's branch
// currents (output_far, OpenPEEC src/farfield.c:42-182) and the transient waveforms synthesised
// from the frequency responses (output_tran, src/transient.c:92-177).  The heavy sums run on the
// GPU; the files, the power integrals and the log lines are written by the host in the
// reference's formats and orders.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <vector>
#include "../peec_internal.h"
#include "hip_check.hpp"

namespace {

// One current cell of the radiation sum: unit direction t, midpoint c, length (the ground
// plane's image is formed in the kernel).
struct radcell {
	double t[3], c[3], len;
};

__device__ __forceinline__ double sinc_d(double u)
{
	return (fabs(u) < 1e-8) ? 1.0 : (sin(u) / u);
}

// Radiation vector of one direction per wave: N = sum_m I_m len_m t_m sinc(k len_m (r.t_m) / 2)
// exp(+j k r.c_m), plus the mirrored cells carrying -I when gp; lanes stride over the cells,
// a fixed butterfly adds the lane sums.  Output per direction: r E_theta and r E_phi without
// exp(-jkr)/r, i.e. -j (omega mu0 / 4 pi) (N . theta^), (N . phi^).
__global__ __launch_bounds__(256) void k_radiation(const radcell *__restrict__ cells, const double2 *__restrict__ cur, int ncell,
                                                   int gp, double gpz, double kw, double coef, int nth, int nph, double thmax,
                                                   double2 *__restrict__ eth, double2 *__restrict__ eph)
{
	const int dir = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
	const int ndir = (nth + 1) * nph;
	if (dir >= ndir) return;   // whole waves leave together
	const int it = dir / nph, ip = dir - it * nph;
	const double th = thmax * it / nth, ph = 2 * PI * ip / nph;
	const double st = sin(th), ct = cos(th), cp = cos(ph), sp = sin(ph);
	const double r[3] = {st * cp, st * sp, ct};
	double nr[3] = {0, 0, 0}, ni[3] = {0, 0, 0};
	for (int pass = 0; pass < (gp ? 2 : 1); pass++) {
		for (int m = lane; m < ncell; m += 64) {
			const radcell e = cells[m];
			double t[3] = {e.t[0], e.t[1], e.t[2]}, c[3] = {e.c[0], e.c[1], e.c[2]};
			double sgn = 1;
			if (pass == 1) {   // image: geometric mirror in z, current -I
				t[2] = -t[2];
				c[2] = (2 * gpz) - c[2];
				sgn = -1;
			}
			const double rt = (r[0] * t[0]) + (r[1] * t[1]) + (r[2] * t[2]);
			const double rc = (r[0] * c[0]) + (r[1] * c[1]) + (r[2] * c[2]);
			const double amp = sgn * e.len * sinc_d(0.5 * kw * e.len * rt);
			double sn, cs;
			sincos(kw * rc, &sn, &cs);
			const double2 I = cur[m];
			const double wr = amp * ((I.x * cs) - (I.y * sn)), wi = amp * ((I.x * sn) + (I.y * cs));
			for (int q = 0; q < 3; q++) {
				nr[q] += t[q] * wr;
				ni[q] += t[q] * wi;
			}
		}
	}
	for (int off = 32; off > 0; off >>= 1)
		for (int q = 0; q < 3; q++) {
			nr[q] += __shfl_xor(nr[q], off);
			ni[q] += __shfl_xor(ni[q], off);
		}
	if (lane == 0) {
		const double th_[3] = {ct * cp, ct * sp, -st}, ph_[3] = {-sp, cp, 0};
		double tr = 0, ti = 0, pr = 0, pi = 0;
		for (int q = 0; q < 3; q++) {
			tr += th_[q] * nr[q];
			ti += th_[q] * ni[q];
			pr += ph_[q] * nr[q];
			pi += ph_[q] * ni[q];
		}
		eth[dir] = make_double2(coef * ti, -coef * tr);
		eph[dir] = make_double2(coef * pi, -coef * pr);
	}
}

// y[n] = df (g0 + 2 sum_k Re(g_k exp(j 2 pi f_k t_n))), t_n = n dt: one thread per sample, the
// frequencies summed in order
__global__ void k_synth(const double2 *__restrict__ g, const double *__restrict__ fk, int nf, double g0, double df, double dt,
                        int m, double *__restrict__ y)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= m) return;
	const double t = n * dt;
	double sum = g0;
	for (int k = 0; k < nf; k++) {
		double sn, cs;
		sincos(2 * PI * fk[k] * t, &sn, &cs);
		sum += 2 * ((g[k].x * cs) - (g[k].y * sn));
	}
	y[n] = df * sum;
}

template <typename T>
T *dev_upload(const std::vector<T> &h)
{
	T *d = nullptr;
	HIPCHK(hipMalloc((void **)&d, (h.empty() ? 1 : h.size()) * sizeof(T)));
	if (!h.empty()) HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
	return d;
}

}  // namespace

extern "C" {

// Far field of port #1's currents (farfield.c): r E_theta, r E_phi on a (ffnth + 1) x ffnph grid
// of directions (theta to 90 deg over a ground plane, else 180), directivity D = 4 pi U / Prad
// and gain G = 4 pi U / Pin (Pin = Re Zin / 2 for 1 A), Prad by the trapezoidal rule in theta and
// the periodic rectangle rule in phi.  Writes fn (far.csv) and one log line per frequency.
int output_far(const peec_t *p, const char *fn, FILE *fp_log)
{
	if ((p->ffnth <= 0) || (p->ffnph <= 0)) return 0;
	if ((p->segi == NULL) || (p->nseg <= 0) || (p->nport <= 0)) return 0;
	if (!pgpu_device_ok()) pgpu_fatal("no HIP device");
	const int nth = p->ffnth, nph = p->ffnph, ndir = (nth + 1) * nph, ns = p->nseg;
	const double thmax = p->gp ? (PI / 2) : PI;
	const double eta0 = MU0 * C0;
	FILE *fp = fopen(fn, "w");
	if (fp == NULL) {
		printf("*** file %s open error.\n", fn);
		return 1;
	}
	if (!p->retardation && fp_log)
		fprintf(fp_log, "*** warning : farfield with retardation = 0 (quasi-static currents do not radiate; efficiency is not meaningful)\n");
	std::vector<radcell> cells(ns);
	for (int m = 0; m < ns; m++) {
		const seg_t *s = &p->seg[m];
		for (int q = 0; q < 3; q++) {
			cells[m].t[q] = (s->x2[q] - s->x1[q]) / s->len;
			cells[m].c[q] = 0.5 * (s->x1[q] + s->x2[q]);
		}
		cells[m].len = s->len;
	}
	radcell *dcells = dev_upload(cells);
	double2 *dcur = nullptr, *deth = nullptr, *deph = nullptr;
	HIPCHK(hipMalloc((void **)&dcur, (size_t)ns * sizeof(double2)));
	HIPCHK(hipMalloc((void **)&deth, (size_t)ndir * sizeof(double2)));
	HIPCHK(hipMalloc((void **)&deph, (size_t)ndir * sizeof(double2)));
	std::vector<double2> eth(ndir), eph(ndir);
	std::vector<double> uu(ndir);
	fprintf(fp, "frequency[Hz],theta[deg],phi[deg],rEtheta_real[V],rEtheta_imag[V],rEphi_real[V],rEphi_imag[V],D[dBi],G[dBi]\n");
	for (int i = 0; i < p->nfreq; i++) {
		const double f = freq_at(p, i);
		const double kw = 2 * PI * f / C0;
		const double coef = 2 * PI * f * MU0 / (4 * PI);
		HIPCHK(hipMemcpy(dcur, &p->segi[DIDX(p, i, 0, 0)], (size_t)ns * sizeof(double2), hipMemcpyHostToDevice));
		hipLaunchKernelGGL(k_radiation, dim3((unsigned)((ndir + 3) / 4)), dim3(256), 0, nullptr, (const radcell *)dcells,
		                   (const double2 *)dcur, ns, p->gp, p->gpz, kw, coef, nth, nph, thmax, deth, deph);
		HIPCHK(hipMemcpy(eth.data(), deth, (size_t)ndir * sizeof(double2), hipMemcpyDeviceToHost));
		HIPCHK(hipMemcpy(eph.data(), deph, (size_t)ndir * sizeof(double2), hipMemcpyDeviceToHost));
		int imax = 0;
		for (int d = 0; d < ndir; d++) {
			uu[d] = ((eth[d].x * eth[d].x) + (eth[d].y * eth[d].y) + (eph[d].x * eph[d].x) + (eph[d].y * eph[d].y)) / (2 * eta0);
			if (uu[d] > uu[imax]) imax = d;
		}
		double prad = 0;
		for (int it = 0; it <= nth; it++) {
			double ring = 0;
			for (int ip = 0; ip < nph; ip++) ring += uu[(size_t)it * nph + ip];
			prad += (((it == 0) || (it == nth)) ? 0.5 : 1) * sin(thmax * it / nth) * ring;
		}
		prad *= (thmax / nth) * (2 * PI / nph);
		const double pin = 0.5 * p->zin[i].r;
		for (int d = 0; d < ndir; d++) {
			const int it = d / nph, ip = d % nph;
			const double dd = (prad > 0) ? (4 * PI * uu[d] / prad) : 0;
			const double gg = (pin > 0) ? (4 * PI * uu[d] / pin) : 0;
			fprintf(fp, "%.9e,%.4f,%.4f,%.9e,%.9e,%.9e,%.9e,%.4f,%.4f\n", f, 180 / PI * thmax * it / nth, 360.0 * ip / nph, eth[d].x,
			        eth[d].y, eph[d].x, eph[d].y, (dd > 1e-30) ? (10 * log10(dd)) : -600, (gg > 1e-30) ? (10 * log10(gg)) : -600);
		}
		const double dmax = (prad > 0) ? (4 * PI * uu[imax] / prad) : 0;
		if (fp_log)
			fprintf(fp_log, "farfield : f = %.5e Hz, Prad = %.5e W, Pin = %.5e W, eff = %.5f, max D = %.3f dBi at (theta, phi) = (%.1f, %.1f) deg\n",
			        f, prad, pin, (pin > 0) ? (prad / pin) : 0, (dmax > 1e-30) ? (10 * log10(dmax)) : -600,
			        180 / PI * thmax * (imax / nph) / nth, 360.0 * (imax % nph) / nph);
	}
	if (fp_log) fflush(fp_log);
	void *bb[] = {dcells, dcur, deth, deph};
	for (void *x : bb) HIPCHK(hipFree(x));
	fclose(fp);
	return 0;
}

// Transient response (transient.c): the sweep must be f_k = k df (k = 1..N, linear); a Gaussian
// pulse whose spectrum is down tranatt dB at f_N drives every S entry and the plane wave's
// port voltages; y(t) = df (H(0) X(0) + 2 Re sum_k H(f_k) X(f_k) exp(j 2 pi f_k t)) on 2N
// samples of dt = 1 / (2 f_N), H(0) extrapolated from the two lowest points (real at DC).
int output_tran(const peec_t *p, const char *fn, FILE *fp_log)
{
	if (!p->tran) return 0;
	const char *bad = nullptr;
	const int nf = p->nfreq;
	const double df_ = (nf > 1) ? (p->f1 - p->f0) / (nf - 1) : 0;
	if (p->flog) bad = "*** transient : logarithmic sweep is not supported (use a linear sweep)";
	else if (nf < 3) bad = "*** transient : at least 3 frequency points are required";
	if (bad) {
		printf("%s\n", bad);
		if (fp_log) fprintf(fp_log, "%s\n", p->flog ? "*** transient : logarithmic sweep is not supported" : bad);
		return 1;
	}
	if (fabs(p->f0 - df_) > 1e-6 * df_) {
		printf("*** transient : sweep must be f_k = k df (start = spacing); f0 = %.6e, df = %.6e\n", p->f0, df_);
		if (fp_log) fprintf(fp_log, "*** transient : sweep must be f_k = k df (start = spacing); f0 = %.6e, df = %.6e\n", p->f0, df_);
		return 1;
	}
	if (!pgpu_device_ok()) pgpu_fatal("no HIP device");
	const double df = p->f0, fmax = p->f1, dt = 1 / (2 * fmax);
	const int m = 2 * nf;
	const double sigma = sqrt(p->tranatt * log(10.0) / (20 * 2 * PI * PI)) / fmax;
	const double t0 = 4 * sigma, x0 = sigma * sqrt(2 * PI);
	std::vector<d_complex_t> xs(nf);
	std::vector<double> fk(nf);
	for (int k = 0; k < nf; k++) {
		fk[k] = freq_at(p, k);
		const double a = x0 * exp(-2 * PI * PI * sigma * sigma * fk[k] * fk[k]);
		const double ph = -2 * PI * fk[k] * t0;
		xs[k] = d_complex(a * cos(ph), a * sin(ph));
	}
	FILE *fp = fopen(fn, "w");
	if (fp == NULL) {
		printf("*** file %s open error.\n", fn);
		return 1;
	}
	fprintf(fp, "type,i,j,time[s],value\n");
	double *dfk = dev_upload(fk);
	double2 *dg = nullptr;
	double *dy = nullptr;
	HIPCHK(hipMalloc((void **)&dg, (size_t)nf * sizeof(double2)));
	HIPCHK(hipMalloc((void **)&dy, (size_t)m * sizeof(double)));
	std::vector<d_complex_t> g(nf);
	std::vector<double> y(m);
	// one waveform: h NULL is the pulse itself (H = 1)
	auto series = [&](const char *type, int i, int j, const d_complex_t *h, size_t stride) {
		const double h0 = (h == nullptr) ? 1.0 : ((2 * h[0].r) - h[stride].r);
		for (int k = 0; k < nf; k++) g[k] = (h == nullptr) ? xs[k] : d_mul(h[(size_t)k * stride], xs[k]);
		HIPCHK(hipMemcpy(dg, g.data(), (size_t)nf * sizeof(double2), hipMemcpyHostToDevice));
		hipLaunchKernelGGL(k_synth, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, nullptr, (const double2 *)dg,
		                   (const double *)dfk, nf, h0 * x0, df, dt, m, dy);
		HIPCHK(hipMemcpy(y.data(), dy, (size_t)m * sizeof(double), hipMemcpyDeviceToHost));
		for (int n = 0; n < m; n++) fprintf(fp, "%s,%d,%d,%.9e,%.9e\n", type, i, j, n * dt, y[n]);
	};
	series("X", 0, 0, nullptr, 0);
	if (p->smat != NULL)
		for (int e = 0; e < p->nport * p->nport; e++) {
			const int i = e / p->nport, j = e % p->nport;
			series("S", i + 1, j + 1, &p->smat[ZIDX(p, 0, i, j)], (size_t)p->nport * p->nport);
		}
	if (p->voc != NULL)
		for (int i = 0; i < p->nport; i++) series("V", i + 1, 0, &p->voc[(size_t)i * nf], 1);
	if (fp_log) {
		fprintf(fp_log, "transient : %d samples, dt = %.5e s, span = %.5e s, gaussian sigma = %.5e s (t0 = %.5e s, -%.1f dB at %.5e Hz)\n",
		        m, dt, m * dt, sigma, t0, p->tranatt, fmax);
		fflush(fp_log);
	}
	void *bb[] = {dfk, dg, dy};
	for (void *x : bb) HIPCHK(hipFree(x));
	fclose(fp);
	return 0;
}

}  // extern "C"
