/*
 * skin_host.c — internal (skin-effect) impedance of a branch, a caller-side input of the MNA
 * stamps (only decks with skineffect = 1 use it).  Same model as the reference (OpenPEEC
 * src/skin.c): round wire Z = len k I0(ka) / (2 pi a sigma I1(ka)), k = (1 + j)/delta, with the
 * Bessel ratio from its power series for |ka| < 15 and the asymptotic series beyond; a
 * rectangular section blends the DC and skin-depth limits; volume cells keep R_dc.
 */
#include "peec.h"

#define SERIES_LIMIT 15.0
#define SERIES_TERMS 60

/* I0(z) / I1(z) for complex z */
static d_complex_t bessel_ratio(d_complex_t z)
{
	if (d_abs(z) < SERIES_LIMIT) {
		const d_complex_t z2 = d_mul(z, z);
		d_complex_t i0 = d_complex(1, 0), t = d_complex(1, 0);
		for (int k = 1; k <= SERIES_TERMS; k++) {
			t = d_rmul(1.0 / (4.0 * k * k), d_mul(t, z2));
			i0 = d_add(i0, t);
			if (d_abs(t) < 1e-17 * d_abs(i0)) break;
		}
		d_complex_t s = d_complex(1, 0);
		t = d_complex(1, 0);
		for (int k = 1; k <= SERIES_TERMS; k++) {
			t = d_rmul(1.0 / (4.0 * k * (k + 1)), d_mul(t, z2));
			s = d_add(s, t);
			if (d_abs(t) < 1e-17 * d_abs(s)) break;
		}
		return d_div(i0, d_mul(d_rmul(0.5, z), s));
	}
	/* 1 + 1/(2z) + 3/(8z^2) + 3/(8z^3) */
	const d_complex_t u = d_inv(z);
	const d_complex_t u2 = d_mul(u, u);
	const d_complex_t u3 = d_mul(u2, u);
	return d_add(d_add(d_complex(1, 0), d_rmul(0.5, u)), d_add(d_rmul(0.375, u2), d_rmul(0.375, u3)));
}

d_complex_t zint_round(double len, double a, double sigma, double freq)
{
	if ((sigma <= 0) || (len <= 0) || (a <= 0)) return d_complex(0, 0);
	const double rdc = len / (sigma * PI * a * a);
	const double omega = 2 * PI * freq;
	if (omega <= 0) return d_complex(rdc, 0);
	const double delta = sqrt(2 / (omega * MU0 * sigma));
	const double q = a / delta;
	const d_complex_t ratio = bessel_ratio(d_complex(q, q));
	const double kf = 1 / (2 * PI * a * delta * sigma);
	return d_rmul(len, d_mul(d_complex(kf, kf), ratio));
}

d_complex_t zint_bar(double len, double area, double perim, double sigma, double freq)
{
	if ((sigma <= 0) || (len <= 0) || (area <= 0) || (perim <= 0)) return d_complex(0, 0);
	const double rdc = len / (sigma * area);
	const double omega = 2 * PI * freq;
	if (omega <= 0) return d_complex(rdc, 0);
	const double delta = sqrt(2 / (omega * MU0 * sigma));
	const double rhf = len / (sigma * delta * perim);
	const double ldc = MU0 * len / (8 * PI);
	const double lhf = rhf / omega;
	const double r = sqrt((rdc * rdc) + (rhf * rhf));
	const double l = ldc * lhf / sqrt((ldc * ldc) + (lhf * lhf));
	return d_complex(r, omega * l);
}

d_complex_t zint_seg(const seg_t *s, double freq)
{
	if (s->diel) return d_complex(0, 0);
	if (s->vol) return d_complex(s->res, 0);
	return (s->shape == SHAPE_ROUND) ? zint_round(s->len, s->radius, s->sigma, freq)
	                                 : zint_bar(s->len, s->area, s->perim, s->sigma, freq);
}
