/*
 * complex.h — interleaved FP64 complex scalar used across the PEEC C ABI.
 *
 * Layout and semantics match the reference type d_complex_t {double r, i}
 * (include/complex.h:13 of OpenPEEC): 16 bytes, real part first, so every
 * matrix in the ABI is a flat row-major array of (re, im) pairs.
 * Division and inversion return 0 for an exactly-zero divisor
 * (reference complex.h:42,55); the device code relies on the same rule.
 */
#ifndef PEEC_MI355X_COMPLEX_H
#define PEEC_MI355X_COMPLEX_H

#include <math.h>

typedef struct { double r, i; } d_complex_t;

static inline d_complex_t d_complex(double r, double i)
{
	d_complex_t z;
	z.r = r;
	z.i = i;
	return z;
}

static inline d_complex_t d_add(d_complex_t a, d_complex_t b) { return d_complex(a.r + b.r, a.i + b.i); }
static inline d_complex_t d_sub(d_complex_t a, d_complex_t b) { return d_complex(a.r - b.r, a.i - b.i); }

static inline d_complex_t d_mul(d_complex_t a, d_complex_t b)
{
	return d_complex((a.r * b.r) - (a.i * b.i), (a.r * b.i) + (a.i * b.r));
}

static inline d_complex_t d_div(d_complex_t a, d_complex_t b)
{
	if ((fabs(b.r) <= 0) && (fabs(b.i) <= 0)) return d_complex(0, 0);
	const double den = (b.r * b.r) + (b.i * b.i);
	return d_complex(((a.r * b.r) + (a.i * b.i)) / den, ((a.i * b.r) - (a.r * b.i)) / den);
}

static inline d_complex_t d_rmul(double s, d_complex_t z) { return d_complex(s * z.r, s * z.i); }

static inline d_complex_t d_inv(d_complex_t z)
{
	const double n2 = (z.r * z.r) + (z.i * z.i);
	if (n2 <= 0) return d_complex(0, 0);
	return d_complex(z.r / n2, -z.i / n2);
}

static inline d_complex_t d_conj(d_complex_t z) { return d_complex(z.r, -z.i); }
static inline double d_abs(d_complex_t z) { return sqrt((z.r * z.r) + (z.i * z.i)); }
static inline double d_norm(d_complex_t z) { return (z.r * z.r) + (z.i * z.i); }

/* r * exp(j * deg), deg in degrees */
static inline d_complex_t d_polar_deg(double r, double deg)
{
	const double rad = deg * atan(1.0) / 45.0;
	return d_complex(r * cos(rad), r * sin(rad));
}

#endif
