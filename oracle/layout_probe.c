/*
 * layout_probe.c — prints sizeof/offsetof of the ABI structs, one "name value" per line.
 * Built once against the reference header (oracle/Makefile: _ref/layout_ref) and once
 * against include/peec.h (tests/test_abi.py); the two listings must be identical for the
 * product to be a drop-in for callers compiled against the reference.
 */
#include <stddef.h>
#include "peec.h"

#define SZ(T) printf("sizeof(" #T ") %zu\n", sizeof(T))
#define OF(T, m) printf("offsetof(" #T "," #m ") %zu\n", offsetof(T, m))

int main(void)
{
	SZ(d_complex_t); SZ(rc_t); SZ(ind_t); SZ(mut_t); SZ(src_t); SZ(port_t);
	SZ(wire_t); SZ(panel_t); SZ(plate_t); SZ(diel_t); SZ(seg_t); SZ(peec_t);
	OF(seg_t, n1); OF(seg_t, x1); OF(seg_t, len); OF(seg_t, shape); OF(seg_t, wid);
	OF(seg_t, wv); OF(seg_t, vol); OF(seg_t, npv); OF(seg_t, papex); OF(seg_t, pv);
	OF(seg_t, aL); OF(seg_t, aP); OF(seg_t, area); OF(seg_t, res); OF(seg_t, diel);
	OF(seg_t, cexc); OF(seg_t, frelax);
	OF(peec_t, nres); OF(peec_t, plate); OF(peec_t, ncxyz); OF(peec_t, nseg); OF(peec_t, seg);
	OF(peec_t, lp); OF(peec_t, lpwarn); OF(peec_t, nchg); OF(peec_t, chg); OF(peec_t, cellid);
	OF(peec_t, carea); OF(peec_t, cpt); OF(peec_t, cmat); OF(peec_t, clogged); OF(peec_t, gxyz);
	OF(peec_t, maxid); OF(peec_t, nodemap); OF(peec_t, offL); OF(peec_t, offQ);
	OF(peec_t, nunknown); OF(peec_t, f0); OF(peec_t, nfreq); OF(peec_t, nodetol);
	OF(peec_t, gp); OF(peec_t, gpz); OF(peec_t, pwamp); OF(peec_t, tranatt); OF(peec_t, grading);
	OF(peec_t, compress); OF(peec_t, ctol); OF(peec_t, hlp); OF(peec_t, zin); OF(peec_t, zmat);
	OF(peec_t, smat); OF(peec_t, segi); OF(peec_t, cellq); OF(peec_t, voc); OF(peec_t, segipw);
	OF(panel_t, grading); OF(diel_t, ndivt); OF(plate_t, ndivt); OF(wire_t, ndiv);
	return 0;
}
