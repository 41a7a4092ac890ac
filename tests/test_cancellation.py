"""Closed forms with cancellation, pinned against an extended-precision evaluation.

Two of the reference's static kernels are signed sums of large terms: the parallel-filament form
(4 x fpar, partial.c:131-142) and the Hoer-Love bar form (64 x hl_F, volume.c:44-155).  Their
FP64 results carry the libm's ulp errors multiplied by the cancellation factor
kappa = sum|terms| / |sum|, so two correct FP64 implementations on different libms (glibc on the
host, ocml on gfx950) disagree by about kappa * eps.  Here the same formulas are evaluated in
x87 extended precision (numpy longdouble, eps 1.1e-19) from the same FP64 cell data, and both
the oracle (glibc, bitwise equal to the reference) and the device (ocml) must lie within
8 * kappa * eps of it: that is what justifies the per-deck tolerances of test_gpu_fill.py on
decks made of such pairs, and would catch a wrong term (whose error is not proportional to
kappa * eps).
"""
import ctypes as C

import numpy as np
import pytest

from conftest import deck_text
from openpeec_amd import Peec, abi
from oracle_py import oracle_lib

LD = np.longdouble
EPS = 2.220446049250313e-16
ORC = oracle_lib()
ORC.orc_neumann_pair.argtypes = [C.POINTER(abi.seg_t), C.POINTER(abi.seg_t), C.c_double, C.c_double]
ORC.orc_neumann_pair.restype = C.c_double
ORC.orc_bar_pair.argtypes = [C.POINTER(abi.seg_t), C.POINTER(abi.seg_t)]
ORC.orc_bar_pair.restype = C.c_double
ORC.orc_bar_use_hl.argtypes = [C.POINTER(abi.seg_t), C.POINTER(abi.seg_t)]
ORC.orc_bar_use_hl.restype = C.c_int


def _v(a):
    return np.array([LD(a[0]), LD(a[1]), LD(a[2])])


def parallel_pair_ld(s1, s2, a1, a2):
    """4-term parallel-filament form in extended precision; returns (value, kappa)."""
    aeff = LD(0.5) * (LD(a1) + LD(a2))
    x1a, x2a, x1b, x2b = _v(s1.x1), _v(s1.x2), _v(s2.x1), _v(s2.x2)
    L = LD(s1.len)
    t1 = (x2a - x1a) / L
    r1, r2 = x1b - x1a, x2b - x1a
    lo, hi = np.dot(r1, t1), np.dot(r2, t1)
    if lo > hi:
        lo, hi = hi, lo
    along = np.dot(r1, t1)
    perp = r1 - along * t1
    d = np.sqrt(np.dot(perp, perp) + aeff * aeff)

    def fil(x):
        return x * np.arcsinh(x / d) - np.sqrt(x * x + d * d), abs(x * np.arcsinh(x / d)) + np.sqrt(x * x + d * d)

    terms = [fil(L - lo), fil(L - hi), fil(LD(0) - lo), fil(LD(0) - hi)]
    val = terms[0][0] - terms[1][0] - terms[2][0] + terms[3][0]
    mag = sum(t[1] for t in terms)
    return float(val), float(mag / abs(val))


def hoer_love_ld(x, y, z):
    x2, y2, z2 = x * x, y * y, z * z
    R = np.sqrt(x2 + y2 + z2)
    yz, xz, xy = y2 + z2, x2 + z2, x2 + y2
    parts = []
    if x != 0 and yz > 0:
        parts.append(((y2 * z2 / 4) - (y2 * y2 / 24) - (z2 * z2 / 24)) * x * np.arcsinh(x / np.sqrt(yz)))
    if y != 0 and xz > 0:
        parts.append(((x2 * z2 / 4) - (x2 * x2 / 24) - (z2 * z2 / 24)) * y * np.arcsinh(y / np.sqrt(xz)))
    if z != 0 and xy > 0:
        parts.append(((x2 * y2 / 4) - (x2 * x2 / 24) - (y2 * y2 / 24)) * z * np.arcsinh(z / np.sqrt(xy)))
    parts.append(((x2 * x2) + (y2 * y2) + (z2 * z2) - 3 * ((x2 * y2) + (y2 * z2) + (z2 * x2))) * R / 60)
    if z != 0:
        parts.append(-(x * y * z * z2 / 6) * np.arctan((x * y) / (z * R)))
    if y != 0:
        parts.append(-(x * y * y2 * z / 6) * np.arctan((x * z) / (y * R)))
    if x != 0:
        parts.append(-(x * x2 * y * z / 6) * np.arctan((y * z) / (x * R)))
    return sum(parts), sum(abs(p) for p in parts)


def bar_pair_ld(s1, s2):
    """64-term Hoer-Love sum in extended precision; returns (value, kappa)."""
    def frame(s):
        x1, x2 = _v(s.x1), _v(s.x2)
        t = (x2 - x1) / LD(s.len)
        w = _v(s.wv)
        n = np.array([t[1] * w[2] - t[2] * w[1], t[2] * w[0] - t[0] * w[2], t[0] * w[1] - t[1] * w[0]])
        return t, w, n, LD(0.5) * (x1 + x2)
    t1, w1, n1, m1 = frame(s1)
    _, _, _, m2 = frame(s2)
    a, b, c = LD(s1.thick), LD(s1.wid), LD(s1.len)
    d, e, f = LD(s2.thick), LD(s2.wid), LD(s2.len)
    rc = m2 - m1
    E = np.dot(rc, n1) + (a - d) / 2
    P = np.dot(rc, w1) + (b - e) / 2
    L3 = np.dot(rc, t1) + (c - f) / 2
    q = [E - a, E + d - a, E + d, E]
    r = [P - b, P + e - b, P + e, P]
    s = [L3 - c, L3 + f - c, L3 + f, L3]
    acc, mag = LD(0), LD(0)
    for i in range(4):
        for j in range(4):
            for k in range(4):
                v, m = hoer_love_ld(q[i], r[j], s[k])
                acc += (1 if (i + j + k) % 2 == 0 else -1) * v
                mag += m
    den = a * b * d * e
    return float(acc / den), float(mag / abs(acc))


def _parallel_pairs(m, limit=40):
    segs = m.p.seg
    out = []
    for i in range(m.nseg):
        for j in range(i + 1, m.nseg):
            si, sj = segs[i], segs[j]
            ti = np.subtract(si.x2, si.x1) / si.len
            tj = np.subtract(sj.x2, sj.x1) / sj.len
            if np.linalg.norm(np.cross(ti, tj)) < 1e-9:
                out.append((i, j))
    step = max(1, len(out) // limit)
    return out[::step][:limit]


def _hl_pairs(m, limit=40):
    segs = m.p.seg
    out = [(i, j) for i in range(m.nseg) for j in range(i, m.nseg)
           if ORC.orc_bar_use_hl(C.byref(segs[i]), C.byref(segs[j]))]
    step = max(1, len(out) // limit)
    return out[::step][:limit]


def test_parallel_filament_oracle_within_cancellation_bound():
    m = Peec.from_deck_text(deck_text("dipole_halfwave.peec"))
    m.wire_build()
    pairs = _parallel_pairs(m)
    assert pairs
    for i, j in pairs:
        si, sj = m.p.seg[i], m.p.seg[j]
        ref, kappa = parallel_pair_ld(si, sj, si.aL, sj.aL)
        got = ORC.orc_neumann_pair(C.byref(si), C.byref(sj), si.aL, sj.aL)
        assert abs(got - ref) <= 8 * kappa * EPS * abs(ref) + 1e-300, (i, j, got, ref, kappa)
    m.close()


def test_hoer_love_oracle_within_cancellation_bound():
    m = Peec.from_deck_text(deck_text("plate_thick.peec"))
    m.wire_build()
    pairs = _hl_pairs(m)
    assert pairs
    worst = 0.0
    for i, j in pairs:
        si, sj = m.p.seg[i], m.p.seg[j]
        ref, kappa = bar_pair_ld(si, sj)
        got = ORC.orc_bar_pair(C.byref(si), C.byref(sj))
        assert abs(got - ref) <= 8 * kappa * EPS * abs(ref), (i, j, got, ref, kappa)
        worst = max(worst, kappa)
    assert worst > 1e2   # these pairs do cancel: the bound is doing work
    m.close()


def test_wrong_term_is_caught():
    """A planted mistake (one Hoer-Love corner with the wrong sign) breaks the bound."""
    m = Peec.from_deck_text(deck_text("plate_thick.peec"))
    m.wire_build()
    i, j = _hl_pairs(m, 3)[1]
    si, sj = m.p.seg[i], m.p.seg[j]
    ref, kappa = bar_pair_ld(si, sj)
    got = ORC.orc_bar_pair(C.byref(si), C.byref(sj))
    x = LD(1.7e-3)
    planted = got + 2 * float(hoer_love_ld(x, x, x)[0]) / (si.thick * si.wid * sj.thick * sj.wid)
    assert abs(planted - ref) > 8 * kappa * EPS * abs(ref)
    m.close()


@pytest.mark.gpu
def test_device_closed_forms_within_cancellation_bound(lib):
    m = Peec.from_deck_text(deck_text("dipole_halfwave.peec"), lib)
    m.wire_build()
    for i, j in _parallel_pairs(m):
        si, sj = m.p.seg[i], m.p.seg[j]
        ref, kappa = parallel_pair_ld(si, sj, si.aL, sj.aL)
        got = lib.neumann_pair(C.byref(si), C.byref(sj), si.aL, sj.aL)
        assert abs(got - ref) <= 8 * kappa * EPS * abs(ref) + 1e-300, ("fpar", i, j, got, ref, kappa)
    m.close()
    m = Peec.from_deck_text(deck_text("plate_thick.peec"), lib)
    m.wire_build()
    for i, j in _hl_pairs(m):
        si, sj = m.p.seg[i], m.p.seg[j]
        ref, kappa = bar_pair_ld(si, sj)
        got = lib.bar_pair(C.byref(si), C.byref(sj))
        assert abs(got - ref) <= 8 * kappa * EPS * abs(ref), ("hl", i, j, got, ref, kappa)
    m.close()
