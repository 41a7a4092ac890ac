"""Pinning the oracle (oracle/liboracle.so, the CPU restatement of the reference's fill + dense
solve) — CPU only.

1. Golden vectors written by the reference's own functions (tests/golden/ref_vectors.npz):
   Lp entries, C = P^-1 and Z(f) sweeps must agree bit for bit (same arithmetic, same libm).
2. Live bitwise comparison with the reference build when present.
3. Mathematics, chosen so that a dropped term, wrong sign or transposed operand fails:
   closed forms against brute-force quadrature (neumann_self), the centre potential of the unit
   square (4 ln(1 + sqrt 2)), polygon edge sum = rectangle closed form on a rectangle, the
   thin-wire sum rule (collinear segments add up to the whole), Hoer-Love additivity, LU
   against LAPACK.
4. The reference's own known-answer checks (data/sample/peec_check.sh): wire L, wire Ctotal,
   half-wave dipole resonance and Rin, monopole = dipole / 2, plate capacitance (Richardson).
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import integrate

from conftest import DECKS, deck_text
from openpeec_amd import abi, decks
from oracle_py import oracle_c, oracle_lib, oracle_lp, oracle_peec, oracle_zmat
from refharness import RefPeec, ref_available

GOLD = np.load(DECKS.parent / "ref_vectors.npz")
ORC = oracle_lib()


def _retarded(text):
    return text.replace("OpenPEEC 1 0", "OpenPEEC 1 0\nretardation = 1", 1) if "retardation" not in text else text


def _kw(f):
    return 2 * math.pi * f / 2.99792458e8


# ------------------------------------------------------------------ 1. golden vectors
LP_KEYS = sorted(k for k in GOLD.files if k.startswith("lp/"))


@pytest.mark.parametrize("key", LP_KEYS)
def test_lp_matches_reference_golden(key):
    _, name, fs = key.split("/")
    m = oracle_peec(_retarded(deck_text(name + ".peec")))
    lp = oracle_lp(m, _kw(float(fs)))
    want = GOLD[key]
    if f"lpidx/{name}/{fs}" in GOLD.files:
        ij = GOLD[f"lpidx/{name}/{fs}"]
        got = lp[ij[:, 0], ij[:, 1]]
    else:
        got = lp
    assert np.array_equal(got.view(np.float64), want.view(np.float64))
    m.close()


@pytest.mark.parametrize("key", sorted(k for k in GOLD.files if k.startswith("c/")))
def test_c_matches_reference_golden(key):
    _, name, fs = key.split("/")
    m = oracle_peec(_retarded(deck_text(name + ".peec")))
    c = oracle_c(m, _kw(float(fs)))
    assert np.array_equal(c.view(np.float64), GOLD[key].view(np.float64))
    m.close()


SWEEPS = {
    "dipole_c1": lambda: decks.dipole_deck(),
    "monopole_gp": lambda: deck_text("monopole_gp.peec"),
    "dipole_halfwave": lambda: deck_text("dipole_halfwave.peec"),
    "strip_dipole_ret": lambda: deck_text("strip_dipole_ret.peec"),
    "twoport_dist": lambda: deck_text("twoport_dist.peec"),
    "plate_12": lambda: decks.plate_deck(n=12, ndiv=4),
    "wire_single": lambda: deck_text("wire_single.peec"),
}


@pytest.mark.parametrize("name", sorted(SWEEPS))
def test_sweep_matches_reference_golden(name):
    z = oracle_zmat(SWEEPS[name]())
    assert np.array_equal(z.view(np.float64), GOLD[f"zmat/{name}"].view(np.float64))


# ------------------------------------------------------------------ 2. live reference
@pytest.mark.skipif(not ref_available(), reason="reference build not present")
@pytest.mark.parametrize("name", ["diel_pp", "disk_annulus", "plate_thick", "quad_taper"])
def test_full_fill_bitwise_live(name):
    text = _retarded(deck_text(name + ".peec"))
    r = RefPeec(text)
    r.wire_build()
    m = oracle_peec(text)
    for f in (0.0, 1.1e9):
        assert np.array_equal(oracle_lp(m, _kw(f)).view(np.float64), r.lp_fill(f).view(np.float64))
        if r.p.capacitance and r.p.nchg > 0:
            assert np.array_equal(oracle_c(m, _kw(f)).view(np.float64), r.pot_fill(f).view(np.float64))
    r.close()
    m.close()


# ------------------------------------------------------------------ 3. mathematics
def _ribbon(x1, x2, wid, wv, npv=0, pv=None):
    s = abi.seg_t()
    for k in range(3):
        s.x1[k], s.x2[k], s.wv[k] = x1[k], x2[k], wv[k]
    s.len = math.dist(x1, x2)
    s.wid = wid
    s.npv = npv
    if pv is not None:
        for i, v in enumerate(np.asarray(pv).ravel()):
            s.pv[i] = v
        for k in range(3):
            s.papex[k] = pv[0][k]
    return s


@pytest.mark.parametrize("l,a", [(1.0, 1e-3), (0.01, 1e-3), (0.3, 0.05)])
def test_neumann_self_closed_form_vs_quadrature(l, a):
    # int_0^l int_0^l dx dy / sqrt((x-y)^2 + a^2) = 2 int_0^l (l - u) / sqrt(u^2 + a^2) du
    brute, _ = integrate.quad(lambda u: 2 * (l - u) / math.sqrt(u * u + a * a), 0, l, epsabs=0, epsrel=1e-13, limit=400)
    assert abs(ORC.orc_neumann_self(l, a) - brute) <= 1e-12 * brute


def test_unit_square_centre_potential():
    s = _ribbon((-0.5, 0, 0), (0.5, 0, 0), 1.0, (0, 1, 0))
    pt = (C.c_double * 3)(0, 0, 0)
    assert abs(ORC.orc_rect_potential(C.byref(s), pt) - 4 * math.log(1 + math.sqrt(2))) <= 1e-14


def test_polygon_edge_sum_equals_rectangle_closed_form():
    L, W = 0.3, 0.1
    rect = _ribbon((0, 0, 0), (L, 0, 0), W, (0, 1, 0))
    pv = [(0, -W / 2, 0), (L, -W / 2, 0), (L, W / 2, 0), (0, W / 2, 0)]
    poly = _ribbon((0, 0, 0), (L, 0, 0), W, (0, 1, 0), npv=4, pv=pv)
    rng = np.random.default_rng(3)
    for _ in range(40):
        p = rng.uniform([-0.2, -0.2, -0.1], [0.5, 0.2, 0.1])
        pt = (C.c_double * 3)(*p)
        a = ORC.orc_rect_potential(C.byref(rect), pt)
        b = ORC.orc_poly_potential(C.byref(poly), pt)
        assert abs(a - b) <= 1e-12 * a


def test_thin_wire_sum_rule():
    """Collinear segments of one wire: sum_ij I_ij (8 segments) = I_self(whole wire)."""
    one = deck_text("wire_single.peec")
    eight = one.replace("5.8e7 1\n", "5.8e7 8\n")
    m1, m8 = oracle_peec(one), oracle_peec(eight)
    assert m1.p.nseg == 1 and m8.p.nseg == 8
    l1, l8 = oracle_lp(m1, 0.0), oracle_lp(m8, 0.0)
    assert abs(l8.sum() - l1[0, 0]) <= 1e-12 * abs(l1[0, 0])
    m1.close()
    m8.close()


def test_hoer_love_additivity():
    """Bar B split along its axis into B1 + B2: I(A, B) = I(A, B1) + I(A, B2) (same section)."""
    def bar(x1, x2, wid, thick):
        s = _ribbon(x1, x2, wid, (0, 1, 0))
        s.thick = thick
        s.vol = 1
        return s
    A = bar((0, 0, 0), (0.01, 0, 0), 1e-3, 2e-4)
    B = bar((0.002, 1.5e-3, 3e-4), (0.014, 1.5e-3, 3e-4), 8e-4, 1e-4)
    B1 = bar((0.002, 1.5e-3, 3e-4), (0.007, 1.5e-3, 3e-4), 8e-4, 1e-4)
    B2 = bar((0.007, 1.5e-3, 3e-4), (0.014, 1.5e-3, 3e-4), 8e-4, 1e-4)
    # bar_pair is normalised by the two cross-sections, identical for B, B1 and B2
    whole = ORC.orc_bar_pair(C.byref(A), C.byref(B))
    parts = ORC.orc_bar_pair(C.byref(A), C.byref(B1)) + ORC.orc_bar_pair(C.byref(A), C.byref(B2))
    assert abs(whole - parts) <= 1e-9 * abs(whole)


@pytest.mark.parametrize("n", [5, 64, 200])
def test_oracle_lu_matches_lapack(n):
    rng = np.random.default_rng(n)
    a0 = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    b0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a, b = a0.copy(), b0.copy()
    piv = np.zeros(n, dtype=np.int32)
    P = C.POINTER(abi.d_complex_t)
    assert ORC.orc_lu_decomp(n, a.ctypes.data_as(P), piv.ctypes.data_as(C.POINTER(C.c_int))) == -1
    ORC.orc_lu_solve(n, a.ctypes.data_as(P), piv.ctypes.data_as(C.POINTER(C.c_int)), b.ctypes.data_as(P))
    assert np.abs(b - np.linalg.solve(a0, b0)).max() <= 1e-10 * np.abs(b).max()


# ------------------------------------------------------------------ 4. known answers
def test_wire_inductance_known_answer():
    z = oracle_zmat(deck_text("wire_single.peec"))
    m = oracle_peec(deck_text("wire_single.peec"))
    L = z[0, 0, 0].imag / (2 * math.pi * m.freq(0))
    m.close()
    assert abs(L - 1.320380e-6) <= 1e-3 * 1.320380e-6


def test_wire_capacitance_known_answer():
    m = oracle_peec(deck_text("wire_cap.peec"))
    c = oracle_c(m, 0.0)
    m.close()
    assert abs(c.real.sum() - 8.42673e-12) <= 1e-3 * 8.42673e-12


def test_half_wave_dipole_resonance_and_monopole_image():
    text = deck_text("dipole_halfwave.peec")
    z = oracle_zmat(text)[:, 0, 0]
    m = oracle_peec(text)
    f = np.array([m.freq(i) for i in range(m.p.nfreq)])
    m.close()
    k = int(np.nonzero(np.diff(np.sign(z.imag)))[0][0])
    t = -z.imag[k] / (z.imag[k + 1] - z.imag[k])
    fres = f[k] + t * (f[k + 1] - f[k])
    rin = z.real[k] + t * (z.real[k + 1] - z.real[k])
    assert abs(1.0 * fres / 2.99792458e8 - 0.478) <= 0.05 * 0.478
    assert abs(rin - 73.1) <= 0.10 * 73.1
    zm = oracle_zmat(deck_text("monopole_gp.peec"))[:, 0, 0]
    assert np.max(np.abs(zm - z / 2) / np.abs(z / 2)) <= 1e-6


def test_plate_capacitance_richardson():
    base = deck_text("plate_cap.peec")
    cs = []
    for n in (8, 16):
        m = oracle_peec(base.replace("5.8e7 8 8", f"5.8e7 {n} {n}"))
        cs.append(oracle_c(m, 0.0).real.sum())
        m.close()
    c_inf = 2 * cs[1] - cs[0]
    assert abs(c_inf - 4.08100e-11) <= 0.01 * 4.08100e-11


def test_full_stack_deck_fix():
    """The full-stack configuration as written puts the brick's bottom face on the ground plane:
    those charge cells coincide with their own images, their P rows vanish and the sweep stops on
    a singular potential-coefficient matrix (potential.c:228-233).  With the ground plane 0.1 mm
    below the brick (decks.full_stack_deck's default) the same deck runs; small grids here."""
    small = dict(grid=8, ndivt=2, diel=(4, 4, 2), nd=1)
    with pytest.raises(RuntimeError):
        oracle_zmat(decks.full_stack_deck(gpz=0.0, **small))
    z = oracle_zmat(decks.full_stack_deck(**small))
    assert z.shape == (2, 1, 1) and np.all(np.isfinite(z))
    # a capacitive input at 20 MHz (plate over ground through the brick): Im Z < 0
    assert z[0, 0, 0].imag < 0


def _disk_deck(nr, ns):
    return ("OpenPEEC 1 0\ntitle = disk\ncapacitance = 1\nnode = 1 0 0 0\nnode = 2 0 -0.2 0\n"
            f"disk = 0 0 0  0 0 1  0.1 1e-4 5.8e7 {nr} {ns}\nport = 1 2 50\nfrequency = 1e6 1e6 0\nend\n")


def _disk_pc(text):
    m = oracle_peec(text)
    nc = m.p.ncell
    P = np.zeros((nc, nc), dtype=np.complex128)
    Cm = np.zeros_like(P)
    assert ORC.orc_pot_fill(C.byref(m.p), 0.0, P.ctypes.data_as(C.POINTER(abi.d_complex_t)),
                            Cm.ctypes.data_as(C.POINTER(abi.d_complex_t))) == 0
    out = (m.p.nchg, nc, P, Cm.real.sum())
    m.close()
    return out


@pytest.mark.parametrize("k", [2, 3])
def test_disk_centre_ring_split_equals_single_cell(monkeypatch, k):
    """A disk centre ring of more than POLY_MAX vertices (nsec > 32, beyond the reference) is
    built as k congruent sector cells on the centre node.  Forced on a disk the reference can
    build (nsec = 24), the split must reproduce the single cell: same capacitive cells, the centre
    row of P to 1e-5 (its far entries differ only by which quadrature rule each sub-cell pair
    takes), the centre diagonal and the total capacitance to 1e-7.  A normalisation by the
    whole ring's area instead of the sum of the sub-cells' sqrt(area) would be off by sqrt(k)."""
    text = _disk_deck(4, 24)
    monkeypatch.setenv("PEEC_DISK_CENTRE_CELLS", "1")
    n1, nc1, P1, C1 = _disk_pc(text)
    monkeypatch.setenv("PEEC_DISK_CENTRE_CELLS", str(k))
    nk, nck, Pk, Ck = _disk_pc(text)
    assert (nk, nck) == (n1 + k - 1, nc1)
    assert np.max(np.abs(Pk[0, 1:] - P1[0, 1:]) / np.abs(P1[0, 1:])) <= 1e-5
    assert abs(Pk[0, 0] - P1[0, 0]) <= 1e-7 * abs(P1[0, 0])
    assert abs(Ck / C1 - 1) <= 1e-7


def test_disk_beyond_the_reference_converges_to_the_analytic_capacitance():
    """The 8 x 96 disk (centre ring split in 4): total capacitance within 1 % of the thin disk's
    8 eps0 a (7.08335e-12 F for a = 0.1 m, data/sample/peec_check.sh), closer than 8 x 32."""
    _, _, _, c96 = _disk_pc(_disk_deck(8, 96))
    _, _, _, c32 = _disk_pc(_disk_deck(8, 32))
    ref = 7.08335e-12
    assert abs(c96 / ref - 1) <= 0.01
    assert abs(c96 / ref - 1) < abs(c32 / ref - 1)
