"""Lp and C = P^-1 fills on the GPU against the oracle (the CPU restatement, pinned bit for bit to
the reference's own fills by test_oracle.py) and against the reference's golden vectors
(tests/golden/ref_vectors.npz) where those hold the same deck and frequency.

Per-entry bar for Lp: relative error (complex modulus) <= 1e-12 for every quadrature-evaluated
class.  Entries produced by cancelling closed forms carry the libm ulp differences (device ocml
vs glibc) amplified by their cancellation factor: the 4-term parallel-filament sum (fpar,
~1.4e-12 on collinear far pairs) and the 64-term Hoer-Love sum (~2.5e-12 for the plate_thick
bars, ~3e-8 for the thin dielectric layers of diel_pp).  Those decks carry their measured bar
here; test_cancellation.py pins both sides of those entries against a quad-precision evaluation
of the same formulas.  C = P^-1 goes through a blocked LU (different summation order from
the reference's unblocked elimination), so it is compared at 1e-9 relative to max |C|.
"""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import DECKS, deck_text
from openpeec_amd import Peec, abi, decks
from oracle_py import oracle_c, oracle_lib, oracle_lp, oracle_peec
from refharness import rel_err

pytestmark = pytest.mark.gpu

GOLD = np.load(DECKS.parent / "ref_vectors.npz")

CASES = [
    # deck, frequencies, Lp tolerance
    ("dipole_halfwave.peec", [0.0, 1.45e8, 7e8], 3e-12),
    ("wire_gp.peec", [0.0, 3e8, 7e8], 1e-12),
    ("monopole_gp.peec", [1.4e8, 7e8], 1e-12),
    ("plate_cap.peec", [0.0, 2e8, 7e8], 1e-12),
    ("plate_strip.peec", [0.0, 1e9], 1e-12),
    ("strip_dipole_ret.peec", [2.8e8, 7e8], 1e-12),
    ("quad_square.peec", [0.0, 1e9], 1e-12),
    ("quad_taper.peec", [5e8, 7e8], 1e-12),
    ("disk_cap.peec", [0.0, 1e9, 7e8], 1e-12),
    ("plate_thick.peec", [0.0, 1e9], 5e-12),
    ("twowire_cap.peec", [0.0], 1e-12),
    ("diel_pp.peec", [1e9], 1e-7),
    ("disk_annulus.peec", [7e8], 1e-12),
]


def _retarded(text):
    return text.replace("retardation = 0", "").replace("OpenPEEC 1 0", "OpenPEEC 1 0\nretardation = 1", 1)


def _kw(f):
    return 2 * math.pi * f / 2.99792458e8


def _gold_lp(name, f):
    key = f"lp/{name}/{f:.0e}"
    if key not in GOLD.files:
        return None, None
    idx = f"lpidx/{name}/{f:.0e}"
    return GOLD[key], (GOLD[idx] if idx in GOLD.files else None)


@pytest.mark.parametrize("name,freqs,tol", CASES)
def test_lp_and_c_match_oracle_and_reference(lib, name, freqs, tol):
    text = _retarded(deck_text(name))
    print(f"[fill] {name}", flush=True)
    orc = oracle_peec(text)
    mine = Peec.from_deck_text(text, lib)
    mine.wire_build()
    stem = name[:-5]
    for f in freqs:
        lm = mine.lp_fill(f)
        lo = oracle_lp(orc, _kw(f))
        assert lm.shape == lo.shape
        assert rel_err(lm, lo) <= tol, (name, f, rel_err(lm, lo))
        want, ij = _gold_lp(stem, f)
        if want is not None:
            got = lm if ij is None else lm[ij[:, 0], ij[:, 1]]
            assert rel_err(got, want) <= tol, (name, f, "golden", rel_err(got, want))
        if orc.p.capacitance and orc.p.nchg > 0:
            cm = mine.pot_fill(f)
            co = oracle_c(orc, _kw(f))
            assert np.abs(cm - co).max() / np.abs(co).max() <= 1e-9, (name, f)
            key = f"c/{stem}/{f:.0e}"
            if key in GOLD.files:
                cr = GOLD[key]
                assert np.abs(cm - cr).max() / np.abs(cr).max() <= 1e-9, (name, f, "golden")
    orc.close()
    mine.close()


def test_plate_c2_small_every_entry_matches_oracle(lib):
    """The precision workload's geometry at 16 x 16: every Lp entry and every C entry."""
    text = decks.plate_deck(n=16, ndiv=2)
    orc = oracle_peec(text)
    mine = Peec.from_deck_text(text, lib)
    mine.wire_build()
    f = 3e9
    assert rel_err(mine.lp_fill(f), oracle_lp(orc, _kw(f))) <= 1e-12
    co = oracle_c(orc, _kw(f))
    assert np.abs(mine.pot_fill(f) - co).max() / np.abs(co).max() <= 1e-9
    orc.close()
    mine.close()


def test_lp_entry_and_pot_entry(lib):
    """lp_entry on sampled pairs against the oracle's; pot_entry on every pair against P as the
    oracle's pot_fill aggregates it (before inversion), entry by entry at 1e-12."""
    text = _retarded(deck_text("disk_cap.peec"))
    orc = oracle_peec(text)
    mine = Peec.from_deck_text(text, lib)
    mine.wire_build()
    lib.pot_cells(C.byref(mine.p))
    kw = _kw(1e9)
    for (i, j) in [(0, 0), (3, 17), (17, 3), (100, 255)]:
        a = lib.lp_entry(C.byref(mine.p), i, j, kw)
        b = oracle_lib().orc_lp_entry(C.byref(orc.p), i, j, kw)
        assert abs(complex(a.r, a.i) - complex(b.r, b.i)) <= 1e-12 * abs(complex(b.r, b.i))
    nc = orc.p.ncell
    pm = np.zeros((nc, nc), dtype=np.complex128)
    assert oracle_lib().orc_pot_fill(C.byref(orc.p), kw, pm.ctypes.data_as(C.POINTER(abi.d_complex_t)), None) == 0
    got = np.zeros_like(pm)
    for i in range(nc):
        for j in range(nc):
            e = lib.pot_entry(C.byref(mine.p), i, j, kw)
            got[i, j] = complex(e.r, e.i)
    assert rel_err(got, pm) <= 1e-12, rel_err(got, pm)
    orc.close()
    mine.close()
