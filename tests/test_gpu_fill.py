"""Lp and C = P^-1 fills on the GPU against the reference's own FP64 fills on the same decks.

Per-entry bar for Lp: relative error (complex modulus) <= 1e-12 for every quadrature-evaluated
class.  Entries produced by cancelling closed forms carry the libm ulp differences (device ocml
vs glibc) amplified by their cancellation factor: the 4-term parallel-filament sum (fpar,
~1.4e-12 on collinear far pairs) and the 64-term Hoer-Love sum (~2.5e-12 for the plate_thick
bars, ~3e-8 for the thin dielectric layers of diel_pp).  Those decks carry their measured bar
here; test_cancellation.py pins both sides of those entries against a quad-precision evaluation
of the same formulas.  C = P^-1 goes through a blocked LU (different summation order from
the reference's unblocked elimination), so it is compared at 1e-9 relative to max |C|.
"""
import numpy as np
import pytest

from conftest import deck_text
from openpeec_amd import Peec, decks
from refharness import RefPeec, ref_available, rel_err

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not ref_available(), reason="reference build not present")]

CASES = [
    # deck, frequencies, Lp tolerance
    ("dipole_halfwave.peec", [0.0, 1.45e8], 3e-12),
    ("wire_gp.peec", [0.0, 3e8], 1e-12),
    ("monopole_gp.peec", [1.4e8], 1e-12),
    ("plate_cap.peec", [0.0, 2e8], 1e-12),
    ("plate_strip.peec", [0.0, 1e9], 1e-12),
    ("strip_dipole_ret.peec", [2.8e8], 1e-12),
    ("quad_square.peec", [0.0, 1e9], 1e-12),
    ("quad_taper.peec", [5e8], 1e-12),
    ("disk_cap.peec", [0.0, 1e9], 1e-12),
    ("plate_thick.peec", [0.0, 1e9], 5e-12),
    ("twowire_cap.peec", [0.0], 1e-12),
    ("diel_pp.peec", [1e9], 1e-7),
]


def _retarded(text):
    return text.replace("retardation = 0", "").replace("OpenPEEC 1 0", "OpenPEEC 1 0\nretardation = 1", 1)


@pytest.mark.parametrize("name,freqs,tol", CASES)
def test_lp_and_c_match_reference(lib, name, freqs, tol):
    text = _retarded(deck_text(name))
    print(f"[fill] {name}", flush=True)
    ref = RefPeec(text)
    ref.wire_build()
    mine = Peec.from_deck_text(text, lib)
    mine.wire_build()
    for f in freqs:
        lr = ref.lp_fill(f)
        lm = mine.lp_fill(f)
        assert lm.shape == lr.shape
        assert rel_err(lm, lr) <= tol, (name, f, rel_err(lm, lr))
        if ref.p.capacitance and ref.p.nchg > 0:
            cr = ref.pot_fill(f)
            cm = mine.pot_fill(f)
            assert np.abs(cm - cr).max() / np.abs(cr).max() <= 1e-9, (name, f)
    ref.close()
    mine.close()


def test_plate_c2_small_rows_match_reference(lib):
    """The precision workload's geometry at 16 x 16: every Lp entry and every C entry."""
    text = decks.plate_deck(n=16, ndiv=2)
    ref = RefPeec(text)
    ref.wire_build()
    mine = Peec.from_deck_text(text, lib)
    mine.wire_build()
    f = 3e9
    assert rel_err(mine.lp_fill(f), ref.lp_fill(f)) <= 1e-12
    cr = ref.pot_fill(f)
    assert np.abs(mine.pot_fill(f) - cr).max() / np.abs(cr).max() <= 1e-9


def test_lp_entry_and_pot_entry(lib):
    text = _retarded(deck_text("disk_cap.peec"))
    ref = RefPeec(text)
    ref.wire_build()
    ref.lib.pot_cells(ref.p)
    mine = Peec.from_deck_text(text, lib)
    mine.wire_build()
    lib.pot_cells(mine.p)
    kw = 2 * np.pi * 1e9 / 2.99792458e8
    import ctypes as C
    for (i, j) in [(0, 0), (3, 17), (17, 3), (100, 255)]:
        a = lib.lp_entry(C.byref(mine.p), i, j, kw)
        b = ref.lib.lp_entry(C.byref(ref.p), i, j, kw)
        assert abs(complex(a.r, a.i) - complex(b.r, b.i)) <= 1e-12 * abs(complex(b.r, b.i))
    for (i, j) in [(0, 0), (5, 9), (128, 1)]:
        a = lib.pot_entry(C.byref(mine.p), i, j, kw)
        b = ref.lib.pot_entry(C.byref(ref.p), i, j, kw)
        assert abs(complex(a.r, a.i) - complex(b.r, b.i)) <= 1e-12 * abs(complex(b.r, b.i))
