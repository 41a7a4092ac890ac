"""The full-stack workload at a reduced size — ground plane 0.1 mm below an FR4 brick
(5 x 5 x 2 dielectric cells) under a copper plate on a 12 x 12 grid split into 4 volume layers,
6 points 20 MHz - 3 GHz — through the C ABI in every dense form, against the reference's own
solve() on the same deck (tests/golden/c5_vectors.npz, make_workload_golden.py).

Every form within 1e-6 (the metric's bar): plain, precision (FP32-MFMA LU with FP64 refinement),
accelerated (k-factored fill, FP32 inverse factors, GMRES) and batched.  The plain form's own bar
is 1e-9, met on the same stack with the plate as one layer of ribbons; with the plate split into
4 volume layers (8.75 um thick bars) its Lp entries come from the 64-term Hoer-Love sum, whose
cancellation amplifies the last-bit differences of the device's and glibc's asinh / atan / log
by ~1e8 (test_cancellation.py pins both sides against quad precision), and Z follows at ~3e-7.
"""
import numpy as np
import pytest

from conftest import DECKS
from openpeec_amd import PEEC_FORM_MIXED, PEEC_FORM_PLAIN, PEEC_FORM_PRECISION, Batch, Peec, decks
from refharness import rel_err

pytestmark = pytest.mark.gpu

GOLD = np.load(DECKS.parent / "c5_vectors.npz")
C5_SMALL = dict(grid=12, ndivt=4, diel=(5, 5, 2), nd=5)


@pytest.mark.parametrize("key,ndivt,form,tol", [
    ("c5", 4, PEEC_FORM_PLAIN, 1e-6), ("c5", 4, PEEC_FORM_PRECISION, 1e-6), ("c5", 4, PEEC_FORM_MIXED, 1e-6),
    ("c5_ribbon", 1, PEEC_FORM_PLAIN, 1e-9), ("c5_ribbon", 1, PEEC_FORM_MIXED, 1e-6)])
def test_full_stack_forms_match_reference(lib, key, ndivt, form, tol):
    print(f"[c5] {key} form {form}", flush=True)
    lib.peec_sweep_reset()
    m = Peec.from_deck_text(decks.full_stack_deck(**dict(C5_SMALL, ndivt=ndivt)), lib)
    m.setup()
    z, st = m.sweep(0, m.nfreq, form)
    m.close()
    e = rel_err(z, GOLD[f"{key}/zmat"])
    assert e <= tol, e
    if form == PEEC_FORM_PRECISION:
        assert st.berr <= 1e-10


def test_full_stack_solve_entry_point(lib):
    """The reference's solve() entry point on the same deck (Z, Zin and S written into peec_t)."""
    m = Peec.from_deck_text(decks.full_stack_deck(**C5_SMALL), lib)
    m.setup()
    m.solve()
    assert rel_err(m.zmat(), GOLD["c5/zmat"]) <= 1e-6
    m.close()


def test_full_stack_batched_form(lib):
    m = Peec.from_deck_text(decks.full_stack_deck(**C5_SMALL), lib)
    m.setup()
    if m.p.nunknown + m.p.ncell > 4096:
        m.close()
        pytest.skip("reduced deck above the batched sweep's 4096 charge-form unknowns")
    b = Batch([m], 0, m.nfreq)
    z = b.sweep()[0][0]
    b.close()
    m.close()
    assert rel_err(z, GOLD["c5/zmat"]) <= 1e-6, rel_err(z, GOLD["c5/zmat"])
