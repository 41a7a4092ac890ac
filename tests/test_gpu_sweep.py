"""Device-resident sweep (fill -> P^-1 -> MNA assembly -> LU -> port solves) against the
reference's solve() on the same decks: Zin / Z / S within 1e-9 relative."""
import numpy as np
import pytest

from conftest import deck_text
from openpeec_amd import Peec, decks
from refharness import RefPeec, ref_available, rel_err

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not ref_available(), reason="reference build not present")]


def _ref_zmat(text):
    ref = RefPeec(text)
    ref.setup_like_main()
    ref.solve()
    z = ref.zmat()
    ref.close()
    return z


@pytest.mark.parametrize("text_fn", [
    lambda: decks.dipole_deck(),
    lambda: deck_text("monopole_gp.peec"),
    lambda: deck_text("strip_dipole_ret.peec"),
    lambda: deck_text("twoport_dist.peec"),
    lambda: decks.plate_deck(n=12, ndiv=4),
])
def test_solve_matches_reference(lib, text_fn):
    text = text_fn()
    print(f"[sweep] {text.splitlines()[1]}", flush=True)
    zr = _ref_zmat(text)
    m = Peec.from_deck_text(text, lib)
    m.setup()
    m.solve()
    zm = m.zmat()
    assert rel_err(zm, zr) <= 1e-9, rel_err(zm, zr)
    m.close()


def test_peec_sweep_equals_solve(lib):
    text = decks.dipole_deck(ndiv=6)
    zr = _ref_zmat(text)
    m = Peec.from_deck_text(text, lib)
    m.setup()
    z, st = m.sweep(0, m.nfreq)
    assert rel_err(z, zr) <= 1e-9
    assert st.total_ms > 0 and st.lu_ms > 0
    m.close()
