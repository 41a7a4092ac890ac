"""Device-resident sweep (fill -> P^-1 -> MNA assembly -> LU -> port solves) against the oracle's
sweep (the CPU restatement of the reference's solve(), pinned bit for bit to it by
test_oracle.py) and against the Z(f) the reference's own solve() wrote into
tests/golden/ref_vectors.npz: Zin / Z within 1e-9 relative."""
import numpy as np
import pytest

from conftest import DECKS, deck_text
from openpeec_amd import Peec, decks
from oracle_py import oracle_zmat
from refharness import rel_err

pytestmark = pytest.mark.gpu

GOLD = np.load(DECKS.parent / "ref_vectors.npz")


@pytest.mark.parametrize("key,text_fn", [
    ("dipole_c1", lambda: decks.dipole_deck()),
    ("monopole_gp", lambda: deck_text("monopole_gp.peec")),
    ("strip_dipole_ret", lambda: deck_text("strip_dipole_ret.peec")),
    ("twoport_dist", lambda: deck_text("twoport_dist.peec")),
    ("plate_12", lambda: decks.plate_deck(n=12, ndiv=4)),
])
def test_solve_matches_reference(lib, key, text_fn):
    text = text_fn()
    print(f"[sweep] {key}", flush=True)
    m = Peec.from_deck_text(text, lib)
    m.setup()
    m.solve()
    zm = m.zmat()
    m.close()
    zr = GOLD[f"zmat/{key}"]
    assert rel_err(zm, zr) <= 1e-9, rel_err(zm, zr)
    zo = oracle_zmat(text)
    assert rel_err(zm, zo) <= 1e-9, rel_err(zm, zo)


def test_peec_sweep_equals_solve(lib):
    text = decks.dipole_deck(ndiv=6)
    zo = oracle_zmat(text)
    m = Peec.from_deck_text(text, lib)
    m.setup()
    z, st = m.sweep(0, m.nfreq)
    assert rel_err(z, zo) <= 1e-9
    assert st.total_ms > 0 and st.lu_ms > 0
    m.close()
