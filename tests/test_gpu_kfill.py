"""k-factored fill of the accelerated sweep against the element-by-element device fill (which is
itself pinned to the oracle): for every pair the k-independent moments of the regular-part
quadrature reproduce the reference's (exp(-jkR) - 1)/R sum at any wavenumber of the deck's band.
Bar: entries within 1e-12 of the matrix scale and 1e-10 relative for every entry above 1e-6 of
it (the expansion is truncated below 1e-16; the rest is rounding of a reordered sum)."""
import numpy as np
import pytest

from conftest import deck_text
from openpeec_amd import Peec, decks

pytestmark = pytest.mark.gpu


def _retarded(text: str, f0: float = 1e8, f1: float = 3e9, ndiv: int = 20) -> str:
    """The deck with capacitance on and its sweep replaced by a retarded one over [f0, f1]."""
    keep = [ln for ln in text.splitlines()
            if not ln.strip().startswith(("retardation", "capacitance", "frequency", "end"))]
    tail = ["retardation = 1", "capacitance = 1", f"frequency = {f0} {f1} {ndiv}", "end"]
    return "\n".join(keep + tail) + "\n"


def _cmp(a, b):
    scale = float(np.max(np.abs(b)))
    d = np.abs(a - b)
    big = np.abs(b) > 1e-6 * scale
    rel = float(np.max(d[big] / np.abs(b[big]))) if big.any() else 0.0
    return float(np.max(d)) / scale, rel


@pytest.mark.parametrize("name,text_fn", [
    ("dipole", lambda: decks.dipole_deck()),
    ("monopole_gp", lambda: deck_text("monopole_gp.peec")),
    ("wire_gp", lambda: deck_text("wire_gp.peec")),
    ("plate12", lambda: decks.plate_deck(n=12, ndiv=20)),
    ("plate_thick", lambda: _retarded(deck_text("plate_thick.peec"))),
    ("plate_gp", lambda: decks.plate_deck(n=10, ndiv=20, groundplane=0.0, z=0.002)),
    ("strip_dipole_ret_bar", lambda: deck_text("strip_dipole_ret_bar.peec")),
    ("quad_square", lambda: _retarded(deck_text("quad_square.peec"))),
    ("disk_annulus", lambda: _retarded(deck_text("disk_annulus.peec"))),
    ("compress_plate", lambda: _retarded(deck_text("compress_plate.peec"))),
    ("random_wire", lambda: decks.random_wire_deck(7)),
])
def test_kfill_matches_direct(lib, name, text_fn):
    text = text_fn()
    m = Peec.from_deck_text(text, lib)
    m.setup()
    if not m.p.retardation:
        m.close()
        pytest.skip("static deck: no k dependence")
    nf = m.nfreq
    for i in sorted({0, nf // 2, nf - 1}):
        f = m.freq(i)
        print(f"[kfill] {name} f={f:.4g}", flush=True)
        lp0, P0 = m.fill_matrices(f, 0)
        lp1, P1 = m.fill_matrices(f, 1)
        a, r = _cmp(lp1, lp0)
        assert a <= 1e-12 and r <= 1e-10, (name, "Lp", f, a, r)
        if P0.size:
            a, r = _cmp(P1, P0)
            assert a <= 1e-12 and r <= 1e-10, (name, "P", f, a, r)
    m.close()
