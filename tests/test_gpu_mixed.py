"""Accelerated sweep (charge-unknown system, FP32-stored stale LU preconditioning FP64 GMRES,
batched frequencies) against the plain FP64 device sweep and the oracle: Zin within 1e-6
relative (the bar of the accelerated form), and in practice ~1e-9 at the 1e-10 GMRES tolerance."""
import numpy as np
import pytest

from conftest import deck_text
from openpeec_amd import PEEC_FORM_MIXED, PEEC_FORM_PLAIN, Peec, decks
from oracle_py import oracle_zmat

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


@pytest.mark.parametrize("text_fn,batch", [
    (lambda: decks.dipole_deck(), 4),
    (lambda: decks.dipole_deck(ndiv=9), 1),
    (lambda: deck_text("monopole_gp.peec"), 3),
    (lambda: deck_text("twoport_dist.peec"), 2),
    (lambda: decks.plate_deck(n=10, ndiv=6), 4),
])
def test_mixed_matches_oracle(lib, text_fn, batch):
    text = text_fn()
    print(f"[mixed] {text.splitlines()[1]} batch={batch}", flush=True)
    lib.peec_sweep_reset()
    m = Peec.from_deck_text(text, lib)
    m.setup()
    z, st = m.sweep(0, m.nfreq, PEEC_FORM_MIXED, batch)
    zo = oracle_zmat(text)
    assert _rel(z, zo) <= 1e-6, _rel(z, zo)
    assert st.n_lu >= 1
    m.close()


def test_mixed_matches_plain_on_plate(lib):
    text = decks.plate_deck(n=24, ndiv=40)
    lib.peec_sweep_reset()
    m = Peec.from_deck_text(text, lib)
    m.setup()
    zp, _ = m.sweep(0, 12, PEEC_FORM_PLAIN)
    zm, st = m.sweep(0, 12, PEEC_FORM_MIXED, 4)
    assert _rel(zm, zp) <= 1e-6
    assert st.n_lu <= 3
    m.close()


def test_mixed_short_first_batch(lib):
    """A sweep whose first batch holds fewer systems than the batch size: the unused slots carry
    no system and must not be read (one frequency with batch 8)."""
    text = decks.dipole_deck()
    lib.peec_sweep_reset()
    m = Peec.from_deck_text(text, lib)
    m.setup()
    z, _ = m.sweep(7, 1, PEEC_FORM_MIXED, 8)
    zo = oracle_zmat(text, 7, 1)
    assert _rel(z, zo) <= 1e-6
    m.close()


def test_mixed_fp64_preconditioner_build(lib, monkeypatch):
    """The preconditioner factored and inverted in FP64 (PEEC_MIXED_LU64=1) instead of the
    default FP32 storage with FP32 MFMA updates: the same Z to the form's bar, with no more
    GMRES iterations than the FP32 build needs plus a margin."""
    text = decks.plate_deck(n=24, ndiv=40)
    m = Peec.from_deck_text(text, lib)
    m.setup()
    lib.peec_sweep_reset()
    z32, st32 = m.sweep(0, 12, PEEC_FORM_MIXED, 4)
    monkeypatch.setenv("PEEC_MIXED_LU64", "1")
    lib.peec_sweep_reset()
    z64, st64 = m.sweep(0, 12, PEEC_FORM_MIXED, 4)
    print(f"[mixed] GMRES iterations: FP32 build {st32.gmres_its}, FP64 build {st64.gmres_its}", flush=True)
    assert _rel(z32, z64) <= 1e-6
    assert st32.gmres_its <= 1.25 * st64.gmres_its + 8
    m.close()


@pytest.mark.parametrize("text_fn", [
    lambda: decks.plate_deck(n=24, ndiv=8),          # Lp 1 200 (edge tiles of 48), P 625
    lambda: deck_text("monopole_gp.peec"),           # ground-plane images
])
def test_mixed_symmetric_product_matches_full(lib, monkeypatch, text_fn):
    """The GMRES product read from the upper triangles only (k_mvsym, default) against the
    product over both triangles (k_mvdense): the same Zin to the GMRES tolerance."""
    text = text_fn()
    print(f"[mixed-sym] {text.splitlines()[1]}", flush=True)
    m = Peec.from_deck_text(text, lib)
    m.setup()
    zs = {}
    for sym in ("1", "0"):
        monkeypatch.setenv("PEEC_MV_SYM", sym)
        lib.peec_sweep_reset()
        zs[sym], st = m.sweep(0, m.nfreq, PEEC_FORM_MIXED, 4)
        print(f"  sym={sym}: {st.gmres_its} iterations", flush=True)
    assert _rel(zs["1"], zs["0"]) <= 1e-8, _rel(zs["1"], zs["0"])
    m.close()


def test_upper_fill_preconditioner(lib, monkeypatch):
    """With the symmetric product the batch fill writes only the upper triangles (and diagonal
    64 x 64 blocks) of Lp and P, and a factorisation mirrors the lower triangle of the system it
    takes.  A wrong mirror leaves GMRES with the preconditioner of another matrix: the answer
    still converges, but the iterations do not match those of the two-triangle product (whose
    fill writes whole matrices)."""
    text = decks.plate_deck(n=24, ndiv=40)
    its = {}
    for sym in ("1", "0"):
        monkeypatch.setenv("PEEC_MV_SYM", sym)
        print(f"[mixed] plate 24x24, symmetric product {sym}", flush=True)
        lib.peec_sweep_reset()
        m = Peec.from_deck_text(text, lib)
        m.setup()
        zp, _ = m.sweep(0, 12, PEEC_FORM_PLAIN)
        zm, st = m.sweep(0, 12, PEEC_FORM_MIXED, 4)
        assert _rel(zm, zp) <= 1e-6
        its[sym] = (st.gmres_its, st.n_lu)
        m.close()
    (i1, n1), (i0, n0) = its["1"], its["0"]
    assert n1 == n0, its
    assert abs(i1 - i0) <= max(2, 0.05 * i0), its
