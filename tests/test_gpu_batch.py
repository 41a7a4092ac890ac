"""Batched sweep of many structures (peec_batch_*): every (structure, frequency) pair factored
and solved by the persistent batched kernels, replayed as hipGraphs.

  * random wire structures of uneven size (the batching workload's generator, small sizes so the
    oracle finishes in seconds) against the oracle's dense FP64 sweep: Z within 1e-8 relative
    (k-factored fill <= 1e-10 per entry, a different elimination order; the metric's bar is 1e-6);
  * mixed decks in one batch -- ground-plane images (two record terms), two ports, ribbons with
    several sub-cells per capacitive cell, a lumped RLC network, and a wide-band dipole whose
    quadrature order changes inside the band (those pairs evaluated by the element code in the
    batched list kernel) -- against the plain FP64 sweep of each structure;
  * a second sweep replays the captured graphs and reproduces the first bit for bit; the
    autotuner timed every candidate and kept the fastest.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import deck_text
from openpeec_amd import PEEC_FORM_PLAIN, PEEC_FORM_PRECISION, Batch, Peec, abi, decks
from oracle_py import oracle_zmat

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.abs(b)))


@pytest.mark.parametrize("nb", [16, 32, 64])
def test_batched_lu_factors(lib, nb):
    """P A = L U for matrices of uneven order (1 .. 300, edges around the panel width) factored
    together, one workgroup each, and the same interchanges as lu_decomp (random matrices have no
    near-ties in |re| + |im|)."""
    rng = np.random.default_rng(5 + nb)
    ns = [1, 2, 15, 16, 17, 31, 32, 33, 64, 65, 100, 131, 300]
    mats = [rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) for n in ns]
    a = np.concatenate([m.ravel() for m in mats]).astype(np.complex128)
    piv = np.zeros(sum(ns), dtype=np.int32)
    narr = (C.c_int * len(ns))(*ns)
    print(f"[batch-lu] nb={nb} orders {ns}", flush=True)
    rc = lib.peec_batch_lu_decomp(len(ns), narr, a.ctypes.data_as(C.POINTER(abi.d_complex_t)),
                                  piv.ctypes.data_as(C.POINTER(C.c_int)), nb)
    assert rc == -1
    off = poff = 0
    for n, m0 in zip(ns, mats):
        lu = a[off:off + n * n].reshape(n, n)
        pv = piv[poff:poff + n]
        L = np.tril(lu, -1) + np.eye(n)
        U = np.triu(lu)
        pa = m0.copy()
        for k in range(n):
            if pv[k] != k:
                pa[[k, pv[k]]] = pa[[pv[k], k]]
        err = np.abs(L @ U - pa).max() / np.abs(m0).max()
        assert err < 1e-12, (n, err)
        ref = m0.copy()
        rpiv = np.zeros(n, dtype=np.int32)
        assert lib.lu_decomp(n, ref.ctypes.data_as(C.POINTER(abi.d_complex_t)), rpiv.ctypes.data_as(C.POINTER(C.c_int))) == -1
        assert np.array_equal(rpiv, pv), n
        off += n * n
        poff += n


def test_batched_lu_singular(lib):
    """A zero column in the second matrix: its index is returned."""
    ns = [8, 8]
    rng = np.random.default_rng(9)
    m1 = rng.standard_normal((8, 8)) + 0j
    m2 = rng.standard_normal((8, 8)) + 0j
    m2[:, 3] = 0
    a = np.concatenate([m1.ravel(), m2.ravel()]).astype(np.complex128)
    piv = np.zeros(16, dtype=np.int32)
    rc = lib.peec_batch_lu_decomp(2, (C.c_int * 2)(*ns), a.ctypes.data_as(C.POINTER(abi.d_complex_t)),
                                  piv.ctypes.data_as(C.POINTER(C.c_int)), 16)
    assert rc == 1


def _sweep(text: str, f0: float, f1: float, nd: int) -> str:
    """The deck with its frequency line replaced by a linear sweep of nd + 1 points."""
    keep = [ln for ln in text.splitlines() if not ln.strip().startswith(("frequency", "end"))]
    return "\n".join(keep + [f"frequency = {f0} {f1} {nd}", "end"]) + "\n"


def _models(lib, texts):
    ms = []
    for t in texts:
        m = Peec.from_deck_text(t, lib)
        m.setup()
        ms.append(m)
    return ms


def test_batch_random_wires_match_oracle(lib):
    texts = [decks.random_wire_deck(1000 + s, nmin=20, nmax=45) for s in range(6)]
    ms = _models(lib, texts)
    print(f"[batch] random wires n = {[m.p.nseg for m in ms]}", flush=True)
    b = Batch(ms, 10, 4)
    zs, st = b.sweep()
    print(f"  jobs {st.jobs} chunks {st.chunks} nb {st.nb} wpc {st.wg_per_cu} order {st.order}", flush=True)
    assert st.jobs == 24 and st.direct_structures == 0
    for t, z in zip(texts, zs):
        zo = oracle_zmat(t, 10, 4)
        r = _rel(z, zo)
        print(f"  rel {r:.2e}", flush=True)
        assert r <= 1e-8, r
    b.close()
    for m in ms:
        m.close()


def test_batch_element_fill_path_matches_oracle(lib, monkeypatch):
    """Structures whose records cannot be built are filled element by element between the two
    graphs (forced here for every structure)."""
    monkeypatch.setenv("PEEC_BATCH_DIRECT", "1")
    texts = [decks.random_wire_deck(3000 + s, nmin=20, nmax=40) for s in range(3)]
    ms = _models(lib, texts)
    print(f"[batch] element-fill path n = {[m.p.nseg for m in ms]}", flush=True)
    b = Batch(ms, 5, 3)
    zs, st = b.sweep()
    assert st.direct_structures == 3
    for t, z in zip(texts, zs):
        r = _rel(z, oracle_zmat(t, 5, 3))
        print(f"  rel {r:.2e}", flush=True)
        assert r <= 1e-8, r
    b.close()
    for m in ms:
        m.close()


def test_batch_mixed_decks_match_plain(lib):
    texts = [
        _sweep(deck_text("monopole_gp.peec"), 1e8, 1e9, 4),
        _sweep(deck_text("twoport_dist.peec"), 1e8, 1e9, 4),
        decks.plate_deck(n=8, ndiv=4),
        _sweep(deck_text("rlc_series.peec"), 1e6, 1e8, 4),
        decks.dipole_deck(f0=1e8, f1=6e9, ndiv=4),
        decks.random_wire_deck(77, nmin=60, nmax=90, ndiv=4),
    ]
    ms = _models(lib, texts)
    nf = 5
    print(f"[batch] mixed decks nunknown = {[m.p.nunknown for m in ms]}, {nf} frequencies", flush=True)
    b = Batch(ms, 0, nf)
    zs, st = b.sweep()
    assert st.direct_pairs > 0   # the wide-band dipole: quadrature order changes inside the band
    for m, z in zip(ms, zs):
        zp, _ = m.sweep(0, nf, PEEC_FORM_PLAIN)
        r = _rel(z, zp)
        print(f"  {m.p.nunknown} unknowns: rel {r:.2e}", flush=True)
        assert r <= 1e-8, r
    b.close()
    for m in ms:
        m.close()


def test_batch_replay_and_autotune(lib):
    texts = [decks.random_wire_deck(2000 + s, nmin=20, nmax=400) for s in range(12)]
    ms = _models(lib, texts)
    print(f"[batch] uneven wires n = {[m.p.nseg for m in ms]}", flush=True)
    b = Batch(ms, 0, 8)
    st0 = b.tune()
    trials = b.trials()
    assert len(trials) == st0.tune_trials == 24
    best = min(trials, key=lambda t: t[4])
    assert (st0.nb, st0.tile_cols, st0.wg_per_cu, st0.order) == best[:4]
    z1, st1 = b.sweep()
    z2, _ = b.sweep()
    for a, c in zip(z1, z2):
        assert np.array_equal(a, c)
    # spot check against the plain sweep of two of them
    for k in (0, 5):
        zp, _ = ms[k].sweep(0, 8, PEEC_FORM_PLAIN)
        r = _rel(z1[k], zp)
        print(f"  structure {k}: rel {r:.2e}", flush=True)
        assert r <= 1e-8, r
    b.close()
    for m in ms:
        m.close()


def test_batch_precision_form(lib):
    """FP32-stored batched factorisation with FP64 refinement inside each workgroup: every port
    solution refined to a componentwise backward error <= 1e-10 without an FP64 redo, Z within
    1e-10 of the FP64 batched sweep and 1e-8 of the oracle."""
    texts = [decks.random_wire_deck(4000 + s, nmin=20, nmax=400) for s in range(8)]
    ms = _models(lib, texts)
    print(f"[batch-precision] n = {[m.p.nseg for m in ms]}", flush=True)
    b = Batch(ms, 0, 6)
    z64, _ = b.sweep()
    b.set_form(PEEC_FORM_PRECISION)
    z32, st = b.sweep()
    print(f"  refine sweeps {st.refine_sweeps}, worst backward error {st.berr_max:.2e}, "
          f"FP64 redos {st.fp64_fallbacks}", flush=True)
    assert st.berr_max <= 1e-10 and st.fp64_fallbacks == 0 and st.refine_sweeps > 0
    for a, c in zip(z32, z64):
        assert _rel(a, c) <= 1e-10, _rel(a, c)
    for k in (0, 3):
        r = _rel(z32[k], oracle_zmat(texts[k], 0, 6))
        print(f"  structure {k} vs oracle: {r:.2e}", flush=True)
        assert r <= 1e-8
    b.close()
    for m in ms:
        m.close()
