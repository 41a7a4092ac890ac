"""Precision form: LU of an FP32 copy of the MNA matrix (FP64 panel arithmetic and pivot search,
FP32 MFMA trailing updates) with FP64 iterative refinement against the FP64 matrix.

  * lu_decomp32 returns factors with P A = L U to the backward error of FP32 elimination,
    ||P A - L U||_F <= n eps32 ||A||_F (eps32 = 2^-24); a wrong MFMA accumulator layout or a
    dropped update leaves O(1) residuals, so this pins the FP32 trailing-update kernel;
  * the refined sweep drives the componentwise (Oettli-Prager) backward error of every port
    solution, max_i |r_i| / (|A| |x| + |b|)_i, below 1e-10 without falling back to an FP64
    factorisation, and its Z agrees with the plain FP64 sweep and the oracle within 1e-6 relative.
"""
import ctypes as C

import numpy as np
import pytest

from conftest import deck_text
from openpeec_amd import PEEC_FORM_PLAIN, PEEC_FORM_PRECISION, Peec, abi, decks
from oracle_py import oracle_zmat

pytestmark = pytest.mark.gpu

EPS32 = 2.0 ** -24


def _zptr(a):
    return a.ctypes.data_as(C.POINTER(abi.d_complex_t))


def _split(lu, piv):
    n = lu.shape[0]
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    perm = np.arange(n)
    for k in range(n):
        if piv[k] != k:
            perm[[k, piv[k]]] = perm[[piv[k], k]]
    return L, U, perm


@pytest.mark.parametrize("gemm", ["rocblas", "hip"])
@pytest.mark.parametrize("n,seed,boost", [(64, 1, 0.0), (200, 2, 0.0), (333, 3, 5.0), (513, 4, 0.0)])
def test_lu_decomp32_backward_error(lib, n, seed, boost, gemm, monkeypatch):
    """Both trailing-update GEMMs: rocBLAS cgemm (default) and the hand-written k_cgemm32."""
    if gemm == "hip":
        monkeypatch.setenv("PEEC_GEMM", "hip")
    print(f"[lu32] n={n} seed={seed} gemm={gemm}", flush=True)
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)) + boost * np.eye(n)
    a = np.ascontiguousarray(a, dtype=np.complex128)
    lu = a.copy()
    piv = np.zeros(n, dtype=np.int32)
    assert lib.lu_decomp32(n, _zptr(lu), piv.ctypes.data_as(C.POINTER(C.c_int))) == -1
    L, U, perm = _split(lu, piv)
    err = np.linalg.norm(a[perm] - L @ U) / np.linalg.norm(a)
    assert err <= n * EPS32, err
    # the factors are FP32 values
    assert np.array_equal(lu.real.astype(np.float32).astype(np.float64), lu.real)
    # and they solve: FP64 refinement with these factors converges to FP64 accuracy
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)

    def apply_inv(r):
        y = np.linalg.solve(L, r[perm])
        return np.linalg.solve(U, y)

    x = apply_inv(b)
    for _ in range(8):
        x = x + apply_inv(b - a @ x)
    berr = np.max(np.abs(b - a @ x)) / (np.max(np.abs(a).sum(axis=1)) * np.max(np.abs(x)) + np.max(np.abs(b)))
    assert berr <= 1e-14, berr


@pytest.mark.parametrize("text_fn", [
    lambda: decks.dipole_deck(),
    lambda: deck_text("monopole_gp.peec"),
    lambda: deck_text("twoport_dist.peec"),
    lambda: decks.plate_deck(n=10, ndiv=6),
])
def test_precision_matches_oracle(lib, text_fn):
    text = text_fn()
    print(f"[precision] {text.splitlines()[1]}", flush=True)
    m = Peec.from_deck_text(text, lib)
    m.setup()
    z, st = m.sweep(0, m.nfreq, PEEC_FORM_PRECISION)
    zo = oracle_zmat(text)
    rel = float(np.max(np.abs(z - zo) / np.abs(zo)))
    print(f"  rel {rel:.2e} berr {st.berr:.2e} sweeps {st.refine_its} fallbacks {st.fp64_fallbacks}", flush=True)
    assert rel <= 1e-6, rel
    assert st.berr <= 1e-10
    assert st.fp64_fallbacks == 0
    m.close()


def test_precision_matches_plain_on_plate(lib):
    text = decks.plate_deck(n=24, ndiv=40)
    print("[precision] plate 24x40", flush=True)
    m = Peec.from_deck_text(text, lib)
    m.setup()
    zp, _ = m.sweep(0, 6, PEEC_FORM_PLAIN)
    zq, st = m.sweep(0, 6, PEEC_FORM_PRECISION)
    rel = float(np.max(np.abs(zq - zp) / np.abs(zp)))
    print(f"  rel {rel:.2e} berr {st.berr:.2e} sweeps {st.refine_its} fallbacks {st.fp64_fallbacks}", flush=True)
    assert st.berr <= 1e-10
    assert st.fp64_fallbacks == 0
    assert st.refine_its >= 6
    assert rel <= 1e-6, rel
    m.close()
