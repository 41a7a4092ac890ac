"""Dense complex LU on the GPU (lu_decomp / lu_solve through the C ABI).

Checks: the solution of A x = b against numpy's LAPACK solve (backward error), the factors'
reconstruction P A = L U, the pivot sequence against the oracle's restatement of lu_decomp
(lu.c, pinned to the reference and to LAPACK by test_oracle.py; same |re|+|im| rule, random
matrices have no near-ties), and the singular-pivot return code.
"""
import ctypes as C

import numpy as np
import pytest

from openpeec_amd import abi
from oracle_py import oracle_lib

pytestmark = pytest.mark.gpu


def _zptr(a):
    return a.ctypes.data_as(C.POINTER(abi.d_complex_t))


def _rand(n, seed):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex128)


@pytest.mark.parametrize("gemm", ["rocblas", "hip"])
@pytest.mark.parametrize("n", [1, 2, 7, 64, 65, 130, 300, 777])
def test_lu_solve_matches_lapack(lib, n, gemm, monkeypatch):
    """Both trailing-update GEMMs: rocBLAS zgemm (default) and the hand-written k_zgemm."""
    if gemm == "hip":
        monkeypatch.setenv("PEEC_GEMM", "hip")
    print(f"[lu] n={n} gemm={gemm}", flush=True)
    a0 = _rand(n, 1000 + n)
    b0 = (np.arange(n) + 1j).astype(np.complex128)
    a = a0.copy()
    piv = np.zeros(n, dtype=np.int32)
    info = lib.lu_decomp(n, _zptr(a), piv.ctypes.data_as(C.POINTER(C.c_int)))
    assert info == -1
    b = b0.copy()
    lib.lu_solve(n, _zptr(a), piv.ctypes.data_as(C.POINTER(C.c_int)), _zptr(b))
    x_ref = np.linalg.solve(a0, b0)
    berr = np.abs(a0 @ b - b0).max() / (np.abs(a0).max() * np.abs(b).max() * n)
    assert berr < 1e-15
    assert np.abs(b - x_ref).max() / np.abs(x_ref).max() < 1e-10
    # P A = L U with the recorded interchanges
    L = np.tril(a, -1) + np.eye(n)
    U = np.triu(a)
    pa = a0.copy()
    for k in range(n):
        if piv[k] != k:
            pa[[k, piv[k]]] = pa[[piv[k], k]]
    assert np.abs(L @ U - pa).max() / np.abs(a0).max() < 1e-12


@pytest.mark.parametrize("n", [50, 200])
def test_lu_pivots_match_reference(lib, n):
    a0 = _rand(n, 7 * n)
    a = a0.copy()
    piv = np.zeros(n, dtype=np.int32)
    assert lib.lu_decomp(n, _zptr(a), piv.ctypes.data_as(C.POINTER(C.c_int))) == -1
    r = a0.copy()
    rpiv = np.zeros(n, dtype=np.int32)
    assert oracle_lib().orc_lu_decomp(n, _zptr(r), rpiv.ctypes.data_as(C.POINTER(C.c_int))) == -1
    assert np.array_equal(piv, rpiv)
    assert np.abs(a - r).max() / np.abs(r).max() < 1e-11


def test_lu_singular_returns_row(lib):
    n = 100
    a = _rand(n, 5)
    a[:, 37] = 0          # column 37 becomes exactly zero below the already eliminated rows
    a[37:, 37] = 0
    piv = np.zeros(n, dtype=np.int32)
    info = lib.lu_decomp(n, _zptr(a), piv.ctypes.data_as(C.POINTER(C.c_int)))
    assert info == 37


def test_gmres_solve_with_a_nearby_lu(lib):
    """gmres_solve (iterative.c:230-252): one GMRES(60) cycle on A preconditioned by the LU of a
    nearby matrix converges to tol |b| and returns the solution (checked against LAPACK); with the
    LU of an unrelated matrix the cycle fails, returns -1 and leaves b untouched."""
    n = 300
    a = _rand(n, 21)
    near = a + 1e-3 * _rand(n, 22)
    b0 = (np.arange(n) * 0.01 + 1j).astype(np.complex128)
    for m0, expect in ((near, True), (_rand(n, 23), False)):
        alu = m0.copy()
        piv = np.zeros(n, dtype=np.int32)
        assert lib.lu_decomp(n, _zptr(alu), piv.ctypes.data_as(C.POINTER(C.c_int))) == -1
        b = b0.copy()
        it = lib.gmres_solve(n, _zptr(a), _zptr(alu), piv.ctypes.data_as(C.POINTER(C.c_int)), _zptr(b), 1e-10)
        if expect:
            assert 1 <= it <= 60
            x = np.linalg.solve(a, b0)
            assert np.abs(a @ b - b0).max() <= 1e-9 * np.abs(b0).max() * 10
            assert np.abs(b - x).max() <= 1e-7 * np.abs(x).max()
        else:
            assert it == -1
            assert np.array_equal(b, b0)
