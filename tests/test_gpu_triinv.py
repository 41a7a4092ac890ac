"""Explicit inverses of LU factors on the GPU (dlu_tri_inverse / slu_tri_inverse), the
preconditioner of the accelerated sweep applies them as two triangular products.

A packed LU buffer (unit lower L below the diagonal, upper U on and above it) is inverted
triangle by triangle, by recursive blocking on rocBLAS triangular multiplies (default) and on the
hand-written GEMM (PEEC_GEMM=hip); L Linv and U Uinv must be the identity to the storage precision,
and Linv / Uinv must be triangular (the other triangle exactly zero).  A wrong sign, a transposed
operand or a product applied on the wrong side in one level of the recursion leaves O(1) errors
in the off-diagonal blocks.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _packed(n, seed):
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    a[np.diag_indices(n)] = 1.5 + rng.random(n) + 0.5j * rng.standard_normal(n)
    return a


@pytest.mark.parametrize("gemm", ["rocblas", "hip"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("n", [64, 65, 200, 700])
def test_triangular_inverses(lib, n, prec, gemm, monkeypatch):
    if gemm == "hip":
        monkeypatch.setenv("PEEC_GEMM", "hip")
    print(f"[triinv] n={n} {prec} gemm={gemm}", flush=True)
    a = _packed(n, 31 + n)
    L = np.tril(a, -1) + np.eye(n)
    U = np.triu(a)
    cdt = torch.complex128 if prec == "f64" else torch.complex64
    dev = torch.device("cuda:0")
    lu = torch.from_numpy(a.astype(np.complex128 if prec == "f64" else np.complex64)).to(dev)
    out = torch.zeros((n, n), dtype=cdt, device=dev)
    ns = ((n + 1) // 2 + 64) ** 2
    scr = torch.zeros(ns, dtype=cdt, device=dev)
    dinv = torch.zeros(((n + 63) // 64) * 64 * 64, dtype=cdt, device=dev)
    fn = lib.dlu_tri_inverse if prec == "f64" else lib.slu_tri_inverse
    fn.restype = None
    tol = 1e-12 if prec == "f64" else 2e-5
    for lower, T in ((1, L), (0, U)):
        torch.cuda.synchronize()
        fn(None, C.c_int(n), C.c_void_p(lu.data_ptr()), C.c_int(n), C.c_int(lower), C.c_void_p(out.data_ptr()),
           C.c_void_p(scr.data_ptr()), C.c_void_p(dinv.data_ptr()))
        torch.cuda.synchronize()
        inv = out.cpu().numpy().astype(np.complex128)
        other = np.triu(inv, 1) if lower else np.tril(inv, -1)
        assert not other.any(), "entries outside the triangle"
        err = np.abs(T @ inv - np.eye(n)).max()
        print(f"  {'L' if lower else 'U'}: max |T Tinv - I| = {err:.2e}", flush=True)
        assert err <= tol, err
