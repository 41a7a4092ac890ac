"""Python access to the oracle (oracle/liboracle.so, the CPU restatement of the reference's
fill + dense solve).  Test infrastructure: used by tests, smoke() and bench.py's cpu_baseline.

The peec_t it works on is built by the product's host-side deck reader, geometry builder and
MNA numbering (pure host code, checked bit-for-bit against the reference's on the sample decks
in test_geometry.py); no device call is made here.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

from openpeec_amd import Peec, abi

ROOT = Path(__file__).resolve().parents[1]
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"
_orc = None


def oracle_lib() -> C.CDLL:
    global _orc
    if _orc is None:
        if not ORACLE_LIB.exists():
            subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "liboracle.so"], check=True)
        lib = C.CDLL(str(ORACLE_LIB))
        PP = C.POINTER(abi.peec_t)
        PS = C.POINTER(abi.seg_t)
        PZ = C.POINTER(abi.d_complex_t)
        lib.orc_sweep.argtypes = [PP, C.c_int, C.c_int, PZ]
        lib.orc_sweep.restype = C.c_int
        lib.orc_lp_fill.argtypes = [PP, C.c_double, PZ]
        lib.orc_lp_fill.restype = None
        lib.orc_pot_fill.argtypes = [PP, C.c_double, PZ, PZ]
        lib.orc_pot_fill.restype = C.c_int
        lib.orc_lp_entry.argtypes = [PP, C.c_int, C.c_int, C.c_double]
        lib.orc_lp_entry.restype = abi.d_complex_t
        lib.orc_lu_decomp.argtypes = [C.c_int, PZ, C.POINTER(C.c_int)]
        lib.orc_lu_decomp.restype = C.c_int
        lib.orc_lu_solve.argtypes = [C.c_int, PZ, C.POINTER(C.c_int), PZ]
        lib.orc_lu_solve.restype = None
        lib.orc_fill_rows.argtypes = [PP, C.c_double, C.c_int, C.c_int, C.c_int]
        lib.orc_fill_rows.restype = None
        lib.orc_neumann_self.argtypes = [C.c_double, C.c_double]
        lib.orc_neumann_self.restype = C.c_double
        lib.orc_neumann_pair.argtypes = [PS, PS, C.c_double, C.c_double]
        lib.orc_neumann_pair.restype = C.c_double
        lib.orc_pair_k.argtypes = [PS, PS, C.c_double, C.c_double, C.c_double]
        lib.orc_pair_k.restype = abi.d_complex_t
        lib.orc_self_k.argtypes = [PS, C.c_double, C.c_double]
        lib.orc_self_k.restype = abi.d_complex_t
        lib.orc_rect_potential.argtypes = [PS, C.POINTER(C.c_double)]
        lib.orc_rect_potential.restype = C.c_double
        lib.orc_poly_potential.argtypes = [PS, C.POINTER(C.c_double)]
        lib.orc_poly_potential.restype = C.c_double
        lib.orc_bar_pair.argtypes = [PS, PS]
        lib.orc_bar_pair.restype = C.c_double
        _orc = lib
    return _orc


def oracle_peec(text: str) -> Peec:
    """A numbered peec_t (geometry + capacitive cells + MNA numbering, no fills)."""
    m = Peec.from_deck_text(text)
    m.wire_build()
    if m.p.capacitance and m.p.nchg > 0:
        assert m.lib.pot_cells(C.byref(m.p)) == 0
    assert m.lib.mna_numbering(C.byref(m.p), None) == 0
    return m


def oracle_zmat(text: str, ifreq0: int = 0, nfreq: int | None = None) -> np.ndarray:
    m = oracle_peec(text)
    nf = m.p.nfreq - ifreq0 if nfreq is None else nfreq
    np_ = m.p.nport
    z = np.zeros((nf, np_, np_), dtype=np.complex128)
    rc = oracle_lib().orc_sweep(C.byref(m.p), ifreq0, nf, z.ctypes.data_as(C.POINTER(abi.d_complex_t)))
    m.close()
    if rc != 0:
        raise RuntimeError(f"oracle sweep failed ({rc})")
    return z


def oracle_lp(m: Peec, kw: float) -> np.ndarray:
    n = m.p.nseg
    lp = np.zeros((n, n), dtype=np.complex128)
    oracle_lib().orc_lp_fill(C.byref(m.p), kw, lp.ctypes.data_as(C.POINTER(abi.d_complex_t)))
    return lp


def oracle_c(m: Peec, kw: float) -> np.ndarray:
    nc = m.p.ncell
    c = np.zeros((nc, nc), dtype=np.complex128)
    rc = oracle_lib().orc_pot_fill(C.byref(m.p), kw, None, c.ctypes.data_as(C.POINTER(abi.d_complex_t)))
    if rc != 0:
        raise RuntimeError("oracle: singular P")
    return c
