"""Test-side access to the checkers: the reference build (oracle/_ref/libpeec_ref.so, compiled
from OpenPEEC's own C sources where they are present) and the oracle restatement
(oracle/liboracle.so).  Test infrastructure only; the product never loads these.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from openpeec_amd import abi, cview

ROOT = Path(__file__).resolve().parents[1]
REF_LIB = ROOT / "oracle" / "_ref" / "libpeec_ref.so"
ORACLE_LIB = ROOT / "oracle" / "liboracle.so"

_libc = C.CDLL(None)
_libc.fmemopen.restype = C.c_void_p
_libc.fmemopen.argtypes = [C.c_void_p, C.c_size_t, C.c_char_p]
_libc.fopen.restype = C.c_void_p
_libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
_libc.fclose.argtypes = [C.c_void_p]
_libc.fflush.argtypes = [C.c_void_p]

_ref = None


def ref_available() -> bool:
    return REF_LIB.exists()


def ref_lib() -> C.CDLL:
    global _ref
    if _ref is None:
        _ref = C.CDLL(str(REF_LIB), mode=os.RTLD_LOCAL)
        abi.declare(_ref, reference=True)
    return _ref


def devnull():
    return _libc.fopen(b"/dev/null", b"w")


class RefPeec:
    """A peec_t driven by the reference build (same struct layout as the product)."""

    def __init__(self, text: str):
        self.lib = ref_lib()
        self.p = abi.peec_t()
        raw = text.encode()
        buf = C.create_string_buffer(raw)
        fp = _libc.fmemopen(buf, len(raw), b"r")
        rc = self.lib.input_data(fp, C.byref(self.p))
        _libc.fclose(fp)
        self._open = True
        self.log = devnull()
        if rc != 0:
            raise ValueError("deck rejected by the reference input_data")

    def close(self):
        if self._open:
            self.lib.peec_free(C.byref(self.p))
            _libc.fclose(self.log)
            self._open = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def wire_build(self):
        assert self.lib.wire_build(C.byref(self.p), self.log) == 0

    def lp_fill(self, f: float) -> np.ndarray:
        self.lib.lp_fill(C.byref(self.p), f, self.log)
        n = self.p.nseg
        return cview(self.p.lp, (n, n)).copy()

    def pot_fill(self, f: float) -> np.ndarray:
        assert self.lib.pot_fill(C.byref(self.p), f, self.log) == 0
        n = self.p.ncell
        return cview(self.p.cmat, (n, n)).copy()

    def setup_like_main(self):
        """The reference main's setup: geometry, static fills, numbering (main.c:78-94)."""
        self.wire_build()
        if not self.p.compress:
            self.lib.lp_fill(C.byref(self.p), 0.0, self.log)
        assert self.lib.pot_fill(C.byref(self.p), 0.0, self.log) == 0
        assert self.lib.mna_numbering(C.byref(self.p), self.log) == 0

    def solve(self):
        assert self.lib.solve(C.byref(self.p), self.log) == 0

    def zmat(self) -> np.ndarray:
        np_ = self.p.nport
        return cview(self.p.zmat, (self.p.nfreq, np_, np_)).copy()

    def zin(self) -> np.ndarray:
        return cview(self.p.zin, (self.p.nport, self.p.nfreq)).copy()


def seg_array(p: abi.peec_t, which: str = "seg") -> list:
    n = p.nseg if which == "seg" else p.nchg
    arr = getattr(p, which)
    return [arr[i] for i in range(n)]


def rel_err(a: np.ndarray, b: np.ndarray) -> float:
    """max |a - b| / |b| over entries (complex modulus), 0/0 counted as 0."""
    a = np.asarray(a)
    b = np.asarray(b)
    d = np.abs(a - b)
    s = np.abs(b)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(s > 0, d / s, np.where(d > 0, np.inf, 0.0))
    return float(r.max()) if r.size else 0.0
