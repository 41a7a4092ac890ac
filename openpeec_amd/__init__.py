"""openpeec_amd — MI355X (gfx950) implementation of OpenPEEC's per-frequency full-wave path.

The product is the C-ABI library libpeec_mi355x.so (C host code + hand-written HIP kernels);
this package loads it and offers a thin Python view of the peec.h interface:

    from openpeec_amd import Peec
    p = Peec.from_deck_text(open("dipole.peec").read())
    p.setup()                       # geometry, capacitive cells, MNA numbering
    z, stats = p.sweep(0, p.nfreq)  # device-resident fill + LU + port solves

There is no CPU compute path: if the library is missing or no HIP device is present, the
compute calls fail loudly.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

from . import abi

__all__ = ["load_library", "Peec", "abi", "LIB_PATH"]

LIB_PATH = Path(__file__).resolve().parent / "libpeec_mi355x.so"
_lib = None

PEEC_FORM_PLAIN = 0
PEEC_FORM_MIXED = 1
PEEC_FORM_PRECISION = 2
PEEC_FORM_LEARNED = 3


def load_library(build_if_missing: bool = True) -> C.CDLL:
    """Load libpeec_mi355x.so (built in-tree); build it first only when it does not exist."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        if not build_if_missing:
            raise RuntimeError(f"{LIB_PATH} is not built (run __graft_entry__.build())")
        from .build import build
        build()
    lib = C.CDLL(str(LIB_PATH), mode=os.RTLD_LOCAL)
    abi.declare(lib)
    if lib.peec_sizeof_peec() != C.sizeof(abi.peec_t) or lib.peec_sizeof_seg() != C.sizeof(abi.seg_t):
        raise RuntimeError("ctypes mirror of peec_t/seg_t does not match the library layout")
    _lib = lib
    return lib


def cview(ptr, shape) -> np.ndarray:
    """numpy complex128 view (no copy) of a d_complex_t* of the given shape."""
    n = int(np.prod(shape))
    if not ptr or n == 0:
        return np.zeros(shape, dtype=np.complex128)
    raw = C.cast(ptr, C.POINTER(C.c_double))
    return np.ctypeslib.as_array(raw, shape=(n * 2,)).view(np.complex128).reshape(shape)


class Peec:
    """Owns one peec_t filled from a deck; frees it (host arrays and device mirror) on close."""

    def __init__(self, lib: C.CDLL | None = None):
        self.lib = lib if lib is not None else load_library()
        self.p = abi.peec_t()
        self._open = False

    # ---- construction
    @classmethod
    def from_deck_text(cls, text: str, lib: C.CDLL | None = None) -> "Peec":
        obj = cls(lib)
        rc = obj.lib.peec_load_deck_string(text.encode(), C.byref(obj.p))
        obj._open = True
        if rc != 0:
            obj.close()
            raise ValueError("deck rejected by input_data")
        return obj

    @classmethod
    def from_deck_file(cls, path: str, lib: C.CDLL | None = None) -> "Peec":
        return cls.from_deck_text(Path(path).read_text(), lib)

    def close(self) -> None:
        if self._open:
            self.lib.peec_free(C.byref(self.p))
            self._open = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- setup / fills
    def setup(self) -> None:
        if self.lib.peec_setup(C.byref(self.p), None) != 0:
            raise RuntimeError("peec_setup failed")

    def wire_build(self) -> None:
        if self.lib.wire_build(C.byref(self.p), None) != 0:
            raise RuntimeError("wire_build failed")

    def lp_fill(self, f: float) -> np.ndarray:
        self.lib.lp_fill(C.byref(self.p), f, None)
        n = self.p.nseg
        return cview(self.p.lp, (n, n)).copy()

    def pot_fill(self, f: float) -> np.ndarray:
        if self.lib.pot_fill(C.byref(self.p), f, None) != 0:
            raise RuntimeError("pot_fill failed (singular P)")
        n = self.p.ncell
        return cview(self.p.cmat, (n, n)).copy()

    def fill_matrices(self, f: float, mode: int = 0):
        """(Lp, P) at f from the device fill: mode 0 element by element, 1 k-factored."""
        ns, nc = self.p.nseg, self.p.ncell if self.p.capacitance else 0
        lp = np.zeros((ns, ns), dtype=np.complex128)
        P = np.zeros((nc, nc), dtype=np.complex128)
        rc = self.lib.peec_fill_matrices(C.byref(self.p), f, mode, lp.ctypes.data_as(C.POINTER(abi.d_complex_t)),
                                         P.ctypes.data_as(C.POINTER(abi.d_complex_t)) if nc else None)
        if rc != 0:
            raise RuntimeError(f"peec_fill_matrices failed (code {rc})")
        return lp, P

    def kw(self, f: float) -> float:
        return 2 * np.pi * f / 2.99792458e8 if self.p.retardation else 0.0

    # ---- sweeps
    def freq(self, ifreq: int) -> float:
        p = self.p
        if p.nfreq <= 1:
            return p.f0
        if p.flog:
            return p.f0 * (p.f1 / p.f0) ** (ifreq / (p.nfreq - 1))
        return p.f0 + (p.f1 - p.f0) * ifreq / (p.nfreq - 1)

    def sweep(self, ifreq0: int, nfreq: int, form: int = PEEC_FORM_PLAIN, batch: int = 4):
        np_ = self.p.nport
        z = np.zeros((nfreq, np_, np_), dtype=np.complex128)
        st = abi.peec_stats_t()
        zp = z.ctypes.data_as(C.POINTER(abi.d_complex_t))
        rc = self.lib.peec_sweep_ex(C.byref(self.p), ifreq0, nfreq, form, batch, zp, C.byref(st))
        if rc != 0:
            raise RuntimeError(f"peec_sweep failed (code {rc})")
        return z, st

    def solve(self) -> None:
        if self.lib.solve(C.byref(self.p), None) != 0:
            raise RuntimeError("solve failed")

    # ---- views
    @property
    def nseg(self) -> int:
        return self.p.nseg

    @property
    def nfreq(self) -> int:
        return self.p.nfreq

    @property
    def nunknown(self) -> int:
        return self.p.nunknown

    def zin(self) -> np.ndarray:
        return cview(self.p.zin, (self.p.nport, self.p.nfreq)).copy()

    def zmat(self) -> np.ndarray:
        np_ = self.p.nport
        return cview(self.p.zmat, (self.p.nfreq, np_, np_)).copy()

    def smat(self) -> np.ndarray:
        np_ = self.p.nport
        return cview(self.p.smat, (self.p.nfreq, np_, np_)).copy()


class Batch:
    """Many set-up structures swept together (peec_batch_*): every (structure, frequency) pair is
    one system of the persistent batched kernels.  Frequencies ifreq0 .. ifreq0 + nfreq - 1 of
    each structure's own sweep."""

    def __init__(self, models: "list[Peec]", ifreq0: int, nfreq: int):
        if not models:
            raise ValueError("no structures")
        self.lib = models[0].lib
        self.models = models
        self.nfreq = nfreq
        arr = (C.POINTER(abi.peec_t) * len(models))(*[C.pointer(m.p) for m in models])
        self._arr = arr
        self.h = self.lib.peec_batch_create(arr, len(models), ifreq0, nfreq)
        if not self.h:
            raise RuntimeError("peec_batch_create rejected the structures")

    def set_form(self, form: int) -> None:
        """PEEC_FORM_PLAIN (FP64 factorisation) or PEEC_FORM_PRECISION (FP32 + FP64 refinement)."""
        if self.lib.peec_batch_set_form(self.h, form) != 0:
            raise ValueError("batched sweep: form must be plain or precision")

    def tune(self) -> abi.peec_batch_stats_t:
        st = abi.peec_batch_stats_t()
        if self.lib.peec_batch_tune(self.h, C.byref(st)) != 0:
            raise RuntimeError("peec_batch_tune failed")
        return st

    def trials(self) -> "list[tuple[int, int, int, int, float]]":
        """(panel width, tile columns, workgroups per CU, order, LU ms) of every candidate."""
        cap = 64
        nb, tc, wpc, order = (C.c_int * cap)(), (C.c_int * cap)(), (C.c_int * cap)(), (C.c_int * cap)()
        ms = (C.c_double * cap)()
        n = self.lib.peec_batch_trials(self.h, nb, tc, wpc, order, ms, cap)
        return [(nb[k], tc[k], wpc[k], order[k], ms[k]) for k in range(min(n, cap))]

    def sweep(self):
        """[Z per structure (nfreq x nport x nport)], stats."""
        zs = [np.zeros((self.nfreq, m.p.nport, m.p.nport), dtype=np.complex128) for m in self.models]
        ptrs = (C.POINTER(abi.d_complex_t) * len(zs))(*[z.ctypes.data_as(C.POINTER(abi.d_complex_t)) for z in zs])
        st = abi.peec_batch_stats_t()
        rc = self.lib.peec_batch_sweep(self.h, ptrs, C.byref(st))
        if rc != 0:
            raise RuntimeError(f"peec_batch_sweep failed (code {rc})")
        return zs, st

    def close(self) -> None:
        if self.h:
            self.lib.peec_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
