"""The reference program's whole flow through the library's C ABI -- setup, solve(), then every
result writer (output_zin / spara tables, zin.csv, peec.sNp, dist.csv, far.csv, pw.csv,
tran.csv, the sources' node voltages in peec.log) -- against the files the reference's own
program wrote for the same decks (tests/golden/output_vectors.npz, make_output_golden.py).

Bars: 1e-9 relative (to each column's largest magnitude) for the dense plain path; the far
field's radiation sums run in a different order on the device, so its fields are held to 1e-9 of
the pattern's largest field and its D / G (printed to 1e-4 dB) to 2e-4 dB where the pattern is
within 60 dB of its peak; node voltages are printed with 5 significant digits.  The compressed
deck (ACA 1e-10, GMRES to 1e-10) and the acceleration = 1 deck (GMRES to 1e-10 preconditioned by
a kept LU) are iterative on both sides and are held to 1e-7.
"""
import ctypes as C
import os
import tempfile
from pathlib import Path

import numpy as np
import pytest

from conftest import DECKS
from openpeec_amd import Peec
from refharness import _libc

pytestmark = pytest.mark.gpu

GOLD = np.load(DECKS.parent / "output_vectors.npz")
sys_path = str(DECKS.parent)
import sys  # noqa: E402

sys.path.insert(0, sys_path)
from make_output_golden import _node_voltages, _table, _touchstone, output_decks  # noqa: E402

ITERATIVE = {"compress_dist": 1e-7, "dipole_accel": 1e-7}


def _close(a, b, rtol):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    scale = np.maximum(np.abs(b).max(axis=0), 1e-300) if b.ndim == 2 else max(np.abs(b).max(), 1e-300)
    return float(np.max(np.abs(a - b) / scale)) if a.size else 0.0


def run_library(lib, text: str, work: Path) -> dict:
    m = Peec.from_deck_text(text, lib)
    cwd = os.getcwd()
    os.chdir(work)
    try:
        log = _libc.fopen(b"peec.log", b"w")
        m.setup()
        assert lib.solve(C.byref(m.p), log) == 0
        lib.output_zin(C.byref(m.p), log)
        lib.output_spara(C.byref(m.p), log)
        assert lib.output_csv(C.byref(m.p), b"zin.csv") == 0
        assert lib.output_touchstone(C.byref(m.p)) == 0
        assert lib.output_dist(C.byref(m.p), b"dist.csv") == 0
        assert lib.output_far(C.byref(m.p), b"far.csv", log) == 0
        assert lib.output_pw(C.byref(m.p), b"pw.csv", log) == 0
        assert lib.output_tran(C.byref(m.p), b"tran.csv", log) == 0
        _libc.fclose(log)
    finally:
        os.chdir(cwd)
        m.close()
    out = {}
    for fn, codes in (("zin.csv", None), ("far.csv", None), ("pw.csv", None), ("dist.csv", {"I": 0, "Ipw": 1, "Q": 2}),
                      ("tran.csv", {"X": 0, "S": 1, "V": 2})):
        if (work / fn).exists():
            out[fn] = _table(work / fn, codes)
    for snp in work.glob("peec.s*p"):
        out["snp"] = _touchstone(snp)
    nv = _node_voltages(work / "peec.log")
    if nv.size:
        out["vnode"] = nv
    return out


@pytest.mark.parametrize("name", sorted(output_decks()))
def test_result_files_match_reference(lib, name):
    print(f"[outputs] {name}", flush=True)
    want = {k.split("/")[-1]: GOLD[k] for k in GOLD.files if k.startswith(f"out/{name}/")}
    with tempfile.TemporaryDirectory() as tmp:
        got = run_library(lib, output_decks()[name], Path(tmp))
    assert sorted(got) == sorted(want), (sorted(got), sorted(want))
    rtol = ITERATIVE.get(name, 1e-9)
    for fn, w in want.items():
        g = got[fn]
        if fn == "far.csv":
            assert np.array_equal(g[:, :3], w[:, :3])
            field = w[:, 3:7]
            e = np.abs(g[:, 3:7] - field).max() / np.abs(field).max()
            assert e <= rtol, (name, fn, e)
            live = w[:, 7] > w[:, 7].max() - 60
            assert np.abs(g[live, 7:9] - w[live, 7:9]).max() <= 2e-4, (name, fn)
        elif fn == "vnode":
            assert np.array_equal(g[:, 0], w[:, 0])
            assert np.abs(g[:, 1:] - w[:, 1:]).max() <= 1e-4 * np.abs(w[:, 1:]).max() + 1e-12, (name, fn)
        elif fn == "snp":
            assert _close(g, w, rtol) <= rtol, (name, fn, _close(g, w, rtol))
        else:
            e = _close(g, w, rtol)
            assert e <= rtol, (name, fn, e)
