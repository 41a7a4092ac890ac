"""Wall time of the oracle (oracle/liboracle.so, the C restatement that bench.py's cpu_baseline
times on the GPU box's host) against the reference's own C build (oracle/_ref/libpeec_ref.so,
compiled in place from its sources) on the same inputs in this container, same compiler flags
and thread count.  The ratio says how far the cpu_baseline, which cannot run the reference on
the GPU box, stands from the reference's own CPU path.

Run here (needs the reference build):  make -C oracle ref && python tests/golden/time_oracle_vs_reference.py
Writes tests/golden/oracle_timing.json, which bench.py quotes beside its cpu_baseline.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parents[1]))
from openpeec_amd import abi, decks  # noqa: E402
from oracle_py import oracle_c, oracle_lib, oracle_lp, oracle_peec  # noqa: E402
from refharness import RefPeec, ref_available, ref_lib  # noqa: E402

P = C.POINTER


def best_of(fn, reps: int = 2) -> float:
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        t.append(time.perf_counter() - t0)
    return min(t)


def main() -> None:
    if not ref_available():
        raise SystemExit("reference build missing: run `make -C oracle ref` first")
    out = {"threads": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "cases": {}}
    f = 3e9
    kw = 2 * np.pi * f / 2.99792458e8
    for name, text in [("plate_32x32", decks.plate_deck(n=32)),
                       ("full_stack_16x16", decks.full_stack_deck(grid=16, ndivt=4, diel=(6, 6, 2)))]:
        r = RefPeec(text)
        r.wire_build()
        t_rl = best_of(lambda: r.lp_fill(f))
        t_rp = best_of(lambda: r.pot_fill(f))
        r.close()
        m = oracle_peec(text)
        t_ol = best_of(lambda: oracle_lp(m, kw))
        t_op = best_of(lambda: oracle_c(m, kw))
        out["cases"][name] = {"nseg": m.p.nseg, "ncell": m.p.ncell, "reference_lp_s": t_rl, "reference_pot_s": t_rp,
                              "oracle_lp_s": t_ol, "oracle_pot_s": t_op, "ratio_lp": t_ol / t_rl, "ratio_pot": t_op / t_rp}
        m.close()
        print(name, out["cases"][name], flush=True)
    ns = 2048
    rng = np.random.default_rng(1)
    a0 = (rng.standard_normal((ns, ns)) + 1j * rng.standard_normal((ns, ns))).astype(np.complex128)
    lu = {}
    for lab, fn in [("reference", ref_lib().lu_decomp), ("oracle", oracle_lib().orc_lu_decomp)]:
        def run():
            a = a0.copy()
            piv = np.zeros(ns, dtype=np.int32)
            fn(ns, a.ctypes.data_as(P(abi.d_complex_t)), piv.ctypes.data_as(P(C.c_int)))
        lu[lab] = best_of(run)
    out["lu_2048"] = {"reference_s": lu["reference"], "oracle_s": lu["oracle"], "ratio": lu["oracle"] / lu["reference"]}
    print("lu", out["lu_2048"], flush=True)
    (HERE / "oracle_timing.json").write_text(json.dumps(out, indent=1) + "\n")


if __name__ == "__main__":
    main()
