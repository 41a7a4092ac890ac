#!/usr/bin/env python3
"""Benchmark: frequency points per second of the full-wave PEEC path on one MI355X.

Headline workload (the full-stack configuration, the largest that runs on one GPU): ground plane
0.1 mm below an FR4-like brick (eps_r 4.4, 60 x 60 x 1.6 mm, 20 x 20 x 4 cells) under a copper
plate 60 x 60 mm, 35 um, on a 48 x 48 grid split into 4 volume layers; port from the plate's
corner to ground; retarded + capacitive; 150 points f_k = k 20 MHz (24 780 current cells,
4 581 capacitive cells, 29 361 MNA unknowns).  The timed region is that whole sweep, cold (the
k-factored fill records, factorisations and inversions are inside it), split into K steps:
step i is a contiguous run of the sweep's points (whole lock-step batches where K allows), so
ms_per_step = wall time / K and value = 150 / wall time.  W warm-up points run first and the
cached preconditioner and records are dropped before timing.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--no-cpu] ...
    python bench.py --only batch | learned | plate     (one side workload alone)

Form: the accelerated sweep -- k-factored Lp/P fill, charge-unknown system, A(f_ref) factored
as the preconditioner through its Schur complement on the charges (C = P^-1 and an LU of the
plain system on FP32 storage, FP64 panels and pivoting, FP32 MFMA trailing updates, factors
inverted in FP32), FP64 flexible GMRES to a 1e-10 relative residual in lock-step batches of B
frequencies, a fresh factorisation when the extra iterations have cost as much as one.  Zin of the
first, middle and last points is compared with the plain FP64 form (the reference's algorithm:
element fill, P^-1, dense FP64 LU); the run fails when the error exceeds 1e-6.  The precision
form (FP32-stored LU with FP64 refinement) runs on the same points.  roofline: the kernel with
the most GPU time in the timed window among the sweep's two heavy kernels -- the GMRES product
k_mvsym (HBM bound) and the LU trailing update (FP32 MFMA bound) -- timed with HIP events on the
library's stream around every launch; the other is reported beside it.  Side objects: the
precision workload (64 x 64 plate, 201 points), the batching workload (256 random wire
structures x 64 points), the learned-preconditioner workload (disk + quad, compressed, 101
points).  cpu_baseline times the oracle (the reference's arithmetic restated in C, OpenMP on
all host cores) on a bounded sample of one full-stack point and extrapolates.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "freq points/s (Lp/P fill+solve) at Zin rel err ≤1e-6 vs FP64; fill GFLOP/s"
PEAK_FP64_MFMA = 78.6     # TFLOP/s, MI355X dense FP64 matrix
PEAK_FP64_VECTOR = 78.6   # TFLOP/s, MI355X FP64 vector
PEAK_FP32_MFMA = 157.3    # TFLOP/s, MI355X dense FP32 matrix
PEAK_HBM_GBS = 8000.0     # GB/s, MI355X HBM3E


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="steps the full-stack sweep is split into (all timed)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed warm-up points before the cold sweep")
    ap.add_argument("--batch", type=int, default=4, help="frequencies per lock-step batch (1..8)")
    ap.add_argument("--grid", type=int, default=64, help="plate grid of the precision workload (64)")
    ap.add_argument("--no-plate", action="store_true", help="skip the precision workload (64 x 64 plate)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--no-check", action="store_true", help="skip the small-deck oracle check")
    ap.add_argument("--no-batch", action="store_true", help="skip the many-structure batching workload")
    ap.add_argument("--no-compare", action="store_true",
                    help="headline sweep only, without the plain / precision comparison points (profiling runs)")
    ap.add_argument("--batch-structs", type=int, default=256, help="structures of the batching workload")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner workload")
    ap.add_argument("--learned-points", type=int, default=101, help="frequency points of the learned workload (101)")
    ap.add_argument("--aca-tol", type=float, default=1e-12,
                    help="ACA tolerance of the learned workload (1e-12: Zin within 1e-6 of dense over the whole band; "
                         "at 1e-6 and 1e-10 the truncation alone exceeds that)")
    ap.add_argument("--only", choices=["batch", "learned", "plate"], default=None,
                    help="run one side workload alone")
    ap.add_argument("--fullstack-points", type=int, default=150, help="points of the full-stack sweep (150)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (calibrates the HBM byte counters)")
    return ap.parse_args()


def cpu_baseline(m, f: float, label: str) -> dict:
    """Oracle timing on a bounded sample of one frequency point of the workload in m (the plain
    form's stages: Lp and charge-cell fills, P LU and inverse, MNA LU, port solve), extrapolated
    to the whole point by pair count and by n^3 / n^2."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_py import oracle_lib
    from openpeec_amd import abi

    orc = oracle_lib()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    p = m.p
    kw = 2 * np.pi * f / 2.99792458e8
    t_total = 0.0
    parts = {}
    # fills: row blocks spread over the triangle, extrapolated by pair count
    for pot, n in ((0, p.nseg), (1, p.nchg)):
        rows = 24
        starts = [0, n // 3, (2 * n) // 3]
        t = 0.0
        sampled = 0
        for s0 in starts:
            t0 = time.perf_counter()
            orc.orc_fill_rows(C.byref(p), kw, s0, s0 + rows, pot)
            t += time.perf_counter() - t0
            sampled += sum(n - i for i in range(s0, s0 + rows))
        total_pairs = n * (n + 1) // 2
        parts["lp_fill_s" if pot == 0 else "p_fill_s"] = t * total_pairs / sampled
    # dense LU and P^-1 (LU + ncell unit-column solves) from a 3072 sample, large enough
    # (151 MB) to stream from DRAM like the full-size factorisation
    ns = 3072
    rng = np.random.default_rng(1)
    a = (rng.standard_normal((ns, ns)) + 1j * rng.standard_normal((ns, ns))).astype(np.complex128)
    piv = np.zeros(ns, dtype=np.int32)
    t0 = time.perf_counter()
    orc.orc_lu_decomp(ns, a.ctypes.data_as(C.POINTER(abi.d_complex_t)), piv.ctypes.data_as(C.POINTER(C.c_int)))
    t_lu = time.perf_counter() - t0
    b = np.ones(ns, dtype=np.complex128)
    nsolve = 16
    t0 = time.perf_counter()
    for _ in range(nsolve):
        orc.orc_lu_solve(ns, a.ctypes.data_as(C.POINTER(abi.d_complex_t)), piv.ctypes.data_as(C.POINTER(C.c_int)),
                         b.ctypes.data_as(C.POINTER(abi.d_complex_t)))
    t_sv = (time.perf_counter() - t0) / nsolve
    n, nc = p.nunknown, p.ncell
    parts["mna_lu_s"] = t_lu * (n / ns) ** 3
    parts["p_lu_s"] = t_lu * (nc / ns) ** 3
    parts["p_inverse_solves_s"] = t_sv * nc * (nc / ns) ** 2
    parts["port_solve_s"] = t_sv * (n / ns) ** 2
    t_total = sum(parts.values())
    out = {
        "value": 1.0 / t_total, "unit": "freq points/s", "cores": threads, "kind": "port",
        "sample": (f"oracle (reference arithmetic, bitwise-equal on the sample decks) at {f / 1e9:g} GHz: "
                   f"3 x 24 rows of the Lp and charge-cell triangles, LU and 16 solves of a {ns}^2 "
                   f"complex matrix; extrapolated by pair count and n^3 / n^2 to one point of the "
                   f"{label} (n = {n}, ncell = {nc})"),
        "seconds_per_point": t_total, "parts_s": {k: round(v, 3) for k, v in parts.items()},
    }
    # the oracle's wall time against the reference's own C build on the same inputs, measured
    # where the reference sources are (tests/golden/time_oracle_vs_reference.py)
    tj = ROOT / "tests" / "golden" / "oracle_timing.json"
    if tj.exists():
        t = json.loads(tj.read_text())
        r = [c[k] for c in t["cases"].values() for k in ("ratio_lp", "ratio_pot")] + [t["lu_2048"]["ratio"]]
        out["oracle_over_reference_time"] = {"min": min(r), "max": max(r), "threads": t["threads"],
                                             "cases": sorted(t["cases"]) + ["lu_2048"]}
    return out


def small_check(lib) -> dict:
    """Device sweep vs the oracle on the same plate at 12 x 12 (oracle-sized), 2 points."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_py import oracle_zmat
    from openpeec_amd import Peec, decks

    text = decks.plate_deck(n=12, ndiv=200)
    m = Peec.from_deck_text(text, lib)
    m.setup()
    idx = [0, 200]
    err = 0.0
    for i in idx:
        z, _ = m.sweep(i, 1)
        zo = oracle_zmat(text, i, 1)
        err = max(err, float(np.max(np.abs(z - zo) / np.abs(zo))))
    m.close()
    return {"deck": "plate 12x12, 10 MHz and 3 GHz", "zin_max_rel_err_vs_oracle": err}


def plain_points(m, idx):
    """Plain FP64 form at the given points: Zin per point and the elapsed time."""
    from openpeec_amd import PEEC_FORM_PLAIN
    zs = []
    st = None
    t0 = time.perf_counter()
    for i in idx:
        z, st = m.sweep(i, 1, PEEC_FORM_PLAIN)
        zs.append(z[0, 0, 0])
    return np.array(zs), time.perf_counter() - t0, st


def precision_points(m, idx, zp):
    """Precision form (FP32-stored LU with FP32 MFMA trailing updates, FP64 refinement) at the
    given points: rate, and the error of each Zin against the plain FP64 form."""
    from openpeec_amd import PEEC_FORM_PRECISION
    m.sweep(idx[0], 1, PEEC_FORM_PRECISION)   # first use allocates the FP32 copy
    zs, sts = [], []
    t0 = time.perf_counter()
    for i in idx:
        z, st = m.sweep(i, 1, PEEC_FORM_PRECISION)
        zs.append(z[0, 0, 0])
        sts.append(st)
    elapsed = time.perf_counter() - t0
    zs = np.array(zs)
    errs = np.abs(zs - zp) / np.abs(zp)
    gemm_ms = sum(s.lu_gemm_ms for s in sts)
    gemm_flop = sum(s.lu_gemm_flop for s in sts)
    tf = gemm_flop / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else None
    last = sts[-1]
    return {
        "points_per_s": len(idx) / elapsed,
        "zin_rel_err_vs_plain_fp64": [float(e) for e in errs],
        "max_backward_error": max(s.berr for s in sts),
        "backward_error_kind": "componentwise, max_i |b - A x|_i / (|A| |x| + |b|)_i, FP64 residual",
        "refinement_sweeps": [int(s.refine_its) for s in sts],
        "fp64_fallbacks": int(sum(s.fp64_fallbacks for s in sts)),
        "stages_ms_last_point": {"lp_fill": last.fill_ms, "charge_fill": last.pot_ms, "p_lu_and_inverse": last.cinv_ms,
                                 "mna_assembly": last.asm_ms, "mna_lu_fp32": last.lu_ms,
                                 "solve_and_refine": last.solve_ms, "total": last.total_ms},
        "lu_fp32_gemm": {"TFLOP_per_s": tf, "frac_of_fp32_mfma_peak": (tf / PEAK_FP32_MFMA) if tf else None,
                         "launches": int(sum(s.lu_gemm_launches for s in sts))},
    }


def lu_roofline(achieved: float, workload: str) -> dict:
    """Roofline of the trailing update of the accelerated form's LU (the Schur complement of A(f_ref)): FP32 storage by
    default (rocBLAS cgemm on FP32 MFMA; the hand-written k_cgemm32 with PEEC_GEMM=hip), FP64 with
    PEEC_MIXED_LU64=1 (rocBLAS zgemm / k_zgemm on FP64 MFMA).  Traffic from the workload's profile
    (profiles/<kernel>_<workload>.md) when one was taken."""
    hip = os.environ.get("PEEC_GEMM", "") == "hip"
    if os.environ.get("PEEC_MIXED_LU64", "0") == "1":
        peak, alg = PEAK_FP64_MFMA, "16 (2 M N + (M + N) K) B"
        name = "k_zgemm" if hip else "rocBLAS zgemm"
        prof = f"{'k_zgemm' if hip else 'zgemm_lu_update'}_{workload}.md"
        name += " (LU trailing update of the preconditioner, FP64 MFMA)"
    else:
        peak, alg = PEAK_FP32_MFMA, "8 (2 M N + (M + N) K) B"
        name = "k_cgemm32" if hip else "rocBLAS cgemm"
        prof = f"{'k_cgemm32' if hip else 'cgemm_lu_update'}_{workload}.md"
        name += " (LU trailing update of the preconditioner, FP32 MFMA)"
    traffic = lu_traffic(prof)
    return {"kernel": name, "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": traffic,
            "traffic_note": (f"HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a copy of known size, "
                             f"profiles/{prof}); algorithmic {alg}") if traffic else f"not profiled here; algorithmic {alg}"}


def mv_roofline(gbs: float | None, launches: int, avg_ms: float, avg_bytes: float, prof: str, what: str) -> dict:
    """Roofline of the GMRES product's dense blocks (Lp and P of every system of a batch, FP64),
    HBM bound -- k_mvsym, which reads the upper triangles only (k_mvdense, both triangles, with
    PEEC_MV_SYM=0); timed with HIP events on the library's stream around every launch of the
    timed sweep."""
    sym = os.environ.get("PEEC_MV_SYM", "1") != "0"
    if not sym:
        prof = prof.replace("k_mvsym", "k_mvdense")
    name = (f"k_mvsym (GMRES product A x: upper triangles of the symmetric Lp and P {what}, FP64)" if sym
            else f"k_mvdense (GMRES product A x, dense Lp and P blocks {what}, FP64)")
    traffic = lu_traffic(prof)
    return {"kernel": name, "bound": "hbm",
            "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": (gbs / PEAK_HBM_GBS) if gbs else None,
            "traffic": traffic, "launches": launches, "avg_launch_ms": avg_ms, "avg_launch_bytes": avg_bytes,
            "traffic_note": (f"HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a copy of known size, "
                             f"profiles/{prof}); " if traffic else "not profiled here; ") +
                            "algorithmic 16 B per entry read of Lp and P per system (upper triangle with whole "
                            "diagonal tiles for k_mvsym), x once, partial sums written"}


def lu_traffic(name: str) -> float | None:
    """HBM bytes per LU trailing-update launch, measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    (calibrated on a copy of known size) and recorded in the named profile."""
    f = ROOT / "profiles" / name
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def log(msg: str) -> None:
    """Progress to stderr (the JSON line alone goes to stdout)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def bench_batch(lib, nstruct: int = 256, seed0: int = 0x5EED0003, plain_sample: int = 2) -> dict:
    """The batching workload: nstruct random wire structures (20-400 segments), 64 frequencies
    each, swept together by the batched kernels; the plain FP64 form of each structure on
    plain_sample of its frequencies for the rate and Zin comparison."""
    from openpeec_amd import PEEC_FORM_PLAIN, Batch, Peec, decks
    t0 = time.perf_counter()
    ms = []
    for s in range(nstruct):
        m = Peec.from_deck_text(decks.random_wire_deck(seed0 + s), lib)
        m.setup()
        ms.append(m)
    t_setup = time.perf_counter() - t0
    nf = ms[0].p.nfreq
    t0 = time.perf_counter()
    b = Batch(ms, 0, nf)
    t_create = time.perf_counter() - t0
    log(f"batching: {nstruct} structures set up")
    st_t = b.tune()
    trials = b.trials()
    b.sweep()                      # warm: graphs captured, buffers sized
    t0 = time.perf_counter()
    zs, st = b.sweep()
    elapsed = time.perf_counter() - t0
    log(f"batching: sweep {elapsed:.2f} s")
    jobs = st.jobs
    # the precision form of the batched factorisation on the same jobs (tuned for it first)
    from openpeec_amd import PEEC_FORM_PRECISION
    b.set_form(PEEC_FORM_PRECISION)
    b.tune()
    b.sweep()
    t0 = time.perf_counter()
    zq, stq = b.sweep()
    t_prec = time.perf_counter() - t0
    err_q = max(float(np.max(np.abs(a - c) / np.abs(c))) for a, c in zip(zq, zs))
    log(f"batching: precision form {t_prec:.2f} s")
    # plain FP64 form: every structure at plain_sample frequencies spread over its sweep
    idx = [int(round(x)) for x in np.linspace(0, nf - 1, plain_sample)]
    t0 = time.perf_counter()
    err = 0.0
    for m, z in zip(ms, zs):
        for i in idx:
            zp, _ = m.sweep(i, 1, PEEC_FORM_PLAIN)
            err = max(err, float(np.max(np.abs(z[i] - zp[0]) / np.abs(zp[0]))))
    t_plain = time.perf_counter() - t0
    plain_rate = nstruct * len(idx) / t_plain
    nseg = [m.p.nseg for m in ms]
    order = [m.p.nunknown + (m.p.ncell if m.p.capacitance else 0) for m in ms]
    b.close()
    for m in ms:
        m.close()
    lu_tf = st.lu_flop / (st.lu_ms * 1e-3) / 1e12 if st.lu_ms > 0 else None
    return {
        "workload": f"{nstruct} random wire structures (seed {seed0:#x}+s), 20-400 segments of 5-20 mm in a 0.3 m cube, "
                    f"radius 0.5-2 mm, 1 mm feed gap, retarded + capacitive, {nf} frequencies 50 MHz-2 GHz each",
        "points": jobs, "points_per_s": jobs / elapsed, "ms": 1e3 * elapsed,
        "plain_fp64_points_per_s": plain_rate, "speedup_vs_plain_fp64": (jobs / elapsed) / plain_rate,
        "zin_max_rel_err_vs_plain_fp64": err, "zin_checked_points": nstruct * len(idx),
        "segments": {"min": min(nseg), "max": max(nseg), "mean": float(np.mean(nseg))},
        "system_order": {"min": min(order), "max": max(order)},
        "stages_ms": {"fill_graphs": st.fill_ms, "stamp_host": st.stamp_ms, "lu_solve_graphs": st.lu_ms,
                      "total": st.total_ms},
        "lu": {"TFLOP_per_s": lu_tf, "frac_of_fp64_mfma_peak": (lu_tf / PEAK_FP64_MFMA) if lu_tf else None,
               "flop": st.lu_flop},
        "chunks": st.chunks, "direct_pair_evaluations": st.direct_pairs,
        "precision_form": {"points_per_s": jobs / t_prec, "ms": 1e3 * t_prec, "lu_solve_graphs_ms": stq.lu_ms,
                           "zin_max_rel_err_vs_fp64_batched": err_q, "max_backward_error_accepted": stq.berr_max,
                           "refinement_sweeps_per_port": stq.refine_sweeps / max(1, jobs),
                           "fp64_redos": stq.fp64_fallbacks, "tuned_panel_width": stq.nb,
                           "tuned_tile_cols": stq.tile_cols, "tuned_wg_per_cu": stq.wg_per_cu},
        "tuned": {"panel_width": st.nb, "tile_cols": st.tile_cols, "wg_per_cu": st.wg_per_cu,
                  "order": "largest first" if st.order == 0 else "interleaved",
                  "trials": [{"nb": t[0], "tile_cols": t[1], "wpc": t[2], "order": t[3], "lu_ms": t[4]} for t in trials],
                  "tune_ms": st_t.tune_ms},
        "setup_s": {"decks_and_geometry": t_setup, "records_and_tables": t_create},
    }


LPC_TRAIN_GEOMS = [dict(nr=8, nsec=16, qx=8, qy=4), dict(nr=10, nsec=20, qx=10, qy=6, gap=0.004),
                   dict(nr=12, nsec=24, qx=12, qy=6, gap=0.007), dict(nr=24, nsec=32, qx=20, qy=10, gap=0.006)]
LPC_TRAIN_PAIRS = [(fr, fr * s) for fr in (4e8, 1e9, 2e9) for s in (1.0, 1.3, 1.8, 2.5)]


def bench_learned(lib, points: int = 101, aca_tol: float = 1e-12) -> dict:
    """The learned-preconditioner workload: disk 48 x 96 under a 32 x 16 quad, compressed at ACA
    tolerance aca_tol, swept over `points` of 100 MHz - 4 GHz three ways: the reference's rule
    (precond.c's preconditioner rebuilt at every frequency, reference GMRES termination), the
    same preconditioner reused across frequencies without the network (theta = 0), and with the
    network trained on the device on small geometries (training time reported separately).  Zin
    of every point is compared with the same geometry uncompressed (accelerated dense form, itself
    checked against the plain dense FP64 form on 5 points across the band).  The truncation alone
    leaves the reference's own compressed solve up to 1.7e-4 from dense at ACA 1e-6 (reduced
    deck), and 4.1e-6 from its ACA-1e-12 answer at 1e-10 (this deck at 4 GHz, |Zin| = 6.6 ohm)."""
    import ctypes as C
    from openpeec_amd import PEEC_FORM_LEARNED, PEEC_FORM_PLAIN, Peec, decks
    PD = C.POINTER(C.c_double)
    t0 = time.perf_counter()
    Xs, Cs = [], []
    for g in LPC_TRAIN_GEOMS:
        m = Peec.from_deck_text(decks.disk_quad_deck(**g, f0=1e8, f1=4e9, ndiv=4), lib)
        m.setup()
        fr = np.array([a for a, _ in LPC_TRAIN_PAIRS])
        fv = np.array([b for _, b in LPC_TRAIN_PAIRS])
        X = np.zeros(len(fr) * 8)
        coef = np.zeros(len(fr) * 4)
        lib.peec_lpc_samples(C.byref(m.p), fr.ctypes.data_as(PD), fv.ctypes.data_as(PD), len(fr),
                             X.ctypes.data_as(PD), coef.ctypes.data_as(PD), None)
        Xs.append(X)
        Cs.append(coef)
        m.close()
        log(f"learned: samples of {g} done")
    t_samples = time.perf_counter() - t0
    X = np.ascontiguousarray(np.concatenate(Xs))
    coef = np.ascontiguousarray(np.concatenate(Cs))
    S = len(coef) // 4
    W = np.zeros(lib.peec_lpc_nparams())
    hist = np.zeros(1501)
    t0 = time.perf_counter()
    lib.peec_lpc_fit(X.ctypes.data_as(PD), coef.ctypes.data_as(PD), S, 1500, 3e-3, 7, W.ctypes.data_as(PD),
                     hist.ctypes.data_as(PD))
    t_fit = time.perf_counter() - t0
    text = decks.disk_quad_deck(ndiv=points - 1, ctol=aca_tol)
    m = Peec.from_deck_text(text, lib)
    m.setup()
    out = {"workload": f"disk r = 50 mm, 48 rings x 96 sectors, under a 80 x 40 mm quad (32 x 16) 5 mm above, "
                       f"compressed (ACA {aca_tol:g}), retarded + capacitive, {points} points 100 MHz-4 GHz",
           "unknowns": m.p.nunknown, "branches": m.p.nseg, "charge_cells": m.p.ncell,
           "training": {"samples": S, "geometries": len(LPC_TRAIN_GEOMS), "sample_s": t_samples, "fit_s": t_fit,
                        "loss_first": float(hist[0]), "loss_last": float(hist[-1])}}
    runs = {}
    # untimed: the band's k-factored records and the device buffers are set up once per
    # geometry, by whichever sweep comes first
    m.sweep(0, points, PEEC_FORM_PLAIN)
    lib.peec_lpc_set_model(W.ctypes.data_as(PD))
    t0 = time.perf_counter()
    z_l, st = m.sweep(0, points, PEEC_FORM_LEARNED)
    runs["learned"] = (time.perf_counter() - t0, st)
    log(f"learned: sweep with the network {runs['learned'][0]:.1f} s, {st.gmres_its} iterations")
    lib.peec_lpc_set_model(None)
    t0 = time.perf_counter()
    z_0, st = m.sweep(0, points, PEEC_FORM_LEARNED)
    runs["reuse_without_network"] = (time.perf_counter() - t0, st)
    log(f"learned: sweep without the network {runs['reuse_without_network'][0]:.1f} s, {st.gmres_its} iterations")
    t0 = time.perf_counter()
    z_r, st = m.sweep(0, points, PEEC_FORM_PLAIN)
    runs["reference_rule"] = (time.perf_counter() - t0, st)
    log(f"learned: reference-rule sweep {runs['reference_rule'][0]:.1f} s, {st.gmres_its} iterations")
    for k, (t, st) in runs.items():
        out[k] = {"s": t, "points_per_s": points / t, "gmres_its": int(st.gmres_its),
                  "its_per_point": st.gmres_its / points, "preconditioner_builds": int(st.pc_builds) if k != "reference_rule" else points,
                  "fallbacks": int(st.lpc_fallbacks), "stages_ms": {"compressed_build": st.fill_ms, "preconditioner": st.lu_ms,
                                                                   "gmres": st.solve_ms}}
    out["learned"]["mean_abs_theta"] = runs["learned"][1].theta_abs_sum / points
    out["zin_max_rel_err"] = {"learned_vs_reference_rule": float(np.max(np.abs(z_l - z_r) / np.abs(z_r))),
                              "without_vs_reference_rule": float(np.max(np.abs(z_0 - z_r) / np.abs(z_r)))}
    out["gain_vs_without_network"] = {"iterations": runs["reuse_without_network"][1].gmres_its / max(1, runs["learned"][1].gmres_its),
                                      "time": runs["reuse_without_network"][0] / runs["learned"][0]}
    out["speedup_vs_reference_rule"] = runs["reference_rule"][0] / runs["learned"][0]
    m.close()
    # the same geometry uncompressed: the accelerated form (all points, after an untimed run)
    # and the plain dense FP64 form (first 5 points), against each other and against the
    # compressed result (which carries the ACA truncation)
    import re
    from openpeec_amd import PEEC_FORM_MIXED
    dm = Peec.from_deck_text(re.sub(r"compression = 1 \S+", "compression = 0", text), lib)
    dm.setup()
    pidx = sorted({int(round(x)) for x in np.linspace(0, points - 1, min(points, 5))})
    dm.sweep(0, points, PEEC_FORM_MIXED)
    t0 = time.perf_counter()
    z_m, _ = dm.sweep(0, points, PEEC_FORM_MIXED)
    t_m = time.perf_counter() - t0
    z_p, t_p, _ = plain_points(dm, pidx)
    rel = lambda a, b: float(np.max(np.abs(a - b) / np.abs(b)))
    per_point = np.abs(z_l[:, 0, 0] - z_m[:, 0, 0]) / np.abs(z_m[:, 0, 0])
    out["uncompressed_same_geometry"] = {
        "unknowns": dm.p.nunknown, "accelerated_points_per_s": points / t_m, "plain_fp64_points_per_s": len(pidx) / t_p,
        "plain_points_hz": [dm.freq(i) for i in pidx], "accelerated_vs_plain_fp64": (points / t_m) / (len(pidx) / t_p),
        "accelerated_vs_compressed_reference_rule": (points / t_m) / out["reference_rule"]["points_per_s"],
        "zin_max_rel_err": {"accelerated_vs_plain_fp64": rel(z_m[pidx, 0, 0], z_p),
                            "learned_vs_plain_fp64": rel(z_l[pidx, 0, 0], z_p),
                            "compressed_reference_rule_vs_accelerated_all_points": rel(z_r, z_m),
                            "learned_vs_accelerated_all_points": float(per_point.max())}}
    out["zin_check"] = {"learned_vs_dense_per_point_max": float(per_point.max()),
                        "worst_point_hz": dm.freq(int(np.argmax(per_point))), "bar": 1e-6,
                        "ok": bool(per_point.max() <= 1e-6 and rel(z_m[pidx, 0, 0], z_p) <= 1e-6)}
    log(f"learned: same geometry uncompressed, accelerated {points / t_m:.1f} points/s, plain FP64 {len(pidx) / t_p:.2f} points/s")
    dm.close()
    return out


def sweep_chunks(npts: int, k: int, batch: int) -> list[tuple[int, int]]:
    """The sweep's points 0..npts-1 as k contiguous runs (i0, count), whole lock-step batches of
    `batch` points each where k allows (the last run takes the remainder)."""
    nbat = (npts + batch - 1) // batch
    if k <= nbat:
        sizes = [(nbat // k + (1 if i < nbat % k else 0)) * batch for i in range(k)]
    else:
        k = min(k, npts)
        sizes = [npts // k + (1 if i < npts % k else 0) for i in range(k)]
    out, i0 = [], 0
    for sz in sizes:
        c = min(sz, npts - i0)
        if c > 0:
            out.append((i0, c))
        i0 += c
    return out


def headline_roofline(st, workload: str = "fullstack_sweep", what: str = "of 4 systems of the full stack") -> tuple[dict, dict]:
    """The timed window's two heavy kernels with their HIP-event times: the GMRES product
    k_mvsym (HBM) and the LU trailing update (rocBLAS cgemm on FP32 MFMA, k_cgemm32 with
    PEEC_GEMM=hip); returns (dominant, other)."""
    mv_gbs = (st.mv_bytes / (st.mv_ms * 1e-3)) / 1e9 if st.mv_ms > 0 else None
    mv = mv_roofline(mv_gbs, st.mv_launches, st.mv_ms / max(1, st.mv_launches), st.mv_bytes / max(1, st.mv_launches),
                     f"k_mvsym_{workload}.md", what)
    mv["total_ms_in_timed_window"] = st.mv_ms
    gemm_ms = st.lu_gemm_ms / max(1, st.lu_gemm_launches)
    gemm_flop = st.lu_gemm_flop / max(1, st.lu_gemm_launches)
    achieved = (gemm_flop / (gemm_ms * 1e-3)) / 1e12 if gemm_ms > 0 else 0.0
    lu = lu_roofline(achieved, workload) | {"launches": st.lu_gemm_launches, "avg_launch_ms": gemm_ms,
                                                       "avg_launch_flop": gemm_flop,
                                                       "total_ms_in_timed_window": st.lu_gemm_ms}
    return (mv, lu) if st.mv_ms >= st.lu_gemm_ms else (lu, mv)


def device_sync() -> None:
    """Waits for every stream of the device, the library's own included, so the timed region
    ends after all the work it started."""
    C.CDLL("libamdhip64.so").hipDeviceSynchronize()


def bench_headline(lib, args) -> dict:
    """The full-stack workload swept cold in K steps (see the module docstring)."""
    from openpeec_amd import PEEC_FORM_MIXED, Peec, decks
    t0 = time.perf_counter()
    m = Peec.from_deck_text(decks.full_stack_deck(), lib)
    m.setup()
    log(f"full stack: {m.p.nseg} current cells, {m.p.ncell} capacitive cells, {m.p.nunknown} unknowns "
        f"({time.perf_counter() - t0:.1f} s setup)")
    nf = m.p.nfreq
    npts = max(2, min(args.fullstack_points, nf))
    chunks = sweep_chunks(npts, max(1, args.steps), args.batch)
    if args.warmup > 0:
        m.sweep(0, min(args.warmup, npts), PEEC_FORM_MIXED, args.batch)
    lib.peec_sweep_reset()
    z = np.zeros((npts, m.p.nport, m.p.nport), dtype=np.complex128)
    tot = None
    step_ms = []
    t0 = time.perf_counter()
    for (i0, c) in chunks:
        ts = time.perf_counter()
        zc, st = m.sweep(i0, c, PEEC_FORM_MIXED, args.batch)
        step_ms.append(1e3 * (time.perf_counter() - ts))
        z[i0:i0 + c] = zc
        tot = st if tot is None else add_stats(tot, st)
    device_sync()
    elapsed = time.perf_counter() - t0
    log(f"full stack: cold sweep of {npts} points in {len(chunks)} steps, {elapsed:.1f} s ({tot.n_lu} factorisations, "
        f"{tot.gmres_its} GMRES iterations)")
    if args.no_compare:
        roof, other = headline_roofline(tot)
        return {"value": npts / elapsed, "unit": "freq points/s", "steps": len(chunks),
                "ms_per_step": 1e3 * elapsed / len(chunks), "step_ms": step_ms, "n_lu": tot.n_lu,
                "gmres_its_per_point": tot.gmres_its / npts, "zin_check": {"skipped": True, "ok": True},
                "stages_ms": {"fill_incl_kf_setup": tot.fill_ms, "kf_setup": tot.kf_setup_ms, "lu_and_inverse": tot.lu_ms,
                              "gmres": tot.solve_ms, "total": tot.total_ms},
                "roofline": roof, "second_kernel": other}, m
    idx = [0, npts // 2, npts - 1]
    zp, t_plain, stp = plain_points(m, idx)
    log(f"full stack: plain FP64 form {t_plain / len(idx):.1f} s per point")
    prec = precision_points(m, idx, zp)
    zm = np.array([z[i, 0, 0] for i in idx])
    errs = np.abs(zm - zp) / np.abs(zp)
    err = float(np.max(errs))
    fill_flop_top = lib.peec_fill_flops(C.byref(m.p), m.freq(idx[-1]))
    plain_fill_ms = stp.fill_ms + stp.pot_ms
    roof, other = headline_roofline(tot)
    out = {
        "value": npts / elapsed, "unit": "freq points/s", "n_gpus": 1, "steps": len(chunks), "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / len(chunks), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (fill, GMRES, residuals); preconditioner factors stored f32",
        "data": "synthetic deck: full-stack geometry generated in-process (no dataset)",
        "config": {
            "workload": "full stack: ground plane 0.1 mm under an FR4 brick (eps_r 4.4, 60 x 60 x 1.6 mm, 20 x 20 x 4 "
                        "cells); copper plate 60 x 60 mm, 35 um, 48 x 48 grid in 4 volume layers on the brick; port "
                        f"plate corner to ground; retarded + capacitive; {npts} points f_k = k 20 MHz; one step = one "
                        f"contiguous run of {npts / len(chunks):.1f} points on average (the whole sweep is timed)",
            "unknowns": m.p.nunknown, "current_cells": m.p.nseg, "capacitive_cells": m.p.ncell, "form": "accelerated",
            "batch": args.batch, "points_hz": [m.freq(0), m.freq(npts - 1)],
            "step_points": [c for _, c in chunks]},
        "zin_check": {"points_hz": [m.freq(i) for i in idx], "rel_err_vs_plain_fp64": [float(e) for e in errs],
                      "max_rel_err_vs_plain_fp64": err, "bar": 1e-6, "ok": err <= 1e-6},
        "plain_fp64_points_per_s": len(idx) / t_plain,
        "speedup_vs_plain_fp64": (npts / elapsed) / (len(idx) / t_plain),
        "precision_form": prec,
        "stages_ms": {"fill_incl_kf_setup": tot.fill_ms, "kf_setup": tot.kf_setup_ms, "lu_and_inverse": tot.lu_ms,
                      "gmres": tot.solve_ms, "total": tot.total_ms},
        "step_ms": step_ms,
        "n_lu": tot.n_lu, "gmres_its_per_point": tot.gmres_its / npts, "kf_direct_pairs": tot.kf_direct_pairs,
        "fill_nominal_gflop_per_point": lib.peec_fill_flops(C.byref(m.p), m.freq(npts - 1)) / 1e9,
        "plain_fill": {"gflops": fill_flop_top / (plain_fill_ms * 1e-3) / 1e9,
                       "frac_of_fp64_vector_peak": fill_flop_top / (plain_fill_ms * 1e-3) / 1e12 / PEAK_FP64_VECTOR,
                       "note": "element-by-element FP64 fill (Lp and charge cells) of the plain form at 3 GHz"},
        "plain_stages_ms_last_point": {"lp_fill": stp.fill_ms, "charge_fill": stp.pot_ms, "p_lu_and_inverse": stp.cinv_ms,
                                       "mna_assembly": stp.asm_ms, "mna_lu": stp.lu_ms, "total": stp.total_ms},
        "roofline": roof,
        "second_kernel": other,
    }
    return out, m


def add_stats(a, b):
    """Sum of two peec_stats_t of consecutive sweep calls (times, counts, flops, bytes)."""
    from openpeec_amd import abi
    r = abi.peec_stats_t()
    for name, _ in abi.peec_stats_t._fields_:
        va, vb = getattr(a, name), getattr(b, name)
        if name in ("form", "ising", "kf_direct_pairs"):
            setattr(r, name, vb)
        elif name == "berr":
            setattr(r, name, max(va, vb))
        else:
            setattr(r, name, va + vb)
    return r


def bench_plate(lib, grid: int = 64, batch: int = 4) -> dict:
    """The precision workload (64 x 64 plate, 201 points 10 MHz - 3 GHz): the accelerated form
    over the whole sweep, cold; the plain FP64 and precision forms at its first, middle and last
    points, with the accelerated form's Zin error against the plain one."""
    from openpeec_amd import PEEC_FORM_MIXED, Peec, decks
    m = Peec.from_deck_text(decks.plate_deck(n=grid), lib)
    m.setup()
    nf = m.p.nfreq
    m.sweep(0, 1, PEEC_FORM_MIXED, batch)
    lib.peec_sweep_reset()
    t0 = time.perf_counter()
    z, st = m.sweep(0, nf, PEEC_FORM_MIXED, batch)
    elapsed = time.perf_counter() - t0
    idx = [0, nf // 2, nf - 1]
    zp, t_plain, stp = plain_points(m, idx)
    prec = precision_points(m, idx, zp)
    zm = np.array([z[i, 0, 0] for i in idx])
    err = float(np.max(np.abs(zm - zp) / np.abs(zp)))
    fill_flop_top = lib.peec_fill_flops(C.byref(m.p), m.freq(idx[-1]))
    plain_fill_ms = stp.fill_ms + stp.pot_ms
    roof, other = headline_roofline(st, "plate64_mixed_sweep", "of 4 systems of the 64 x 64 plate")
    out = {
        "workload": f"copper plate 100x100 mm, 35 um, {grid}x{grid} grid, retarded + capacitive, corner-to-corner port, "
                    f"10 MHz-3 GHz sweep of {nf} points, accelerated form, cold",
        "unknowns": m.p.nunknown, "points": nf, "s": elapsed, "points_per_s": nf / elapsed,
        "zin_max_rel_err_vs_plain_fp64": err, "plain_fp64_points_per_s": len(idx) / t_plain,
        "speedup_vs_plain_fp64": (nf / elapsed) / (len(idx) / t_plain), "precision_form": prec,
        "stages_ms": {"fill_incl_kf_setup": st.fill_ms, "kf_setup": st.kf_setup_ms, "lu_and_inverse": st.lu_ms,
                      "gmres": st.solve_ms, "total": st.total_ms},
        "n_lu": st.n_lu, "gmres_its_per_point": st.gmres_its / nf,
        "plain_fill": {"gflops": fill_flop_top / (plain_fill_ms * 1e-3) / 1e9,
                       "frac_of_fp64_vector_peak": fill_flop_top / (plain_fill_ms * 1e-3) / 1e12 / PEAK_FP64_VECTOR},
        "gmres_product": roof if roof["unit"] == "GB/s" else other,
        "lu_update": roof if roof["unit"] == "TFLOP/s" else other,
    }

    m.close()
    return out


def copy_probe(mib: int) -> None:
    """A copy of known size through one elementwise kernel (y = 1.0 * x): its FETCH_SIZE /
    WRITE_SIZE per dispatch against mib MiB read and written scales the counters of the path's
    kernels to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    x = torch.empty(n, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        torch.mul(x, 1.0, out=y)   # one elementwise kernel: reads and writes mib MiB
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_mib": mib, "bytes_read": 4 * n, "bytes_written": 4 * n, "dispatches": 3}))


def main():
    args = parse()
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this library runs on one GPU; sharding the frequencies over several "
                                   "GPUs is not part of it", "n_gpus": args.gpus}))
        sys.exit(2)
    from openpeec_amd import load_library

    lib = load_library(build_if_missing=True)
    if not lib.peec_device_available():
        raise SystemExit("no HIP device: the library has no CPU path")
    if args.only == "batch":
        print(json.dumps({"batching": bench_batch(lib, args.batch_structs)}))
        return
    if args.only == "learned":
        print(json.dumps({"learned_preconditioner": bench_learned(lib, args.learned_points, args.aca_tol)}))
        return
    if args.only == "plate":
        print(json.dumps({"precision_workload": bench_plate(lib, args.grid, args.batch)}))
        return
    result = {"metric": METRIC}
    head, m = bench_headline(lib, args)
    result.update(head)
    if not args.no_check:
        result["small_deck_check"] = small_check(lib)
    if not args.no_plate:
        log("precision workload (64 x 64 plate)")
        result["precision_workload"] = bench_plate(lib, args.grid, args.batch)
    if not args.no_batch:
        result["batching"] = bench_batch(lib, args.batch_structs)
    if not args.no_learned:
        result["learned_preconditioner"] = bench_learned(lib, args.learned_points, args.aca_tol)
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        from oracle_py import oracle_peec
        from openpeec_amd import decks
        m_cpu = oracle_peec(decks.full_stack_deck())
        result["cpu_baseline"] = cpu_baseline(m_cpu, m.freq(m.p.nfreq - 1), "full stack (29 361 unknowns)")
        m_cpu.close()
    m.close()
    print(json.dumps(result))
    if not result["zin_check"]["ok"]:
        log(f"Zin check failed: {result['zin_check']['max_rel_err_vs_plain_fp64']:.3e} > 1e-6")
        sys.exit(3)


if __name__ == "__main__":
    main()
