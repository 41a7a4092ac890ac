#!/usr/bin/env python3
"""Benchmark: frequency points per second of the full-wave PEEC path on one MI355X.

Workload (the precision configuration): copper plate 100 x 100 mm, 35 um, 64 x 64 grid
(8320 ribbon branches, 4225 charge cells, 12545 MNA unknowns), retardation and capacitance
on, corner-to-corner port, sweep 10 MHz - 3 GHz in 201 points.  A step is one frequency point
of that sweep, device resident from the cell arrays to Z(f).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--form mixed|plain] [--batch B] [--no-cpu]
    python bench.py --only batch | learned | fullstack     (one side workload alone)

Forms.  mixed (default): the sweep's points 0..K-1 in one accelerated sweep -- k-factored Lp/P
fill, charge-unknown system, an LU of A(f_ref) on FP32 storage (FP64 panels and pivoting, FP32
MFMA trailing updates) whose factors are inverted in FP32, FP64 flexible GMRES to a 1e-10
relative residual in lock-step batches of B frequencies, a fresh factorisation whenever the
extra iterations have cost about as much as one; the LUs, the inversions and the fill
preparation run inside the timed region (cold start).  plain: every
point independently (fill, P LU and inverse, assembly, FP64 LU, solve), K points spread over
the sweep.  Either way value = K / (wall time of the K points).

Zin of the mixed form is compared with the plain FP64 form on three of the timed points (the
metric is quoted at Zin rel err <= 1e-6); the plain form's own rate on those points is reported
beside it.  The roofline block is the sweep's dominant kernel, k_mvsym (the dense blocks of the
GMRES product A x from their upper triangles, HBM bound), timed with HIP events on the library's stream around every launch;
the LU trailing update (rocBLAS cgemm on FP32 MFMA) is reported the same way in lu_update.  The precision form (FP32-stored LU, FP32 MFMA trailing updates, FP64 refinement)
runs on the same three points and reports its rate and its Zin error against the plain form per
point.  cpu_baseline times the oracle (the reference's arithmetic restated in C, OpenMP on
all host cores) on a bounded sample and extrapolates to one frequency point.
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
    ap.add_argument("--steps", type=int, default=None, help="points timed (default: 201 mixed, 3 plain)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--form", choices=["plain", "mixed"], default="mixed")
    ap.add_argument("--batch", type=int, default=4, help="mixed form: frequencies per lock-step batch (1..8)")
    ap.add_argument("--grid", type=int, default=64, help="plate grid (64 = the precision workload)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--no-check", action="store_true", help="skip the small-deck oracle check")
    ap.add_argument("--no-batch", action="store_true", help="skip the many-structure batching workload")
    ap.add_argument("--no-compare", action="store_true",
                    help="accelerated form only, without the plain / precision comparison points (profiling runs)")
    ap.add_argument("--batch-structs", type=int, default=256, help="structures of the batching workload")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner workload")
    ap.add_argument("--learned-points", type=int, default=21,
                    help="frequency points of the learned workload (21: every fifth of its 101)")
    ap.add_argument("--only", choices=["batch", "learned", "fullstack"], default=None,
                    help="run one side workload alone (fullstack: the full-stack configuration, not in the default run)")
    ap.add_argument("--fullstack-points", type=int, default=150, help="points of the full-stack sweep")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB three times (calibrates the HBM byte counters)")
    return ap.parse_args()


def cpu_baseline(m, budget_s: float = 20.0) -> dict:
    """Oracle timing on a bounded sample of one C2 frequency point, extrapolated."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_py import oracle_lib
    from openpeec_amd import abi

    orc = oracle_lib()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    p = m.p
    f = 3e9
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
    # dense LU (12545) and P^-1 (LU + ncell unit-column solves at 4225) from a 3072 sample,
    # large enough (151 MB) to stream from DRAM like the full-size factorisation
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
    return {
        "value": 1.0 / t_total, "unit": "freq points/s", "cores": threads, "kind": "port",
        "sample": (f"oracle (reference arithmetic, bitwise-equal on the sample decks) at 3 GHz: "
                   f"3 x 24 rows of the Lp and charge-cell triangles, LU and 16 solves of a {ns}^2 "
                   f"complex matrix; extrapolated by pair count and n^3 / n^2 to one point of the "
                   f"64 x 64 plate (n = {n}, ncell = {nc})"),
        "seconds_per_point": t_total, "parts_s": {k: round(v, 3) for k, v in parts.items()},
    }


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


def lu_roofline(achieved: float) -> dict:
    """Roofline of the accelerated sweep's dominant kernel, the trailing update of the LU of
    A(f_ref): FP32 storage by default (rocBLAS cgemm on FP32 MFMA; the hand-written k_cgemm32 with
    PEEC_GEMM=hip), FP64 with PEEC_MIXED_LU64=1 (rocBLAS zgemm / k_zgemm on FP64 MFMA)."""
    hip = os.environ.get("PEEC_GEMM", "") == "hip"
    if os.environ.get("PEEC_MIXED_LU64", "0") == "1":
        peak, alg = PEAK_FP64_MFMA, "16 (2 M N + (M + N) K) B"
        name = "k_zgemm" if hip else "rocBLAS zgemm"
        prof = "k_zgemm_plate64_mixed_sweep.md" if hip else None
        name += " (LU trailing update of A(f_ref), FP64 MFMA)"
    else:
        peak, alg = PEAK_FP32_MFMA, "8 (2 M N + (M + N) K) B"
        name = "k_cgemm32" if hip else "rocBLAS cgemm"
        prof = "k_cgemm32_plate64_mixed_sweep.md" if hip else "cgemm_lu_update_plate64_mixed_sweep.md"
        name += " (LU trailing update of A(f_ref), FP32 MFMA)"
    return {"kernel": name, "bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
            "frac": achieved / peak, "traffic": lu_traffic(prof) if prof else None,
            "traffic_note": (f"HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a copy of known size, "
                             f"profiles/{prof}); algorithmic {alg}") if prof else f"not profiled; algorithmic {alg}"}


def mv_roofline(gbs: float | None, launches: int, avg_ms: float, avg_bytes: float) -> dict:
    """Roofline of the accelerated sweep's dominant kernel (most GPU time per sweep): the dense
    blocks of the GMRES product A x (Lp and P of every system of a batch, FP64), HBM bound --
    k_mvsym, which reads the upper triangles only (k_mvdense, both triangles, with
    PEEC_MV_SYM=0); timed with HIP events on the library's stream around every launch of the
    timed sweep."""
    sym = os.environ.get("PEEC_MV_SYM", "1") != "0"
    prof = "k_mvsym_plate64_mixed_sweep.md" if sym else "k_mvdense_plate64_mixed_sweep.md"
    name = ("k_mvsym (GMRES product A x: upper triangles of the symmetric Lp and P of 4 systems, FP64)" if sym
            else "k_mvdense (GMRES product A x, dense Lp and P blocks of 4 systems, FP64)")
    return {"kernel": name, "bound": "hbm",
            "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": (gbs / PEAK_HBM_GBS) if gbs else None,
            "traffic": lu_traffic(prof), "launches": launches, "avg_launch_ms": avg_ms, "avg_launch_bytes": avg_bytes,
            "traffic_note": f"HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE scaled by a copy of known size, "
                            f"profiles/{prof}); algorithmic 16 B per entry read of Lp and P per system (upper "
                            f"triangle with whole diagonal tiles for k_mvsym), x once, partial sums written"}


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


def bench_learned(lib, points: int = 101) -> dict:
    """The learned-preconditioner workload: disk 48 x 32 under a 32 x 16 quad, compressed, swept
    over `points` of 100 MHz - 4 GHz three ways: the reference's rule (precond.c's preconditioner
    rebuilt at every frequency, reference GMRES termination), the same preconditioner reused
    across frequencies without the network (theta = 0), and with the network trained on the
    device on small geometries (training time reported separately)."""
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
    text = decks.disk_quad_deck(ndiv=points - 1)
    m = Peec.from_deck_text(text, lib)
    m.setup()
    out = {"workload": f"disk r = 50 mm, 48 rings x 32 sectors, under a 80 x 40 mm quad (32 x 16) 5 mm above, "
                       f"compressed (ACA 1e-6), retarded + capacitive, {points} points 100 MHz-4 GHz",
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
    kp = min(points, 5)
    dm.sweep(0, points, PEEC_FORM_MIXED)
    t0 = time.perf_counter()
    z_m, _ = dm.sweep(0, points, PEEC_FORM_MIXED)
    t_m = time.perf_counter() - t0
    t0 = time.perf_counter()
    z_p, _ = dm.sweep(0, kp, PEEC_FORM_PLAIN)
    t_p = time.perf_counter() - t0
    rel = lambda a, b: float(np.max(np.abs(a - b) / np.abs(b)))
    out["uncompressed_same_geometry"] = {
        "unknowns": dm.p.nunknown, "accelerated_points_per_s": points / t_m, "plain_fp64_points_per_s": kp / t_p,
        "plain_points": kp, "accelerated_vs_plain_fp64": (points / t_m) / (kp / t_p),
        "accelerated_vs_compressed_reference_rule": (points / t_m) / out["reference_rule"]["points_per_s"],
        "zin_max_rel_err": {"accelerated_vs_plain_fp64": rel(z_m[:kp], z_p),
                            "compressed_vs_plain_fp64": rel(z_r[:kp], z_p),
                            "compressed_vs_accelerated_all_points": rel(z_r, z_m)}}
    log(f"learned: same geometry uncompressed, accelerated {points / t_m:.1f} points/s, plain FP64 {kp / t_p:.2f} points/s")
    dm.close()
    return out


def bench_fullstack(lib, npoints: int = 150) -> dict:
    """The full-stack configuration (ground plane, 4-layer volume plate over a dielectric brick,
    n = 29 361 unknowns): the accelerated form over npoints points of its 20 MHz-step sweep (cold,
    one warm-up point first), the plain FP64 and the precision form at its first and last points,
    and each form's Zin error against the plain form there."""
    from openpeec_amd import PEEC_FORM_MIXED, PEEC_FORM_PLAIN, Peec, decks
    t0 = time.perf_counter()
    m = Peec.from_deck_text(decks.full_stack_deck(), lib)
    m.setup()
    log(f"full stack: {m.p.nseg} current cells, {m.p.ncell} capacitive cells, {m.p.nunknown} unknowns "
        f"({time.perf_counter() - t0:.1f} s setup)")
    nf = m.p.nfreq
    npoints = max(2, min(npoints, nf))
    m.sweep(0, 1, PEEC_FORM_MIXED, 4)
    lib.peec_sweep_reset()
    t0 = time.perf_counter()
    z, st = m.sweep(0, npoints, PEEC_FORM_MIXED, 4)
    elapsed = time.perf_counter() - t0
    log(f"full stack: accelerated sweep of {npoints} points {elapsed:.1f} s ({st.n_lu} factorisations, "
        f"{st.gmres_its} GMRES iterations)")
    idx = [0, npoints - 1]
    zp, t_plain, stp = plain_points(m, idx)
    log(f"full stack: plain FP64 form {t_plain / len(idx):.1f} s per point")
    prec = precision_points(m, idx, zp)
    zm = np.array([z[i, 0, 0] for i in idx])
    err = float(np.max(np.abs(zm - zp) / np.abs(zp)))
    out = {
        "workload": "ground plane 0.1 mm under an FR4 brick (eps_r 4.4, 60 x 60 x 1.6 mm, 20 x 20 x 4 cells); "
                    "copper plate 60 x 60 mm, 35 um, 48 x 48 grid in 4 volume layers on the brick; port plate "
                    f"corner to ground; retarded + capacitive; points 20 MHz + k 20 MHz, k = 0..{npoints - 1}",
        "current_cells": m.p.nseg, "capacitive_cells": m.p.ncell, "unknowns": m.p.nunknown,
        "accelerated": {"points": npoints, "s": elapsed, "points_per_s": npoints / elapsed, "n_lu": st.n_lu,
                        "gmres_its_per_point": st.gmres_its / npoints, "kf_direct_pairs": st.kf_direct_pairs,
                        "stages_ms": {"fill_incl_kf_setup": st.fill_ms, "kf_setup": st.kf_setup_ms,
                                      "lu_and_inverse": st.lu_ms, "gmres": st.solve_ms},
                        "zin_max_rel_err_vs_plain_fp64": err},
        "plain_fp64_points_per_s": len(idx) / t_plain,
        "speedup_vs_plain_fp64": (npoints / elapsed) / (len(idx) / t_plain),
        "plain_stages_ms_last_point": {"lp_fill": stp.fill_ms, "charge_fill": stp.pot_ms, "p_lu_and_inverse": stp.cinv_ms,
                                       "mna_lu": stp.lu_ms, "total": stp.total_ms},
        "precision_form": prec,
        "zin_points_hz": [m.freq(i) for i in idx],
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
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU frequency sharding is "
                                   "described in DESIGN.md but not part of this build", "n_gpus": args.gpus}))
        sys.exit(2)
    from openpeec_amd import PEEC_FORM_MIXED, PEEC_FORM_PLAIN, Peec, decks, load_library

    lib = load_library(build_if_missing=True)
    if not lib.peec_device_available():
        raise SystemExit("no HIP device: the library has no CPU path")
    if args.only == "batch":
        print(json.dumps({"batching": bench_batch(lib, args.batch_structs)}))
        return
    if args.only == "learned":
        print(json.dumps({"learned_preconditioner": bench_learned(lib, args.learned_points)}))
        return
    if args.only == "fullstack":
        print(json.dumps({"full_stack": bench_fullstack(lib, args.fullstack_points)}))
        return
    text = decks.plate_deck(n=args.grid)
    m = Peec.from_deck_text(text, lib)
    m.setup()
    nf = m.p.nfreq
    n = m.p.nunknown
    mixed = args.form == "mixed"
    steps = args.steps if args.steps is not None else (nf if mixed else 3)
    steps = max(1, min(steps, nf))
    result = {"metric": METRIC}
    if mixed:
        # warm-up (kernel loading, allocations), then a cold sweep: preconditioner and fill
        # records are dropped so the timed region rebuilds them
        if args.warmup > 0:
            m.sweep(0, min(args.warmup, nf), PEEC_FORM_MIXED, args.batch)
        lib.peec_sweep_reset()
        t0 = time.perf_counter()
        z, st = m.sweep(0, steps, PEEC_FORM_MIXED, args.batch)
        elapsed = time.perf_counter() - t0
        idx = sorted({0, steps // 2, steps - 1})
        zm = np.array([z[i, 0, 0] for i in idx])
        if args.no_compare:
            print(json.dumps({"steps": steps, "ms_per_step": 1e3 * elapsed / steps, "value": steps / elapsed,
                              "stages_ms": {"fill": st.fill_ms, "lu_and_inverse": st.lu_ms, "gmres": st.solve_ms},
                              "gmres_its": st.gmres_its, "n_lu": st.n_lu,
                              "k_zgemm_avg_launch_ms": st.lu_gemm_ms / max(1, st.lu_gemm_launches),
                              "k_zgemm_avg_launch_flop": st.lu_gemm_flop / max(1, st.lu_gemm_launches),
                              "k_zgemm_launches": st.lu_gemm_launches,
                              "mv_avg_launch_ms": st.mv_ms / max(1, st.mv_launches), "mv_launches": st.mv_launches,
                              "mv_GB_per_s": (st.mv_bytes / (st.mv_ms * 1e-3)) / 1e9 if st.mv_ms > 0 else None}))
            m.close()
            return
        zp, t_plain, stp = plain_points(m, idx)
        prec = precision_points(m, idx, zp)
        fill_flop_top = lib.peec_fill_flops(C.byref(m.p), m.freq(idx[-1]))
        plain_fill_ms = stp.fill_ms + stp.pot_ms
        err = float(np.max(np.abs(zm - zp) / np.abs(zp)))
        gemm_ms = st.lu_gemm_ms / max(1, st.lu_gemm_launches)
        gemm_flop = st.lu_gemm_flop / max(1, st.lu_gemm_launches)
        achieved = (gemm_flop / (gemm_ms * 1e-3)) / 1e12 if gemm_ms > 0 else 0.0
        mv_gbs = (st.mv_bytes / (st.mv_ms * 1e-3)) / 1e9 if st.mv_ms > 0 else None
        fill_flop = lib.peec_fill_flops(C.byref(m.p), m.freq(steps - 1))
        result.update({
            "value": steps / elapsed, "unit": "freq points/s", "n_gpus": 1, "steps": steps,
            "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None,
            "dtype": "f64 (GMRES, residuals, fill); preconditioner factors stored f32",
            "data": "synthetic deck: plate geometry generated in-process (no dataset)",
            "config": {
                "workload": f"copper plate 100x100 mm, 35 um, {args.grid}x{args.grid} grid, retarded + capacitive, "
                            f"corner-to-corner port, 10 MHz-3 GHz sweep of {nf} points; one step = one frequency point",
                "unknowns": n, "charge_unknowns": m.p.ncell, "branches": m.p.nseg, "form": "mixed",
                "batch": args.batch, "points": [m.freq(0), m.freq(steps - 1)],
            },
            "zin_check": {"points_hz": [m.freq(i) for i in idx], "max_rel_err_vs_plain_fp64": err,
                          "bar": 1e-6, "ok": err <= 1e-6},
            "plain_fp64_points_per_s": len(idx) / t_plain,
            "speedup_vs_plain_fp64": (steps / elapsed) / (len(idx) / t_plain),
            "precision_form": prec,
            "stages_ms": {"fill_incl_kf_setup": st.fill_ms, "kf_setup": st.kf_setup_ms,
                          "lu_and_inverse": st.lu_ms, "gmres": st.solve_ms, "total": st.total_ms},
            "n_lu": st.n_lu, "gmres_its_per_point": st.gmres_its / steps, "kf_direct_pairs": st.kf_direct_pairs,
            "lu_update": (lu_roofline(achieved) | {
                "launches": st.lu_gemm_launches, "avg_launch_ms": gemm_ms, "avg_launch_flop": gemm_flop,
            }),
            "fill_nominal_gflop_per_point": fill_flop / 1e9,
            "plain_fill": {"gflops": fill_flop_top / (plain_fill_ms * 1e-3) / 1e9,
                           "frac_of_fp64_vector_peak": fill_flop_top / (plain_fill_ms * 1e-3) / 1e12 / PEAK_FP64_VECTOR,
                           "note": "element-by-element FP64 fill of the plain form at the last checked point"},
            "kfactored_fill_ms_per_point": (st.fill_ms - st.kf_setup_ms) / steps,
            "roofline": mv_roofline(mv_gbs, st.mv_launches, st.mv_ms / max(1, st.mv_launches),
                                    st.mv_bytes / max(1, st.mv_launches)),
        })
    else:
        form = PEEC_FORM_PLAIN
        idx = [int(round(x)) for x in np.linspace(0, nf - 1, steps + 2)[1:-1]]
        for w in range(args.warmup):
            m.sweep(nf - 1 - w, 1, form)
        t0 = time.perf_counter()
        for i in idx:
            m.sweep(i, 1, form)
        elapsed = time.perf_counter() - t0
        z_top, st = m.sweep(nf - 1, 1, form)
        fill_flop = lib.peec_fill_flops(C.byref(m.p), m.freq(nf - 1))
        fill_ms = st.fill_ms + st.pot_ms
        gemm_avg_ms = st.lu_gemm_ms / max(1, st.lu_gemm_launches)
        gemm_avg_flop = st.lu_gemm_flop / max(1, st.lu_gemm_launches)
        achieved = (gemm_avg_flop / (gemm_avg_ms * 1e-3)) / 1e12 if gemm_avg_ms > 0 else 0.0
        result.update({
            "value": len(idx) / elapsed, "unit": "freq points/s", "n_gpus": 1, "steps": len(idx),
            "warmup": args.warmup, "ms_per_step": 1e3 * elapsed / len(idx), "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64",
            "data": "synthetic deck: plate geometry generated in-process (no dataset)",
            "config": {
                "workload": f"copper plate 100x100 mm, 35 um, {args.grid}x{args.grid} grid, retarded + capacitive, "
                            f"corner-to-corner port, 10 MHz-3 GHz sweep; one step = one frequency point",
                "unknowns": n, "branches": m.p.nseg, "charge_cells": m.p.ncell, "form": "plain",
                "frequencies_timed_hz": [m.freq(i) for i in idx],
            },
            "stages_ms_top_freq": {
                "lp_fill": st.fill_ms, "charge_fill": st.pot_ms, "p_lu_and_inverse": st.cinv_ms,
                "mna_assembly": st.asm_ms, "mna_lu": st.lu_ms, "port_solve": st.solve_ms, "total": st.total_ms,
            },
            "fill_gflops": fill_flop / (fill_ms * 1e-3) / 1e9 if fill_ms > 0 else None,
            "fill_frac_of_fp64_vector_peak": (fill_flop / (fill_ms * 1e-3) / 1e12) / PEAK_FP64_VECTOR if fill_ms > 0 else None,
            "roofline": {
                "kernel": ("k_zgemm" if os.environ.get("PEEC_GEMM", "") == "hip" else "rocBLAS zgemm")
                          + " (LU trailing update, FP64 MFMA)",
                "bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_MFMA, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_MFMA, "traffic": None,
                "launches": st.lu_gemm_launches, "avg_launch_ms": gemm_avg_ms, "avg_launch_flop": gemm_avg_flop,
            },
        })
    if not args.no_check:
        result["small_deck_check"] = small_check(lib)
    if not args.no_batch:
        result["batching"] = bench_batch(lib, args.batch_structs)
    if not args.no_learned:
        result["learned_preconditioner"] = bench_learned(lib, args.learned_points)
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        from oracle_py import oracle_peec
        m_cpu = oracle_peec(text)
        result["cpu_baseline"] = cpu_baseline(m_cpu)
        m_cpu.close()
    m.close()
    print(json.dumps(result))


if __name__ == "__main__":
    main()
