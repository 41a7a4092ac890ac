"""Learned near-field correction of the compressed solver's preconditioner.

  P^-1 = M^-1 - theta M^-1 N M^-1: M is precond.c's preconditioner (possibly built at another
  frequency of the sweep), N the near-field blocks off the block diagonal, theta a complex weight
  from a small network (8 -> 16 -> 16 -> 2) trained on the device by the library's own kernels.

CPU: the oracle's statement of the network (oracle/mlp.c) is pinned by central finite
differences of its own loss (every parameter) and by the closed-form optimum of one sample with
the hidden layers switched off.
GPU: forward and gradient kernels against the oracle (1e-12); training on samples the solver
labels itself (the weight with the fewest GMRES iterations over a grid, per training system of
three geometries and twelve frequency pairs, the reference frequency up to 2.5 times below) brings the loss most of the way to the labels; and
on a held-out geometry the learned sweep needs fewer GMRES iterations than the same sweep
without the network, with the same Z (true residual 1e-8 both ways) as the reference-rule
compressed sweep.
"""
import ctypes as C

import numpy as np
import pytest

from openpeec_amd import PEEC_FORM_LEARNED, PEEC_FORM_PLAIN, Peec, decks
from oracle_py import oracle_lib

NF, NP = 8, 450
PD = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(PD)


def _orc():
    lib = oracle_lib()
    lib.orc_mlp_loss_grad.restype = C.c_double
    lib.orc_mlp_loss_grad.argtypes = [PD, PD, PD, C.c_int, PD]
    lib.orc_mlp_forward.argtypes = [PD, PD, C.c_int, PD]
    lib.orc_mlp_nparams.restype = C.c_int
    return lib


def _problem(seed, S=37):
    rng = np.random.default_rng(seed)
    W = rng.standard_normal(NP) * 0.4
    X = rng.standard_normal((S, NF))
    c = rng.uniform(0.5, 2.0, S)
    d = rng.uniform(1.0, 3.0, S)
    p = (rng.standard_normal(S) + 1j * rng.standard_normal(S)) * np.sqrt(c * d) * 0.5
    coef = np.stack([p.real, p.imag, c, d], axis=1).ravel()
    return W, np.ascontiguousarray(X.ravel()), np.ascontiguousarray(coef)


def test_oracle_gradient_matches_finite_differences():
    lib = _orc()
    assert lib.orc_mlp_nparams() == NP
    W, X, coef = _problem(1)
    g = np.zeros(NP)
    lib.orc_mlp_loss_grad(_p(W), _p(X), _p(coef), 37, _p(g))
    h = 1e-6
    fd = np.zeros(NP)
    for e in range(NP):
        wp, wm = W.copy(), W.copy()
        wp[e] += h
        wm[e] -= h
        fd[e] = (lib.orc_mlp_loss_grad(_p(wp), _p(X), _p(coef), 37, None) -
                 lib.orc_mlp_loss_grad(_p(wm), _p(X), _p(coef), 37, None)) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))


def test_oracle_single_sample_optimum():
    """Hidden layers off (W1, b1, W2, b2, W3 = 0): theta = b3, whose optimum for one sample is
    -conj(p) / c with loss 1 - |p|^2 / (c d)."""
    lib = _orc()
    p, c, d = 0.3 - 0.4j, 1.7, 2.0
    coef = np.array([p.real, p.imag, c, d])
    X = np.zeros(NF)
    W = np.zeros(NP)
    th = -np.conj(p) / c
    W[NP - 2], W[NP - 1] = th.real, th.imag
    g = np.zeros(NP)
    loss = lib.orc_mlp_loss_grad(_p(W), _p(X), _p(coef), 1, _p(g))
    assert abs(loss - (1 - abs(p) ** 2 / (c * d))) <= 1e-15
    assert np.max(np.abs(g[NP - 2:])) <= 1e-15
    theta = np.zeros(2)
    lib.orc_mlp_forward(_p(W), _p(X), 1, _p(theta))
    assert abs(theta[0] - th.real) <= 1e-16 and abs(theta[1] - th.imag) <= 1e-16


@pytest.mark.gpu
def test_device_network_matches_oracle(lib):
    orc = _orc()
    assert lib.peec_lpc_nparams() == NP and lib.peec_lpc_nfeatures() == NF
    for seed in (2, 3):
        W, X, coef = _problem(seed, S=300)
        g0 = np.zeros(NP)
        l0 = orc.orc_mlp_loss_grad(_p(W), _p(X), _p(coef), 300, _p(g0))
        g1 = np.zeros(NP)
        l1 = np.zeros(1)
        assert lib.peec_lpc_mlp_grad(_p(W), _p(X), _p(coef), 300, _p(g1), _p(l1)) == 0
        assert abs(l1[0] - l0) <= 1e-12 * abs(l0)
        assert np.max(np.abs(g1 - g0)) <= 1e-12 * np.max(np.abs(g0))
        t0, t1 = np.zeros(600), np.zeros(600)
        orc.orc_mlp_forward(_p(W), _p(X), 300, _p(t0))
        lib.peec_lpc_mlp_forward(_p(W), _p(X), 300, _p(t1))
        assert np.max(np.abs(t1 - t0)) <= 1e-13 * max(1.0, np.max(np.abs(t0)))


def _samples(lib, geoms, pairs):
    Xs, Cs = [], []
    for g in geoms:
        m = Peec.from_deck_text(decks.disk_quad_deck(**g, f0=1e8, f1=4e9, ndiv=4), lib)
        m.setup()
        fr = np.array([a for a, _ in pairs], dtype=np.float64)
        fv = np.array([b for _, b in pairs], dtype=np.float64)
        X = np.zeros(len(pairs) * NF)
        coef = np.zeros(len(pairs) * 4)
        its = np.zeros(len(pairs) * 7, dtype=np.int32)
        print(f"[lpc] samples {g}", flush=True)
        assert lib.peec_lpc_samples(C.byref(m.p), _p(fr), _p(fv), len(pairs), _p(X), _p(coef),
                                    its.ctypes.data_as(C.POINTER(C.c_int))) == 0
        print(f"  iterations over theta 0..0.6: {its.reshape(-1, 7).tolist()}", flush=True)
        Xs.append(X.reshape(-1, NF))
        Cs.append(coef.reshape(-1, 4))
        m.close()
    return np.ascontiguousarray(np.concatenate(Xs).ravel()), np.ascontiguousarray(np.concatenate(Cs).ravel())


TRAIN_GEOMS = [dict(nr=8, nsec=16, qx=8, qy=4), dict(nr=10, nsec=20, qx=10, qy=6, gap=0.004),
               dict(nr=12, nsec=24, qx=12, qy=6, gap=0.007)]
TRAIN_PAIRS = [(fr, fr * s) for fr in (4e8, 1e9, 2e9) for s in (1.0, 1.3, 1.8, 2.5)]


@pytest.mark.gpu
def test_training_and_gain_on_held_out_geometry(lib):
    X, coef = _samples(lib, TRAIN_GEOMS, TRAIN_PAIRS)
    S = len(coef) // 4
    W = np.zeros(NP)
    hist = np.zeros(1501)
    print(f"[lpc] fit on {S} samples", flush=True)
    assert lib.peec_lpc_fit(_p(X), _p(coef), S, 1500, 3e-3, 7, _p(W), _p(hist)) == 0
    c = coef.reshape(-1, 4)
    best = float(np.mean(1 / c[:, 3]))
    print(f"  loss {hist[0]:.4f} -> {hist[-1]:.4f} (labels met exactly: {best:.4f})", flush=True)
    assert hist[-1] < hist[0] and hist[-1] - best < 0.25 * (hist[0] - best)
    lib.peec_lpc_set_model(_p(W))
    text = decks.disk_quad_deck(nr=14, nsec=28, qx=14, qy=8, gap=0.0055, f0=4e8, f1=3.6e9, ndiv=9)
    m = Peec.from_deck_text(text, lib)
    m.setup()
    print(f"[lpc] held-out geometry: {m.p.nunknown} unknowns", flush=True)
    z_l, st_l = m.sweep(0, m.nfreq, PEEC_FORM_LEARNED)
    lib.peec_lpc_set_model(None)
    z_0, st_0 = m.sweep(0, m.nfreq, PEEC_FORM_LEARNED)
    z_r, _ = m.sweep(0, m.nfreq, PEEC_FORM_PLAIN)
    print(f"  learned: {st_l.gmres_its} its, {st_l.pc_builds} builds, {st_l.total_ms:.0f} ms, "
          f"mean |theta| {st_l.theta_abs_sum / m.nfreq:.3f}, fallbacks {st_l.lpc_fallbacks}", flush=True)
    print(f"  without: {st_0.gmres_its} its, {st_0.pc_builds} builds, {st_0.total_ms:.0f} ms", flush=True)
    rl = float(np.max(np.abs(z_l - z_r) / np.abs(z_r)))
    r0 = float(np.max(np.abs(z_0 - z_r) / np.abs(z_r)))
    print(f"  Z vs reference-rule sweep: learned {rl:.2e}, without {r0:.2e}", flush=True)
    assert rl <= 1e-6 and r0 <= 1e-6
    assert st_l.gmres_its < st_0.gmres_its
    m.close()
