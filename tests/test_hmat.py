"""Compressed (compression = 1) path against the reference's own outputs (tests/golden/
hmat_vectors.npz, written by tests/golden/make_hmat_golden.py with the reference build) and
against the dense FP64 oracle.

CPU: the cluster tree / block partition (host code) is integer-identical to the reference's.
GPU: ACA ranks, y = H x, z = M^-1 b and Z(f) through the C-ABI:
  * H x within 2e-6 of the reference's (both are ACA approximations at ctol = 1e-6 of the same
    entries; identical ranks and pivots make them agree to rounding);
  * the preconditioner uses no approximated block (leaf diagonal blocks, sparse couplings, the
    Schur complement); the device applies the Schur complement through explicit inverses of its
    LU factors where the reference substitutes, which moves M^-1 b by cond(S) x eps: 1e-8;
  * Z within 2e-6 of the reference's compressed solve and 2e-5 of the dense oracle (the
    reference's own compressed Z is 7.6e-6 from dense on these decks).
"""
import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

from openpeec_amd import Peec, abi

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE / "golden"))
from make_hmat_golden import decks, seeded  # noqa: E402

GOLD = np.load(HERE / "golden" / "hmat_vectors.npz")
DECKS = decks()
P = C.POINTER


def _pts(m, kind):
    if kind == "lp":
        segs = m.p.seg
        return np.array([[0.5 * (segs[i].x1[c] + segs[i].x2[c]) for c in range(3)] for i in range(m.nseg)]).ravel()
    nc = m.p.ncell
    return np.ctypeslib.as_array(m.p.cpt, shape=(3 * nc,)).copy()


def _tree(lib, pts):
    n = len(pts) // 3
    cnt = (C.c_int * 4)()
    pa = np.ascontiguousarray(pts, dtype=np.float64)
    assert lib.hmat_tree_info(n, pa.ctypes.data_as(P(C.c_double)), cnt, None, None, None) == 0
    perm = np.zeros(n, dtype=np.int32)
    far = np.zeros((cnt[0], 4), dtype=np.int32)
    near = np.zeros((cnt[1], 4), dtype=np.int32)
    lib.hmat_tree_info(n, pa.ctypes.data_as(P(C.c_double)), cnt, perm.ctypes.data_as(P(C.c_int)),
                       far.ctypes.data_as(P(C.c_int)), near.ctypes.data_as(P(C.c_int)))
    return perm, far, near


@pytest.mark.parametrize("name", sorted(DECKS))
def test_block_partition_matches_reference(lib, name):
    m = Peec.from_deck_text(DECKS[name], lib)
    m.setup()
    for kind in ("lp", "p"):
        key = f"hm/{name}/{kind}"
        if f"{key}/perm" not in GOLD.files:
            continue
        perm, far, near = _tree(lib, _pts(m, kind))
        assert np.array_equal(perm, GOLD[f"{key}/perm"]), (name, kind)
        assert np.array_equal(far, GOLD[f"{key}/far"]), (name, kind)
        assert np.array_equal(near, GOLD[f"{key}/near"]), (name, kind)
    m.close()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(DECKS))
def test_hmatrix_ranks_and_product(lib, name):
    m = Peec.from_deck_text(DECKS[name], lib)
    m.setup()
    f = m.freq(0)
    for kind, build in (("lp", lib.hmat_build), ("p", lib.hmat_build_pot)):
        key = f"hm/{name}/{kind}"
        if f"{key}/perm" not in GOLD.files:
            continue
        print(f"[hmat] {name} {kind}", flush=True)
        h = build(C.byref(m.p), f, None)
        assert h
        cnt = (C.c_int * 4)()
        lib.hmat_info(h, cnt, None, None)
        ranks = np.zeros(max(1, cnt[0]), dtype=np.int32)
        perm = np.zeros(cnt[3], dtype=np.int32)
        lib.hmat_info(h, cnt, perm.ctypes.data_as(P(C.c_int)), ranks.ctypes.data_as(P(C.c_int)))
        gr = GOLD[f"{key}/rank"]
        assert np.array_equal(perm, GOLD[f"{key}/perm"])
        ours = np.where(ranks[:cnt[0]] < 0, 0, ranks[:cnt[0]])
        if gr.size:
            assert np.mean(ours == gr) >= 0.9 and np.max(np.abs(ours - gr)) <= 3, (ours, gr)
        x = GOLD[f"{key}/x"]
        y = np.zeros_like(x)
        lib.hmat_matvec(h, x.ctypes.data_as(P(abi.d_complex_t)), y.ctypes.data_as(P(abi.d_complex_t)))
        assert _rel(y, GOLD[f"{key}/y"]) <= 2e-6, (name, kind, _rel(y, GOLD[f"{key}/y"]))
        lib.hmat_free(h)
    m.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(DECKS))
def test_bulk_entries_give_the_same_hmatrix(lib, name, monkeypatch):
    """The entries evaluated in bulk before the build (default) and inside ACA and the dense
    blocks (PEEC_HM_FULL=0) are the same numbers, so ranks must be identical and H x must agree
    to rounding; a bulk matrix stored transposed, or indexed by original instead of permuted
    positions, changes the pivots and the product at O(1)."""
    m = Peec.from_deck_text(DECKS[name], lib)
    m.setup()
    f = m.freq(m.nfreq - 1)
    rng = np.random.default_rng(7)
    for kind, build in (("lp", lib.hmat_build), ("p", lib.hmat_build_pot)):
        out = []
        for mode in ("0", "1"):
            monkeypatch.setenv("PEEC_HM_FULL", mode)
            print(f"[hmat bulk] {name} {kind} PEEC_HM_FULL={mode}", flush=True)
            h = build(C.byref(m.p), f, None)
            if not h:
                pytest.skip(f"{name}: no {kind} matrix")
            cnt = (C.c_int * 4)()
            lib.hmat_info(h, cnt, None, None)
            ranks = np.zeros(max(1, cnt[0]), dtype=np.int32)
            lib.hmat_info(h, cnt, None, ranks.ctypes.data_as(P(C.c_int)))
            n = cnt[3]
            x = np.ascontiguousarray(rng.standard_normal(n) + 1j * rng.standard_normal(n)) if mode == "0" else x
            y = np.zeros_like(x)
            lib.hmat_matvec(h, x.ctypes.data_as(P(abi.d_complex_t)), y.ctypes.data_as(P(abi.d_complex_t)))
            out.append((ranks[:cnt[0]].copy(), y))
            lib.hmat_free(h)
        assert np.array_equal(out[0][0], out[1][0]), (name, kind)
        assert _rel(out[1][1], out[0][1]) <= 1e-13, (name, kind, _rel(out[1][1], out[0][1]))
    # over a sweep the bulk entries come from the band's k-factored fill instead
    zs = []
    for mode in ("0", "1"):
        monkeypatch.setenv("PEEC_HM_FULL", mode)
        print(f"[hmat bulk] {name} sweep PEEC_HM_FULL={mode}", flush=True)
        z, _ = m.sweep(0, m.nfreq)
        zs.append(z)
    assert np.max(np.abs(zs[1] - zs[0]) / np.abs(zs[0])) <= 1e-12
    m.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(DECKS))
def test_preconditioner_matches_reference(lib, name):
    m = Peec.from_deck_text(DECKS[name], lib)
    m.setup()
    p = m.p
    f = m.freq(0)
    print(f"[pc] {name}", flush=True)
    blk = (abi.hblock_t * 2)()
    nblk = 0
    hl = lib.hmat_build(C.byref(p), f, None)
    blk[nblk] = abi.hblock_t(p.offS, p.nseg, hl, abi.d_complex_t(0.0, -2 * np.pi * f))
    nblk += 1
    hp = None
    if p.capacitance and p.ncell > 0:
        hp = lib.hmat_build_pot(C.byref(p), f, None)
        blk[nblk] = abi.hblock_t(p.offQ, p.ncell, hp, abi.d_complex_t(1.0, 0.0))
        nblk += 1
    sp = abi.mna_sparse_t()
    lib.mna_sparse(C.byref(p), f, C.byref(sp))
    pc = lib.pc_build(C.byref(p), C.byref(sp), blk, nblk)
    assert pc
    z = GOLD[f"hm/{name}/pc/b"].copy()
    lib.pc_apply(pc, z.ctypes.data_as(P(abi.d_complex_t)))
    assert _rel(z, GOLD[f"hm/{name}/pc/z"]) <= 1e-8, _rel(z, GOLD[f"hm/{name}/pc/z"])
    # GMRES on the compressed operator reproduces the port solution of the dense system
    b = np.zeros(p.nunknown, dtype=np.complex128)
    lib.mna_rhs_port(C.byref(p), 0, b.ctypes.data_as(P(abi.d_complex_t)))
    x = b.copy()
    its = lib.gmres_hmat(C.byref(p), C.byref(sp), blk, nblk, pc, x.ctypes.data_as(P(abi.d_complex_t)), 1e-10)
    assert its > 0
    r = np.zeros_like(x)
    lib.mna_apply(C.byref(p), C.byref(sp), blk, nblk, x.ctypes.data_as(P(abi.d_complex_t)),
                  r.ctypes.data_as(P(abi.d_complex_t)), None)
    # 1e-10 where the deck is well scaled; the quad sheet is shorted by 1e-9 ohm bus resistors
    # (conductance 1e9 beside kernel entries many orders smaller), whose attainable residual is ~1e-5
    bar = 1e-4 if name == "compress_quad_ret" else 1.0001e-10
    assert np.linalg.norm(r - b) <= bar * np.linalg.norm(b), np.linalg.norm(r - b) / np.linalg.norm(b)
    lib.pc_free(pc)
    lib.mna_sparse_free(C.byref(sp))
    lib.hmat_free(hl)
    if hp:
        lib.hmat_free(hp)
    m.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(DECKS))
def test_compressed_sweep_matches_reference_and_dense(lib, name):
    from oracle_py import oracle_zmat
    text = DECKS[name]
    print(f"[hsweep] {name}", flush=True)
    m = Peec.from_deck_text(text, lib)
    m.setup()
    z, st = m.sweep(0, m.nfreq)
    zr = GOLD[f"hm/{name}/zmat"]
    assert float(np.max(np.abs(z - zr) / np.abs(zr))) <= 2e-6
    dense = "\n".join(ln for ln in text.splitlines() if not ln.strip().startswith("compression")) + "\n"
    zo = oracle_zmat(dense)
    assert float(np.max(np.abs(z - zo) / np.abs(zo))) <= 2e-5
    assert st.gmres_its > 0
    m.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(DECKS))
def test_leaf_blocks(lib, name):
    """hmat_nleaf / hmat_leaf_block (hmatrix.c:807-824): the leaves tile [0, n) in the
    reference's cluster order (the golden partition's diagonal near blocks), and each leaf's
    diagonal block holds the entries lp_entry / pot_entry give (never approximated)."""
    m = Peec.from_deck_text(DECKS[name], lib)
    m.setup()
    f = m.freq(0)
    kw = m.kw(f)
    for kind, build, entry in (("lp", lib.hmat_build, lib.lp_entry), ("p", lib.hmat_build_pot, lib.pot_entry)):
        key = f"hm/{name}/{kind}"
        if f"{key}/perm" not in GOLD.files:
            continue
        print(f"[leaf] {name} {kind}", flush=True)
        h = build(C.byref(m.p), f, None)
        assert h
        nleaf = lib.hmat_nleaf(h)
        gperm, gnear = GOLD[f"{key}/perm"], GOLD[f"{key}/near"]
        diag = sorted((int(r0), int(mm)) for r0, mm, c0, nn in gnear if r0 == c0 and mm == nn)
        assert nleaf == len(diag)
        seen = []
        for r in range(nleaf):
            idx = P(C.c_int)()
            mm = C.c_int()
            blk = lib.hmat_leaf_block(h, r, C.byref(idx), C.byref(mm))
            assert blk
            ids = [idx[a] for a in range(mm.value)]
            seen += ids
            assert any(sz == mm.value and list(gperm[r0:r0 + sz]) == ids for r0, sz in diag), (name, kind, r)
            if r not in (0, nleaf // 2, nleaf - 1):
                continue   # entries checked on three leaves (one device call per entry)
            for a in range(mm.value):
                for b in range(mm.value):
                    e = entry(C.byref(m.p), ids[a], ids[b], kw)
                    want = complex(e.r, e.i)
                    got = complex(blk[a * mm.value + b].r, blk[a * mm.value + b].i)
                    assert abs(got - want) <= 1e-12 * max(abs(want), 1e-300), (name, kind, r, a, b, got, want)
        assert sorted(seen) == list(range(len(gperm)))
        lib.hmat_free(h)
    m.close()


@pytest.mark.gpu
def test_build_gen_with_a_caller_kernel(lib):
    """hmat_build_gen (hmatrix.c:403-614) with a kernel the caller supplies (tests/gen_kernel.c):
    the block partition integer-identical to the reference's, ACA ranks as close as for the
    library's kernels, and H x within 2e-6 of the reference's H x and of the dense product."""
    from make_hmat_golden import GEN_TOL, build_gen, gen_dense, gen_points
    pts = gen_points()
    h = build_gen(lib, pts)
    assert h
    cnt = (C.c_int * 4)()
    lib.hmat_info(h, cnt, None, None)
    perm = np.zeros(cnt[3], dtype=np.int32)
    ranks = np.zeros(max(1, cnt[0]), dtype=np.int32)
    lib.hmat_info(h, cnt, perm.ctypes.data_as(P(C.c_int)), ranks.ctypes.data_as(P(C.c_int)))
    assert np.array_equal(perm, GOLD["hm/gen/perm"])
    assert cnt[0] == len(GOLD["hm/gen/far"]) and cnt[1] == len(GOLD["hm/gen/near"])
    gr = GOLD["hm/gen/rank"]
    ours = np.where(ranks[:cnt[0]] < 0, 0, ranks[:cnt[0]])
    assert np.mean(ours == gr) >= 0.9 and np.max(np.abs(ours - gr)) <= 3
    x = GOLD["hm/gen/x"]
    y = np.zeros_like(x)
    lib.hmat_matvec(h, x.ctypes.data_as(P(abi.d_complex_t)), y.ctypes.data_as(P(abi.d_complex_t)))
    lib.hmat_free(h)
    assert _rel(y, GOLD["hm/gen/y"]) <= 2e-6
    dense = gen_dense(pts) @ x
    assert np.linalg.norm(y - dense) <= 10 * GEN_TOL * np.linalg.norm(dense)


def test_disk96_fixture_shows_the_reference_aca_failure(lib):
    """The 48 x 96 disk under a 32 x 16 quad at ACA 1e-12: the reference's own H x (recorded by
    make_hmat_golden.py) is 1e-3 from the reference's dense product, nine orders above the
    tolerance (blocks whose rows are exact multiples of each other stop at a zero residual row
    with O(1) error); the cluster tree here is integer-identical to the one it used."""
    from make_hmat_golden import disk96_text
    key = "hm/disk96/lp"
    y, yd = GOLD[f"{key}/y"], GOLD[f"{key}/ydense"]
    assert _rel(y, yd) > 1e-4
    m = Peec.from_deck_text(disk96_text(), lib)
    m.setup()
    perm, far, near = _tree(lib, _pts(m, "lp"))
    m.close()
    assert np.array_equal(perm, GOLD[f"{key}/perm"])
    assert np.array_equal(far, GOLD[f"{key}/far"]) and np.array_equal(near, GOLD[f"{key}/near"])


@pytest.mark.gpu
def test_aca_column_check_repairs_the_disk96_blocks(lib):
    """Same geometry on the device: the column check at ACA termination resumes the blocks the
    reference leaves unconverged, so H x is within 1e-5 of the reference's dense Lp x (the
    reference's own H x misses that bar a hundredfold; what remains is a few single entries at
    the element code's far-field switch, 20 lmax, which no low-rank block follows), while the
    blocks the reference approximates within tolerance keep the reference's rank (the check
    never fires there)."""
    from make_hmat_golden import DISK96, disk96_text
    key = "hm/disk96/lp"
    m = Peec.from_deck_text(disk96_text(), lib)
    m.setup()
    print("[hmat] disk96 lp", flush=True)
    h = lib.hmat_build(C.byref(m.p), DISK96["f0"], None)
    assert h
    cnt = (C.c_int * 4)()
    lib.hmat_info(h, cnt, None, None)
    ranks = np.zeros(max(1, cnt[0]), dtype=np.int32)
    perm = np.zeros(cnt[3], dtype=np.int32)
    lib.hmat_info(h, cnt, perm.ctypes.data_as(P(C.c_int)), ranks.ctypes.data_as(P(C.c_int)))
    assert np.array_equal(perm, GOLD[f"{key}/perm"])
    gr = GOLD[f"{key}/rank"]
    ours = np.where(ranks[:cnt[0]] < 0, 0, ranks[:cnt[0]])
    assert np.mean(ours == gr) >= 0.95, np.mean(ours == gr)
    x = GOLD[f"{key}/x"]
    y = np.zeros_like(x)
    lib.hmat_matvec(h, x.ctypes.data_as(P(abi.d_complex_t)), y.ctypes.data_as(P(abi.d_complex_t)))
    lib.hmat_free(h)
    m.close()
    err = _rel(y, GOLD[f"{key}/ydense"])
    print(f"[hmat] disk96 |Hx - Ax| / |Ax| = {err:.3e} (reference {_rel(GOLD[f'{key}/y'], GOLD[f'{key}/ydense']):.3e})", flush=True)
    assert err <= 1e-5
