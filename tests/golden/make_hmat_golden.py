"""Golden vectors of the compressed (compression = 1) path, written by the reference's own
functions (oracle/_ref/libpeec_ref.so, compiled in place from the reference's C sources):

  hm/<deck>/<k>/perm, far, near, rank   cluster order, far blocks (r0, m, c0, nn), near blocks,
                                         ACA ranks (0 = stored dense) of the Lp (k = lp) and P
                                         (k = p) H-matrices at the deck's first frequency
  hm/<deck>/<k>/x, y                     y = H x for a fixed seeded x
  hm/<deck>/pc/b, z                      z = M^-1 b (pc_apply) for a fixed seeded b
  hm/<deck>/zmat                         Z(f) of the reference's solve() (compressed path)
  hm/gen/...                             the same records of hmat_build_gen with a caller's kernel
                                         (tests/gen_kernel.c: a helix of 400 points, exp(-jkr) / (r + 0.01))
  hm/disk96/lp/...                       the reference's hmat_build (ACA 1e-12) on the learned-preconditioner
                                         geometry (48 x 96 disk under a 32 x 16 quad) at 1.27 GHz, plus
                                         ydense = Lp x from its lp_fill: its ACA keeps blocks with O(1)
                                         errors there (H x 1.2e-3 from the dense product).  Its parser stops at 32
                                         sectors, so the cells come from this library's wire_build (the
                                         disk's branches are built by the reference's rules for any
                                         sector count; only the centre charge cell differs, and Lp does
                                         not involve it)

Run in this container (the reference sources are read from /root/reference at build time of
oracle/_ref only):  make -C oracle ref && python tests/golden/make_hmat_golden.py
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parents[1]))

from openpeec_amd import abi, cview  # noqa: E402
from refharness import RefPeec, ref_available  # noqa: E402

P = C.POINTER


class ctree_t(C.Structure):
    _fields_ = [("lo", P(C.c_int)), ("hi", P(C.c_int)), ("left", P(C.c_int)), ("right", P(C.c_int)),
                ("cen", P(C.c_double)), ("rad", P(C.c_double)), ("nc", C.c_int), ("cap", C.c_int)]


class ref_hmat_t(C.Structure):
    """Layout of the reference's private struct hmat_t (hmatrix.c:54-75)."""
    _fields_ = [("n", C.c_int), ("perm", P(C.c_int)), ("ct", ctree_t), ("nfar", C.c_int), ("fcap", C.c_int),
                ("fr", P(C.c_int)), ("fc", P(C.c_int)), ("frank", P(C.c_int)),
                ("fu", P(P(abi.d_complex_t))), ("fv", P(P(abi.d_complex_t))),
                ("nnear", C.c_int), ("ncap", C.c_int), ("nr", P(C.c_int)), ("nc2", P(C.c_int)),
                ("nd", P(P(abi.d_complex_t))), ("nleaf", C.c_int), ("leafid", P(C.c_int)),
                ("rfoff", P(C.c_int)), ("rfidx", P(C.c_int)), ("rnoff", P(C.c_int)), ("rnidx", P(C.c_int)),
                ("ftoff", P(C.c_int)), ("ftlen", C.c_int)]


def with_sweep(text: str, f0: float, f1: float, nd: int, retard: int | None = None, cap: int | None = None) -> str:
    keep = [ln for ln in text.splitlines() if not ln.strip().startswith(("frequency", "end"))
            and not (retard is not None and ln.strip().startswith("retardation"))
            and not (cap is not None and ln.strip().startswith("capacitance"))]
    tail = []
    if retard is not None:
        tail.append(f"retardation = {retard}")
    if cap is not None:
        tail.append(f"capacitance = {cap}")
    tail += [f"frequency = {f0} {f1} {nd}", "end"]
    return "\n".join(keep + tail) + "\n"


def decks() -> dict[str, str]:
    d = HERE / "decks"
    return {
        "compress_wire": (d / "compress_wire.peec").read_text(),
        "compress_plate": (d / "compress_plate.peec").read_text(),
        "compress_cap": (d / "compress_cap.peec").read_text(),
        "compress_plate_ret": with_sweep((d / "compress_plate.peec").read_text(), 5e8, 1.5e9, 2, retard=1, cap=1),
        "compress_quad_ret": with_sweep((d / "quad_square.peec").read_text().replace("title", "compression = 1\ntitle", 1),
                                        2e8, 2e8, 0, retard=1, cap=1),
    }


def seeded(n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex128)


def arr(ptr, n) -> np.ndarray:
    return np.array([ptr[i] for i in range(n)], dtype=np.int32)


def blocks(h: ref_hmat_t, rows, cols, count) -> np.ndarray:
    out = np.zeros((count, 4), dtype=np.int32)
    for b in range(count):
        a, c = rows[b], cols[b]
        out[b] = (h.ct.lo[a], h.ct.hi[a] - h.ct.lo[a], h.ct.lo[c], h.ct.hi[c] - h.ct.lo[c])
    return out


def record(out: dict, key: str, lib, hp: int, seed: int) -> None:
    h = C.cast(hp, P(ref_hmat_t)).contents
    out[f"{key}/perm"] = arr(h.perm, h.n)
    out[f"{key}/far"] = blocks(h, h.fr, h.fc, h.nfar)
    out[f"{key}/near"] = blocks(h, h.nr, h.nc2, h.nnear)
    out[f"{key}/rank"] = arr(h.frank, h.nfar)
    x = seeded(h.n, seed)
    y = np.zeros(h.n, dtype=np.complex128)
    lib.hmat_matvec(hp, x.ctypes.data_as(P(abi.d_complex_t)), y.ctypes.data_as(P(abi.d_complex_t)))
    out[f"{key}/x"] = x
    out[f"{key}/y"] = y


# hmat_build_gen with a caller's kernel: points on a helix, A(i, j) = exp(-j k r) / (r + d0)
GEN_N, GEN_KW, GEN_TOL = 400, 12.0, 1e-6


def gen_points() -> np.ndarray:
    t = np.linspace(0, 6 * np.pi, GEN_N)
    return np.ascontiguousarray(np.stack([0.05 * np.cos(t), 0.05 * np.sin(t), 0.02 * t], axis=1).ravel())


def gen_kernel_lib() -> C.CDLL:
    """tests/gen_kernel.c (helix_kernel, an hmat_entry_fn) compiled into a temporary library"""
    import subprocess
    import tempfile
    out = Path(tempfile.mkdtemp()) / "libgenkernel.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", str(HERE.parent / "gen_kernel.c"), "-o", str(out), "-lm"], check=True)
    return C.CDLL(str(out))


def gen_dense(pts: np.ndarray) -> np.ndarray:
    xyz = pts.reshape(-1, 3)
    d = np.linalg.norm(xyz[:, None, :] - xyz[None, :, :], axis=2)
    return np.exp(-1j * GEN_KW * d) / (d + 0.01)


def build_gen(lib, pts: np.ndarray) -> int:
    klib = gen_kernel_lib()
    lib.hmat_build_gen.restype = C.c_void_p
    lib.hmat_build_gen.argtypes = [C.c_int, P(C.c_double), C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_char_p,
                                   C.c_void_p]
    fn = C.cast(klib.helix_kernel, C.c_void_p)
    return lib.hmat_build_gen(GEN_N, pts.ctypes.data_as(P(C.c_double)), fn, pts.ctypes.data_as(C.c_void_p), GEN_KW, GEN_TOL,
                              b"gen", None)


def record_gen(out: dict, lib) -> None:
    pts = gen_points()
    h = build_gen(lib, pts)
    assert h
    record(out, "hm/gen", lib, h, 17)
    lib.hmat_free(h)


DISK96 = dict(nr=48, nsec=96, qx=32, qy=16, ctol=1e-12, f0=1.27e9, f1=1.27e9, ndiv=0)


def disk96_text() -> str:
    from openpeec_amd import decks as dk
    return dk.disk_quad_deck(**DISK96)


def record_disk96(out: dict, lib) -> None:
    from openpeec_amd import Peec, load_library
    from refharness import devnull
    m = Peec.from_deck_text(disk96_text(), load_library(build_if_missing=False))
    m.setup()
    f = DISK96["f0"]
    h = lib.hmat_build(C.byref(m.p), f, None)
    assert h
    record(out, "hm/disk96/lp", lib, h, 23)
    lib.hmat_free(h)
    lib.lp_fill(C.byref(m.p), f, devnull())
    n = m.p.nseg
    out["hm/disk96/lp/ydense"] = cview(m.p.lp, (n, n)) @ out["hm/disk96/lp/x"]
    m.close()


def main() -> None:
    if not ref_available():
        raise SystemExit("reference build missing: run `make -C oracle ref` first")
    out: dict[str, np.ndarray] = {}
    from refharness import ref_lib
    record_gen(out, ref_lib())
    record_disk96(out, ref_lib())
    for name, text in decks().items():
        r = RefPeec(text)
        r.setup_like_main()
        lib, p = r.lib, r.p
        f = abi_freq(p, 0)
        hl = lib.hmat_build(C.byref(p), f, None)
        assert hl, name
        record(out, f"hm/{name}/lp", lib, hl, 11)
        blk = (abi.hblock_t * 2)()
        nblk = 0
        blk[nblk] = abi.hblock_t(p.offS, p.nseg, hl, abi.d_complex_t(0.0, -2 * np.pi * f))
        nblk += 1
        hp = None
        if p.capacitance and p.ncell > 0:
            hp = lib.hmat_build_pot(C.byref(p), f, None)
            assert hp, name
            record(out, f"hm/{name}/p", lib, hp, 13)
            blk[nblk] = abi.hblock_t(p.offQ, p.ncell, hp, abi.d_complex_t(1.0, 0.0))
            nblk += 1
        sp = abi.mna_sparse_t()
        lib.mna_sparse(C.byref(p), f, C.byref(sp))
        pc = lib.pc_build(C.byref(p), C.byref(sp), blk, nblk)
        assert pc, name
        b = seeded(p.nunknown, 12)
        z = b.copy()
        lib.pc_apply(pc, z.ctypes.data_as(P(abi.d_complex_t)))
        out[f"hm/{name}/pc/b"] = b
        out[f"hm/{name}/pc/z"] = z
        lib.pc_free(pc)
        lib.mna_sparse_free(C.byref(sp))
        lib.hmat_free(hl)
        if hp:
            lib.hmat_free(hp)
        r.solve()
        out[f"hm/{name}/zmat"] = r.zmat()
        r.close()
    np.savez_compressed(HERE / "hmat_vectors.npz", **out)
    print(f"wrote {len(out)} arrays to {HERE / 'hmat_vectors.npz'}")


def abi_freq(p, i: int) -> float:
    if p.nfreq <= 1:
        return p.f0
    if p.flog:
        return p.f0 * (p.f1 / p.f0) ** (i / (p.nfreq - 1))
    return p.f0 + (p.f1 - p.f0) * i / (p.nfreq - 1)


if __name__ == "__main__":
    main()
