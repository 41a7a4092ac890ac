"""Golden result files of the reference's whole program (oracle/_ref/peec_ref, its own C sources
with its main.c, compiled in place by oracle/Makefile) on decks that exercise every output:
zin.csv, peec.sNp (S parameters), far.csv (far field), dist.csv (distributions), pw.csv (plane
wave), tran.csv (transient), and the node voltages of the sources solve in peec.log.  Run in
the container that has the reference sources:

    make -C oracle ref && python tests/golden/make_output_golden.py

-> tests/golden/output_vectors.npz: out/<deck>/<file> = the file's numeric columns (header
dropped; dist.csv's type column as 0 = I, 1 = Ipw, 2 = Q; tran.csv's as 0 = X, 1 = S, 2 = V;
peec.log's node-voltage tables as rows (node, re, im, |V|)).  The decks are the reference's
samples under tests/golden/decks plus OUTPUT_DECKS below.
"""
from __future__ import annotations

import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
DECKS = HERE / "decks"
REF_BIN = ROOT / "oracle" / "_ref" / "peec_ref"

SAMPLE_DECKS = ["dipole_ff", "dipole_pw", "bent_pw", "loop_tran", "tdr_resistor", "tnetwork_spara", "tnetwork_l_spara",
                "twoport_dist", "wire_dist"]


def _deck(name: str) -> str:
    return (DECKS / f"{name}.peec").read_text()


def _insert(text: str, lines: list[str]) -> str:
    """text with extra keyword lines after its title line"""
    out = []
    for ln in text.splitlines():
        out.append(ln)
        if ln.strip().startswith("title"):
            out += lines
    return "\n".join(out) + "\n"


def output_decks() -> dict[str, str]:
    """Decks written here (their features have no reference sample): independent sources with the
    node-voltage log, LU reuse (acceleration = 1), a ground-plane far field, and the compressed
    path's distributions and far field."""
    d = {name: _deck(name) for name in SAMPLE_DECKS}
    d["rlc_sources"] = "\n".join([
        "OpenPEEC 1 0", "title = series RLC driven by sources", "resistor = 1 2 50", "inductor = L1 2 3 1e-6",
        "capacitor = 3 0 1e-9", "vsource = 4 0 1.0 30", "resistor = 4 1 10", "isource = 0 2 0.01 -45",
        "port = 1 0 50", "frequency = 5e6 1e7 4", "end", ""])
    d["dipole_accel"] = _insert(_deck("dipole_halfwave"), ["acceleration = 1", "distribution = 1"]).replace(
        "frequency = 1.20e8 1.60e8 12", "frequency = 1.20e8 1.60e8 24")
    d["monopole_ff"] = _insert(_deck("monopole_gp"), ["farfield = 18 24"])
    d["compress_dist"] = "\n".join([
        "OpenPEEC 1 0", "title = compressed capacitive dipole with distributions", "compression = 1 1e-10",
        "capacitance = 1", "retardation = 1", "distribution = 1", "farfield = 12 12", "node = 1 0 0 -0.0005",
        "node = 2 0 0 0.0005", "wire = 0 0 -0.25 0 0 -0.0005 1e-3 5.8e7 40", "wire = 0 0 0.0005 0 0 0.25 1e-3 5.8e7 40",
        "port = 1 2 50", "frequency = 2e8 4e8 2", "end", ""])
    return d


def _table(path: Path, codes: dict[str, int] | None = None) -> np.ndarray:
    rows = []
    for ln in path.read_text().splitlines()[1:]:
        f = ln.split(",")
        if codes is not None:
            f[0] = str(codes[f[0]])
        rows.append([float(v) for v in f])
    return np.array(rows, dtype=np.float64)


def _touchstone(path: Path) -> np.ndarray:
    vals = []
    for ln in path.read_text().splitlines():
        if ln.startswith(("!", "#")) or not ln.strip():
            continue
        vals += [float(v) for v in ln.split()]
    return np.array(vals, dtype=np.float64)


def _node_voltages(path: Path) -> np.ndarray:
    rows, on = [], False
    for ln in path.read_text().splitlines():
        if ln.startswith("=== node voltage ==="):
            on = True
            continue
        if on:
            f = ln.split()
            if len(f) == 4 and f[0].lstrip("-").isdigit():
                rows.append([float(v) for v in f])
            elif not ln.startswith("  node"):
                on = False
    return np.array(rows, dtype=np.float64)


def run_reference(text: str, work: Path) -> dict[str, np.ndarray]:
    (work / "deck.peec").write_text(text)
    subprocess.run([str(REF_BIN), "-n", "8", "deck.peec"], cwd=work, check=True, capture_output=True)
    out: dict[str, np.ndarray] = {}
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


def main() -> None:
    if not REF_BIN.exists():
        raise SystemExit("reference build missing: run `make -C oracle ref` first")
    out: dict[str, np.ndarray] = {}
    for name, text in output_decks().items():
        with tempfile.TemporaryDirectory() as tmp:
            for fn, arr in run_reference(text, Path(tmp)).items():
                out[f"out/{name}/{fn}"] = arr
        print(name, sorted(k.split("/")[-1] for k in out if k.startswith(f"out/{name}/")))
    np.savez_compressed(HERE / "output_vectors.npz", **out)
    print(f"wrote {len(out)} arrays to {HERE / 'output_vectors.npz'}")


if __name__ == "__main__":
    main()
