"""The drop-in boundary at link level: the reference's own caller objects (main.c, input_data.c,
wire.c, mna.c, output.c, farfield.c, transient.c, skin.c, utils.c), compiled from its sources
with its own header, link against libpeec_mi355x.so with no unresolved symbol -- the library
provides every hot-path function they call (lp_fill, pot_fill, solve, lu_decomp, hmat_*, ...).
Link only (the program needs a GPU to run).  Needs the reference sources; skipped without them."""
import subprocess
import tempfile
from pathlib import Path

import pytest

REF = Path("/root/reference")
LIB = Path(__file__).resolve().parents[1] / "openpeec_amd"
CALLERS = ["main", "input_data", "wire", "mna", "output", "farfield", "transient", "skin", "utils"]


@pytest.mark.skipif(not (REF / "src" / "main.c").exists(), reason="reference sources not present")
def test_reference_callers_link_against_the_library():
    if not (LIB / "libpeec_mi355x.so").exists():
        pytest.skip("library not built")
    with tempfile.TemporaryDirectory() as tmp:
        objs = []
        for c in CALLERS:
            o = Path(tmp) / f"{c}.o"
            subprocess.run(["gcc", "-std=c11", "-O1", "-fopenmp", f"-I{REF / 'include'}", "-c", str(REF / "src" / f"{c}.c"),
                            "-o", str(o)], check=True, capture_output=True)
            objs.append(str(o))
        exe = Path(tmp) / "peec_gpu"
        r = subprocess.run(["gcc", "-fopenmp", *objs, f"-L{LIB}", f"-Wl,-rpath,{LIB}", "-Wl,--no-undefined", "-lpeec_mi355x",
                            "-lm", "-o", str(exe)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        assert exe.exists()
