"""The C-ABI library: it loads without a GPU, exports every function include/*.h declares, and its
structs have the reference layout.  No compute call is made here."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from openpeec_amd import abi, load_library

ROOT = Path(__file__).resolve().parents[1]


def declared_functions() -> set[str]:
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//.*", "", text)
        text = re.sub(r"static inline[^{]*\{.*?\n\}", "", text, flags=re.S)   # inline helpers (complex.h, freq_at)
        text = re.sub(r"^typedef[^;]*;", "", text, flags=re.M)                  # function-pointer typedefs
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?[\s\*](\w+)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name in ("if", "for", "while", "return", "sizeof"):
                continue
            names.add(name)
    return names


def exported(lib_path: Path) -> set[str]:
    out = subprocess.run(["nm", "-D", "--defined-only", str(lib_path)], capture_output=True, text=True, check=True)
    return {line.split()[-1] for line in out.stdout.splitlines() if line.strip()}


def test_library_loads_and_reports(lib):
    assert b"gfx950" in lib.peec_mi355x_version()


def test_every_declared_function_is_exported(lib):
    from openpeec_amd import LIB_PATH
    decl = declared_functions()
    assert {"lp_fill", "pot_fill", "lu_decomp", "lu_solve", "solve", "peec_sweep"} <= decl
    missing = sorted(decl - exported(LIB_PATH))
    assert not missing, f"declared in include/*.h but not exported: {missing}"


def test_struct_sizes_match_library(lib):
    assert lib.peec_sizeof_peec() == C.sizeof(abi.peec_t) == 8744
    assert lib.peec_sizeof_seg() == C.sizeof(abi.seg_t) == 1800


@pytest.mark.skipif(not Path("/root/reference/include/peec.h").exists(), reason="reference sources not present")
def test_layout_identical_to_reference(tmp_path):
    """sizeof/offsetof listing of include/peec.h equals the reference header's (same probe source)."""
    probe = ROOT / "oracle" / "layout_probe.c"
    outs = []
    for inc in (ROOT / "include", Path("/root/reference/include")):
        exe = tmp_path / f"probe_{len(outs)}"
        subprocess.run(["gcc", "-std=c11", f"-I{inc}", str(probe), "-o", str(exe), "-lm"], check=True)
        outs.append(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout)
    assert outs[0] == outs[1]


def _c_struct_fields(header: str, name: str) -> list[str]:
    """Member names, in order, of `typedef struct { ... } name;` in an include/ header."""
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    m = re.search(r"typedef struct\s*\{((?:(?!typedef struct).)*?)\}\s*" + name + r"\s*;", text, flags=re.S)
    assert m, name
    out = []
    for decl in m.group(1).split(";"):
        decl = decl.strip()
        if not decl:
            continue
        parts = decl.split(None, 1)
        out += [v.strip().lstrip("*") for v in parts[1].split(",")]
    return out


@pytest.mark.parametrize("header,name,cls", [("peec_mi355x.h", "peec_stats_t", abi.peec_stats_t),
                                             ("peec_mi355x.h", "peec_batch_stats_t", abi.peec_batch_stats_t)])
def test_stats_struct_fields_follow_header(header, name, cls):
    """The ctypes mirrors of the statistics structs list the header's members in its order."""
    assert [f[0] for f in cls._fields_] == _c_struct_fields(header, name)
