"""Deck reader, geometry builder, capacitive cells and MNA numbering (host code) against the
reference: the cell arrays must be bit-identical (every partial element is a function of them).

Checked against the committed sha256 of the reference's seg[]/chg[] arrays for every sample deck
(tests/golden/ref_vectors.npz, written by tests/golden/make_golden.py), and live against the
reference build when it is present."""
import ctypes as C
import hashlib

import numpy as np
import pytest

from conftest import DECKS, deck_text
from openpeec_amd import Peec, abi, decks
from refharness import RefPeec, ref_available

GOLD = np.load(DECKS.parent / "ref_vectors.npz")
GEOM = sorted(k.split("/", 1)[1] for k in GOLD.files if k.startswith("geom/"))


def seg_hash(p, which):
    n = p.nseg if which == "seg" else p.nchg
    raw = C.string_at(getattr(p, which), n * C.sizeof(abi.seg_t)) if n else b""
    return hashlib.sha256(raw).hexdigest()


def without_skin(text):
    """The deck with skineffect switched off: the library refuses skin effect (outside this path),
    and the cells do not depend on it."""
    return "\n".join(ln for ln in text.splitlines() if not ln.strip().startswith("skineffect"))


@pytest.mark.parametrize("name", GEOM)
def test_cells_bitwise_equal_reference_hash(lib, name):
    m = Peec.from_deck_text(without_skin(deck_text(name + ".peec")), lib)
    m.wire_build()
    want = GOLD[f"geom/{name}"]
    assert seg_hash(m.p, "seg") == want[0]
    assert seg_hash(m.p, "chg") == want[1]
    m.close()


@pytest.mark.skipif(not ref_available(), reason="reference build not present")
@pytest.mark.parametrize("text_fn", [lambda: decks.plate_deck(64), lambda: decks.dipole_deck(),
                                     lambda: decks.random_wire_deck(7), lambda: decks.plate_deck(8, ndivt=3)])
def test_generated_decks_match_reference_live(lib, text_fn):
    text = text_fn()
    r = RefPeec(text)
    r.wire_build()
    m = Peec.from_deck_text(text, lib)
    m.wire_build()
    assert seg_hash(m.p, "seg") == seg_hash(r.p, "seg")
    assert seg_hash(m.p, "chg") == seg_hash(r.p, "chg")
    assert [m.p.chgnode[i] for i in range(m.p.nchg)] == [r.p.chgnode[i] for i in range(r.p.nchg)]
    # capacitive cells and numbering
    assert lib.pot_cells(C.byref(m.p)) == 0 and r.lib.pot_cells(C.byref(r.p)) == 0
    assert m.p.ncell == r.p.ncell
    for fld, n in (("cellof", m.p.nchg), ("csoff", m.p.ncell + 1), ("csidx", m.p.nchg)):
        assert [getattr(m.p, fld)[i] for i in range(n)] == [getattr(r.p, fld)[i] for i in range(n)]
    assert [m.p.carea[i] for i in range(m.p.ncell)] == [r.p.carea[i] for i in range(r.p.ncell)]
    assert lib.mna_numbering(C.byref(m.p), None) == 0
    assert r.lib.mna_numbering(C.byref(r.p), r.log) == 0
    assert (m.p.nunknown, m.p.offS, m.p.refnode) == (r.p.nunknown, r.p.offS, r.p.refnode)
    m.close()
    r.close()


def test_workload_sizes(lib):
    """The precision workload: 8320 branches, 4225 charge cells, 12545 unknowns."""
    m = Peec.from_deck_text(decks.plate_deck(64), lib)
    m.wire_build()
    assert lib.pot_cells(C.byref(m.p)) == 0
    assert lib.mna_numbering(C.byref(m.p), None) == 0
    assert (m.p.nseg, m.p.ncell, m.p.nunknown) == (8320, 4225, 12545)
    m.close()


def test_deck_errors_match_reference_rules(lib, capfd):
    bad = "OpenPEEC 1 0\nwire = 0 0 0 0 0 1 1e-3 5.8e7 4\nfrequency = 1e6 1e6 0\nend\n"
    with pytest.raises(ValueError):
        Peec.from_deck_text(bad, lib)
    C.CDLL(None).fflush(None)   # the library's stdout is a buffered C stream
    assert "no port data" in capfd.readouterr().out
    # beyond the reference: a 96-sector disk (its centre ring split into 4 sector cells) is read
    disk96 = "OpenPEEC 1 0\ndisk = 0 0 0 0 0 1 0.05 3.5e-5 5.8e7 4 96\nport = 1 2 50\nfrequency = 1e6 1e6 0\nend\n"
    m = Peec.from_deck_text(disk96, lib)
    m.wire_build()
    assert m.p.nchg == 4 + 4 * 96 and max(m.p.chg[i].npv for i in range(m.p.nchg)) <= 64
    m.close()


def _c_stdout(capfd):
    """what the library printed (C stdio is flushed first)"""
    from refharness import _libc
    _libc.fflush(None)
    return capfd.readouterr().out


def test_skin_effect_is_refused(lib, capfd):
    with pytest.raises(ValueError):
        Peec.from_deck_text(deck_text("dipole_full.peec"), lib)
    assert "skineffect = 1 is not available" in _c_stdout(capfd)


@pytest.mark.parametrize("bad,msg", [
    ("resistor = 1 2 -5", "*** invalid resistor data"),
    ("wire = 0 0 0 0 0 0 1e-3 5.8e7 4", "*** invalid wire data"),
    ("plate = 0 0 0 1 0 0 0.5 0.5 0 1e-4 5.8e7 4 4", "*** plate : the two edge vectors must be perpendicular"),
    ("quad = 0 0 0 1 0 0 1 1 0 0 1 0.5 1e-4 5.8e7 2 2", "*** quad : vertices must be coplanar"),
    ("quad = 0 0 0 1 0 0 0 1 0 1 1 0 1e-4 5.8e7 2 2", "*** invalid quad data"),
    ("quad = 0 0 0 2 0 0 1 0.5 0 1 2 0 1e-4 5.8e7 2 2", "*** quad : must be convex (vertices in cyclic order)"),
    ("disk = 0 0 0 0 0 1 0.05 1e-4 5.8e7 4 2", "*** invalid disk data"),
    ("mutual = LA LB 0.5", "*** mutual : unknown or same inductor : LA LB"),
    ("port = 1 1 50", "*** invalid port data"),
    ("planewave = 90 0 3 1.0", "*** invalid planewave data"),
    ("node = -1 0 0 0", "*** invalid node data"),
])
def test_deck_errors_match_reference_messages(lib, capfd, bad, msg):
    """Rejected decks print the reference's message (input_data.c) and return 1; each case's
    message was checked against the reference build's input_data on the same text."""
    text = "OpenPEEC 1 0\nresistor = 1 0 50\n" + bad + "\nport = 1 0 50\nfrequency = 1e6 1e7 1\nend\n"
    with pytest.raises(ValueError):
        Peec.from_deck_text(text, lib)
    assert msg in _c_stdout(capfd)


def test_deck_reader_line_rules(lib):
    """Comments, blank lines, CR-LF endings, malformed lines and unknown keys are passed over;
    lines after \"end\" are not read; optional fields take their defaults."""
    text = ("# comment\r\nOpenPEEC 1 0\r\n\r\nnot a statement\r\nfuture_key = 3\r\n"
            "title =   two resistors \r\nresistor = 1 2 50\r\nresistor = 2 0 25\r\nvsource = 1 0 2.0\r\n"
            "port = 1 0 50\r\nfrequency = 1e6 1e7 4 log\r\nend\r\nresistor = 7 8 9\r\n")
    m = Peec.from_deck_text(text, lib)
    assert m.p.title.decode() == "two resistors "
    assert (m.p.nres, m.p.nsrc, m.p.nport, m.p.nfreq, m.p.flog, m.p.maxid) == (2, 1, 1, 5, 1, 2)
    assert m.p.src[0].phase == 0 and m.p.src[0].isvsrc == 1 and m.p.src[0].ibr == -1
    assert (m.p.nodetol, m.p.tranatt, m.p.ctol) == (1e-8, 40, 1e-6)
    m.close()


def _netlist(p):
    """every field input_data fills, as bytes: scalars of peec_t up to the cell arrays and the
    element arrays"""
    arrays = [("res", p.nres, abi.rc_t), ("cap", p.ncap, abi.rc_t), ("ind", p.nind, abi.ind_t), ("mut", p.nmut, abi.mut_t),
              ("src", p.nsrc, abi.src_t), ("port", p.nport, abi.port_t), ("wire", p.nwire, abi.wire_t),
              ("plate", p.nplate, abi.plate_t), ("panel", p.npanel, abi.panel_t), ("diel", p.ndiel, abi.diel_t),
              ("ncid", p.nnodexyz, C.c_int), ("ncxyz", 3 * p.nnodexyz, C.c_double)]
    out = {name: C.string_at(getattr(p, name), n * C.sizeof(t)) if n else b"" for name, n, t in arrays}
    skip = {name for name, _, _ in arrays} | {"seg", "chg", "lp", "chgnode", "cellid", "cellof", "carea", "csoff", "csidx",
                                               "cpt", "cmat", "gid", "gxyz", "nodemap", "hlp", "zin", "zmat", "smat", "segi",
                                               "cellq", "voc", "segipw"}
    for name, _ in abi.peec_t._fields_:
        if name not in skip:
            v = getattr(p, name)
            out[name] = bytes(v) if isinstance(v, (bytes, C.Array)) else v
    return out


@pytest.mark.skipif(not ref_available(), reason="reference build not present")
@pytest.mark.parametrize("deck", sorted(p.name for p in DECKS.glob("*.peec")))
def test_netlist_equals_reference_live(lib, deck):
    """The parsed netlist of every sample deck, field by field, against the reference's
    input_data on the same text (skin effect switched off: refused here)."""
    text = without_skin(deck_text(deck))
    mine = Peec.from_deck_text(text, lib)
    ref = RefPeec(text)
    a, b = _netlist(mine.p), _netlist(ref.p)
    assert a.keys() == b.keys()
    for k in a:
        assert a[k] == b[k], (deck, k)
    mine.close()
    ref.close()
