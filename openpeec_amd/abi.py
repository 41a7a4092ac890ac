"""ctypes mirror of the C ABI structs in include/peec.h (byte-identical to OpenPEEC's peec.h).

Used by the Python host layer and by the tests to read peec_t members (cell arrays, matrices,
results) without a compiled binding.  Layout is asserted against the C library at load time
(sizeof(peec_t) and sizeof(seg_t) come from the library itself).
"""
from __future__ import annotations

import ctypes as C

BUFSIZ = 8192
NAMELEN = 32
POLY_MAX = 64


class d_complex_t(C.Structure):
    _fields_ = [("r", C.c_double), ("i", C.c_double)]


class rc_t(C.Structure):
    _fields_ = [("n1", C.c_int), ("n2", C.c_int), ("val", C.c_double)]


class ind_t(C.Structure):
    _fields_ = [("name", C.c_char * NAMELEN), ("n1", C.c_int), ("n2", C.c_int), ("val", C.c_double)]


class mut_t(C.Structure):
    _fields_ = [("name1", C.c_char * NAMELEN), ("name2", C.c_char * NAMELEN),
                ("l1", C.c_int), ("l2", C.c_int), ("k", C.c_double)]


class src_t(C.Structure):
    _fields_ = [("n1", C.c_int), ("n2", C.c_int), ("isvsrc", C.c_int), ("ibr", C.c_int),
                ("amp", C.c_double), ("phase", C.c_double)]


class port_t(C.Structure):
    _fields_ = [("n1", C.c_int), ("n2", C.c_int), ("z0", C.c_double)]


class wire_t(C.Structure):
    _fields_ = [("x1", C.c_double * 3), ("x2", C.c_double * 3), ("shape", C.c_int),
                ("radius", C.c_double), ("width", C.c_double), ("thick", C.c_double),
                ("sigma", C.c_double), ("ndiv", C.c_int)]


class panel_t(C.Structure):
    _fields_ = [("kind", C.c_int), ("v", C.c_double * 12), ("org", C.c_double * 3),
                ("nrm", C.c_double * 3), ("radius", C.c_double), ("ndiva", C.c_int),
                ("ndivb", C.c_int), ("thick", C.c_double), ("sigma", C.c_double),
                ("grading", C.c_int)]


class plate_t(C.Structure):
    _fields_ = [("org", C.c_double * 3), ("ea", C.c_double * 3), ("eb", C.c_double * 3),
                ("thick", C.c_double), ("sigma", C.c_double), ("ndiva", C.c_int),
                ("ndivb", C.c_int), ("ndivt", C.c_int)]


class diel_t(C.Structure):
    _fields_ = [("org", C.c_double * 3), ("ea", C.c_double * 3), ("eb", C.c_double * 3),
                ("thick", C.c_double), ("epsr", C.c_double), ("tand", C.c_double),
                ("epsinf", C.c_double), ("frelax", C.c_double), ("ndiva", C.c_int),
                ("ndivb", C.c_int), ("ndivt", C.c_int)]


class seg_t(C.Structure):
    _fields_ = [("n1", C.c_int), ("n2", C.c_int), ("x1", C.c_double * 3), ("x2", C.c_double * 3),
                ("len", C.c_double), ("shape", C.c_int), ("radius", C.c_double),
                ("width", C.c_double), ("thick", C.c_double), ("wid", C.c_double),
                ("wv", C.c_double * 3), ("vol", C.c_int), ("npv", C.c_int),
                ("papex", C.c_double * 3), ("pv", C.c_double * (3 * POLY_MAX)),
                ("aL", C.c_double), ("aP", C.c_double), ("area", C.c_double),
                ("perim", C.c_double), ("sigma", C.c_double), ("res", C.c_double),
                ("diel", C.c_int), ("cexc", C.c_double), ("gexc", C.c_double),
                ("gfac", C.c_double), ("epss", C.c_double), ("epsinf", C.c_double),
                ("frelax", C.c_double)]


_P = C.POINTER


class peec_t(C.Structure):
    _fields_ = [
        ("title", C.c_char * BUFSIZ),
        ("nres", C.c_int), ("ncap", C.c_int), ("nind", C.c_int), ("nmut", C.c_int),
        ("nsrc", C.c_int), ("nport", C.c_int), ("nwire", C.c_int), ("nplate", C.c_int),
        ("npanel", C.c_int), ("ndiel", C.c_int), ("nnodexyz", C.c_int),
        ("plate", _P(plate_t)), ("panel", _P(panel_t)), ("diel", _P(diel_t)),
        ("res", _P(rc_t)), ("cap", _P(rc_t)), ("ind", _P(ind_t)), ("mut", _P(mut_t)),
        ("src", _P(src_t)), ("port", _P(port_t)), ("wire", _P(wire_t)),
        ("ncid", _P(C.c_int)), ("ncxyz", _P(C.c_double)),
        ("nseg", C.c_int), ("seg", _P(seg_t)), ("lp", _P(d_complex_t)), ("lpwarn", C.c_int),
        ("nchg", C.c_int), ("chg", _P(seg_t)), ("chgnode", _P(C.c_int)), ("ncell", C.c_int),
        ("cellid", _P(C.c_int)), ("cellof", _P(C.c_int)), ("carea", _P(C.c_double)),
        ("csoff", _P(C.c_int)), ("csidx", _P(C.c_int)), ("cpt", _P(C.c_double)),
        ("cmat", _P(d_complex_t)), ("clogged", C.c_int),
        ("ngnode", C.c_int), ("gid", _P(C.c_int)), ("gxyz", _P(C.c_double)),
        ("maxid", C.c_int), ("nnode", C.c_int), ("refnode", C.c_int), ("nodemap", _P(C.c_int)),
        ("offL", C.c_int), ("offV", C.c_int), ("offS", C.c_int), ("offQ", C.c_int),
        ("nunknown", C.c_int),
        ("f0", C.c_double), ("f1", C.c_double), ("nfreq", C.c_int), ("flog", C.c_int),
        ("nodetol", C.c_double), ("gmin", C.c_double), ("skin", C.c_int),
        ("capacitance", C.c_int), ("retardation", C.c_int), ("accel", C.c_int),
        ("dist", C.c_int), ("gp", C.c_int), ("gpz", C.c_double), ("ffnth", C.c_int),
        ("ffnph", C.c_int), ("pw", C.c_int), ("pwth", C.c_double), ("pwph", C.c_double),
        ("pwpol", C.c_int), ("pwamp", C.c_double), ("pwphase", C.c_double), ("tran", C.c_int),
        ("tranatt", C.c_double), ("grading", C.c_int), ("compress", C.c_int),
        ("ctol", C.c_double), ("hlp", C.c_void_p),
        ("zin", _P(d_complex_t)), ("zmat", _P(d_complex_t)), ("smat", _P(d_complex_t)),
        ("segi", _P(d_complex_t)), ("cellq", _P(d_complex_t)), ("voc", _P(d_complex_t)),
        ("segipw", _P(d_complex_t)),
    ]


class mna_sparse_t(C.Structure):
    _fields_ = [("ri", C.POINTER(C.c_int)), ("ci", C.POINTER(C.c_int)), ("val", C.POINTER(d_complex_t)),
                ("nnz", C.c_int), ("cap", C.c_int)]


class hblock_t(C.Structure):
    _fields_ = [("off", C.c_int), ("n", C.c_int), ("h", C.c_void_p), ("scale", d_complex_t)]


class peec_stats_t(C.Structure):
    _fields_ = [("form", C.c_int), ("fill_ms", C.c_double), ("pot_ms", C.c_double),
                ("cinv_ms", C.c_double), ("asm_ms", C.c_double), ("lu_ms", C.c_double),
                ("solve_ms", C.c_double), ("total_ms", C.c_double), ("lu_gemm_ms", C.c_double),
                ("lu_gemm_flop", C.c_double), ("lu_gemm_launches", C.c_int),
                ("fill_flop", C.c_double), ("ising", C.c_int), ("refine_its", C.c_int),
                ("berr", C.c_double), ("n_lu", C.c_int), ("gmres_its", C.c_long),
                ("kf_setup_ms", C.c_double), ("kf_direct_pairs", C.c_long),
                ("mv_ms", C.c_double), ("mv_bytes", C.c_double), ("mv_launches", C.c_long),
                ("fp64_fallbacks", C.c_int), ("pc_builds", C.c_int), ("lpc_fallbacks", C.c_int),
                ("theta_abs_sum", C.c_double)]


class peec_batch_stats_t(C.Structure):
    _fields_ = [("total_ms", C.c_double), ("tune_ms", C.c_double), ("fill_ms", C.c_double),
                ("lu_ms", C.c_double), ("stamp_ms", C.c_double), ("lu_flop", C.c_double),
                ("jobs", C.c_long), ("chunks", C.c_int), ("nb", C.c_int), ("wg_per_cu", C.c_int),
                ("order", C.c_int), ("tune_trials", C.c_int), ("tile_cols", C.c_int),
                ("direct_structures", C.c_int), ("direct_pairs", C.c_long), ("refine_sweeps", C.c_long),
                ("berr_max", C.c_double), ("fp64_fallbacks", C.c_int)]


def declare(lib: C.CDLL, reference: bool = False) -> None:
    """Attach argtypes/restype for the peec.h entry points present in `lib`."""
    PP = C.POINTER(peec_t)
    PS = C.POINTER(seg_t)
    PZ = C.POINTER(d_complex_t)
    sig = {
        "input_data": (C.c_int, [C.c_void_p, PP]),
        "peec_free": (None, [PP]),
        "wire_build": (C.c_int, [PP, C.c_void_p]),
        "lp_fill": (None, [PP, C.c_double, C.c_void_p]),
        "pot_fill": (C.c_int, [PP, C.c_double, C.c_void_p]),
        "pot_cells": (C.c_int, [PP]),
        "lp_entry": (d_complex_t, [PP, C.c_int, C.c_int, C.c_double]),
        "pot_entry": (d_complex_t, [PP, C.c_int, C.c_int, C.c_double]),
        "mna_numbering": (C.c_int, [PP, C.c_void_p]),
        "solve": (C.c_int, [PP, C.c_void_p]),
        "lu_decomp": (C.c_int, [C.c_int, PZ, C.POINTER(C.c_int)]),
        "lu_decomp32": (C.c_int, [C.c_int, PZ, C.POINTER(C.c_int)]),
        "lu_solve": (None, [C.c_int, PZ, C.POINTER(C.c_int), PZ]),
        "hmat_build": (C.c_void_p, [PP, C.c_double, C.c_void_p]),
        "hmat_build_pot": (C.c_void_p, [PP, C.c_double, C.c_void_p]),
        "hmat_build_gen": (C.c_void_p, [C.c_int, C.POINTER(C.c_double), C.c_void_p, C.c_void_p, C.c_double,
                                         C.c_double, C.c_char_p, C.c_void_p]),
        "hmat_free": (None, [C.c_void_p]),
        "hmat_matvec": (None, [C.c_void_p, PZ, PZ]),
        "hmat_nleaf": (C.c_int, [C.c_void_p]),
        "hmat_leaf_block": (PZ, [C.c_void_p, C.c_int, C.POINTER(C.POINTER(C.c_int)), C.POINTER(C.c_int)]),
        "hmat_memory_mb": (C.c_double, [C.c_void_p]),
        "hmat_dense_mb": (C.c_double, [C.c_int]),
        "mna_sparse": (None, [PP, C.c_double, C.POINTER(mna_sparse_t)]),
        "mna_sparse_free": (None, [C.POINTER(mna_sparse_t)]),
        "mna_apply": (None, [PP, C.POINTER(mna_sparse_t), C.POINTER(hblock_t), C.c_int, PZ, PZ, PZ]),
        "pc_build": (C.c_void_p, [PP, C.POINTER(mna_sparse_t), C.POINTER(hblock_t), C.c_int]),
        "pc_free": (None, [C.c_void_p]),
        "pc_apply": (None, [C.c_void_p, PZ]),
        "pc_memory_mb": (C.c_double, [C.c_void_p]),
        "gmres_hmat": (C.c_int, [PP, C.POINTER(mna_sparse_t), C.POINTER(hblock_t), C.c_int, C.c_void_p, PZ,
                                 C.c_double]),
        "neumann_self": (C.c_double, [C.c_double, C.c_double]),
        "neumann_pair": (C.c_double, [PS, PS, C.c_double, C.c_double]),
        "neumann_self_k": (d_complex_t, [PS, C.c_double, C.c_double]),
        "neumann_pair_k": (d_complex_t, [PS, PS, C.c_double, C.c_double, C.c_double]),
        "rect_potential": (C.c_double, [PS, C.POINTER(C.c_double)]),
        "ribbon_static": (C.c_double, [PS, PS, C.c_int, C.c_int]),
        "ribbon_corr": (d_complex_t, [PS, PS, C.c_double, C.c_int]),
        "bar_use_hl": (C.c_int, [PS, PS]),
        "bar_pair": (C.c_double, [PS, PS]),
        "poly_area": (C.c_double, [PS]),
        "poly_potential": (C.c_double, [PS, C.POINTER(C.c_double)]),
        "poly_static": (C.c_double, [PS, PS, C.c_int, C.c_int]),
        "poly_corr": (d_complex_t, [PS, PS, C.c_double]),
        "seg_mirror": (None, [PS, C.c_double, PS]),
        "mna_rhs_port": (None, [PP, C.c_int, PZ]),
        "mna_rhs_sources": (None, [PP, PZ]),
        "mna_rhs_planewave": (None, [PP, C.c_double, PZ]),
        "gmres_solve": (C.c_int, [C.c_int, PZ, PZ, C.POINTER(C.c_int), PZ, C.c_double]),
        "output_zin": (None, [PP, C.c_void_p]),
        "output_csv": (C.c_int, [PP, C.c_char_p]),
        "output_spara": (None, [PP, C.c_void_p]),
        "output_touchstone": (C.c_int, [PP]),
        "output_dist": (C.c_int, [PP, C.c_char_p]),
        "output_pw": (C.c_int, [PP, C.c_char_p, C.c_void_p]),
        "output_far": (C.c_int, [PP, C.c_char_p, C.c_void_p]),
        "output_tran": (C.c_int, [PP, C.c_char_p, C.c_void_p]),
    }
    if not reference:
        sig.update({
            "peec_mi355x_version": (C.c_char_p, []),
            "peec_device_available": (C.c_int, []),
            "peec_load_deck": (C.c_int, [C.c_char_p, PP]),
            "peec_load_deck_string": (C.c_int, [C.c_char_p, PP]),
            "peec_setup": (C.c_int, [PP, C.c_void_p]),
            "peec_sweep": (C.c_int, [PP, C.c_int, C.c_int, C.c_int, PZ, C.POINTER(peec_stats_t)]),
            "peec_sweep_ex": (C.c_int, [PP, C.c_int, C.c_int, C.c_int, C.c_int, PZ, C.POINTER(peec_stats_t)]),
            "peec_sweep_reset": (None, []),
            "peec_sizeof_peec": (C.c_size_t, []),
            "peec_fill_flops": (C.c_double, [PP, C.c_double]),
            "peec_fill_matrices": (C.c_int, [PP, C.c_double, C.c_int, PZ, PZ]),
            "hmat_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "hmat_tree_info": (C.c_int, [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "peec_sizeof_seg": (C.c_size_t, []),
            "peec_batch_create": (C.c_void_p, [C.POINTER(PP), C.c_int, C.c_int, C.c_int]),
            "peec_batch_tune": (C.c_int, [C.c_void_p, C.POINTER(peec_batch_stats_t)]),
            "peec_batch_sweep": (C.c_int, [C.c_void_p, C.POINTER(PZ), C.POINTER(peec_batch_stats_t)]),
            "peec_batch_trials": (C.c_int, [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                            C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int]),
            "peec_batch_free": (None, [C.c_void_p]),
            "peec_batch_set_form": (C.c_int, [C.c_void_p, C.c_int]),
            "peec_lpc_nparams": (C.c_int, []),
            "peec_lpc_nfeatures": (C.c_int, []),
            "peec_lpc_set_model": (None, [C.POINTER(C.c_double)]),
            "peec_lpc_get_model": (C.c_int, [C.POINTER(C.c_double)]),
            "peec_lpc_samples": (C.c_int, [PP, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double),
                                           C.POINTER(C.c_double), C.POINTER(C.c_int)]),
            "peec_lpc_mlp_forward": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double)]),
            "peec_lpc_mlp_grad": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int,
                                            C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "peec_lpc_fit": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int, C.c_int, C.c_double,
                                       C.c_ulonglong, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
            "peec_batch_lu_decomp": (C.c_int, [C.c_int, C.POINTER(C.c_int), PZ, C.POINTER(C.c_int), C.c_int]),
            "lu_decomp_device_noalloc": (C.c_int, [C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
        })
    for name, (res, args) in sig.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
