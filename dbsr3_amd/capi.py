"""ctypes bindings of libdbsr3_amd.so (the C-ABI declared in include/dbsr3_amd.h).

The library is built in-tree by ``dbsr3_amd.build``.  Loading fails loudly
when it is missing: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdbsr3_amd.so")

vp = C.c_void_p
dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
i32 = C.c_int
f64 = C.c_double
ch = C.c_char
sz = C.c_size_t

GEN_MATRIX_FN = C.CFUNCTYPE(None, vp, i32, i32, i32, dp, ip, ip, ip, ip)
KBORT_FN = C.CFUNCTYPE(i32, vp, i32, i32)
QUADR_PQ_FN = C.CFUNCTYPE(f64, vp, i32, i32, i32)
IBORT_FN = C.CFUNCTYPE(i32, vp, i32, i32)
i64 = C.c_longlong
lp = C.POINTER(C.c_longlong)


class MatBlocksC(C.Structure):
    """struct dbs_mat_blocks (MAT3 matrix storage descriptor)"""
    _fields_ = [("nch", i32), ("npert", i32), ("mk", i32),
                ("icc", ip), ("nblocks", i32), ("hch_dev", vp), ("acf_dev", vp),
                ("icb", ip), ("ncb", i32), ("hcp_dev", vp),
                ("ibb", ip), ("nbb", i32), ("hp_dev", vp)]


mbp = C.POINTER(MatBlocksC)


class Mat3AuxC(C.Structure):
    """struct dbs_mat3_aux (bound-orbital and target data of a MAT3 run)"""
    _fields_ = [("ibo", i32), ("ipbs", ip), ("kbort", KBORT_FN), ("kbort_user", vp),
                ("vch_dev", vp), ("nvch", i32), ("fppqq_dev", vp), ("iitar", i32),
                ("iptar", ip), ("htarg", dp), ("quadr_pq", QUADR_PQ_FN), ("quadr_pq_user", vp)]


auxp = C.POINTER(Mat3AuxC)


class LDataAuxC(C.Structure):
    """struct dbs_ldata_aux (L_data's one-electron inputs)"""
    _fields_ = [("mat3", auxp), ("otarg", dp), ("check_target", i32), ("norb", i32), ("lval", dp),
                ("kbs", ip), ("nkap", i32), ("kap_list", ip), ("hd_dev", vp)]


ldp = C.POINTER(LDataAuxC)


class NucleusC(C.Structure):
    """dbs_nucleus: model 0 point, 1 uniform, 2 Fermi; z; r_uniform, a_fermi, c_fermi (a.u.)"""
    _fields_ = [("model", i32), ("z", f64), ("r_uniform", f64), ("a_fermi", f64), ("c_fermi", f64)]


nucp = C.POINTER(NucleusC)


class QpqTableC(C.Structure):
    """struct dbs_qpq_table (tabulated QUADR_PQ)"""
    _fields_ = [("norb", i32), ("mmax", i32), ("q", dp)]


class Mat3OrbsC(C.Structure):
    """struct dbs_mat3_orbs (orbital tables of MAT3's Add_integral)"""
    _fields_ = [("norb", i32), ("lbs", ip), ("kbs", ip), ("ipbs", ip), ("ibo", i32), ("mk", i32),
                ("eps_c", f64), ("check_target", i32)]

# name -> argtypes (all return int status unless listed in _RESTYPE)
PROTOTYPES = {
    "dbs_version": [],
    "dbs_ctx_create": [C.POINTER(vp), i32, i32, i32, i32, dp, i32],
    "dbs_ctx_destroy": [vp],
    "dbs_ctx_dims": [vp, ip],
    "dbs_ctx_sync": [vp],
    "dbs_ctx_tables": [vp, dp, dp, dp, dp],
    "dbs_ctx_set_cache_limit": [vp, sz],
    "dbs_ctx_set_workspace_limit": [vp, sz],
    "dbs_malloc": [vp, C.POINTER(vp), sz],
    "dbs_free": [vp, vp],
    "dbs_memcpy_h2d": [vp, vp, vp, sz],
    "dbs_memcpy_d2h": [vp, vp, vp, sz],
    "dbs_moments_dev": [vp, i32, i32, vp],
    "dbs_diag_cells_dev": [vp, i32, i32, vp],
    "dbs_mrk": [vp, i32, i32, C.POINTER(vp)],
    "dbs_mrk_host": [vp, i32, i32, dp],
    "dbs_density_dev": [vp, vp, vp, ch, vp],
    "dbs_convol_dev": [vp, i32, i32, i32, ch, ch, vp, vp],
    "dbs_sum_amb_dev": [vp, vp, vp, ch, vp],
    "dbs_rk_dev": [vp, vp, vp, vp, vp, i32, vp],
    "dbs_rk_noalloc_dev": [vp, vp, vp, vp, vp, i32, vp],
    "dbs_quadr_dev": [vp, vp, vp, i32, vp],
    "dbs_update_hs_dev": [vp, vp, i32, i32, vp, ch],
    "dbs_density": [vp, dp, dp, ch, dp],
    "dbs_convol": [vp, i32, i32, i32, ch, ch, dp, dp],
    "dbs_sum_amb": [vp, dp, dp, ch, dp],
    "dbs_rk": [vp, dp, dp, dp, dp, i32, dp],
    "dbs_quadr": [vp, dp, dp, i32, dp],
    "dbs_cdata_create": [C.POINTER(vp), i32, i32, i32, i32, i32, i32, f64, GEN_MATRIX_FN, vp],
    "dbs_cdata_destroy": [vp],
    "dbs_cdata_add_coef": [vp, f64, i32, i32, i32, i32, i32, i32],
    "dbs_cdata_add_matrix": [vp, i32, i32],
    "dbs_add_integral_ci": [vp, i32, i32, i32, i32, i32, i32, i32, f64, ip, i32, i32],
    "dbs_smu": [i32, i32, i32, i32, i32, i32, dp, dp],
    "dbs_add_integral_mat": [vp, i32, i32, i32, i32, i32, f64, ip, ip, i32, i32, f64],
    "dbs_add_integral_mat3": [vp, C.POINTER(Mat3OrbsC), i32, i32, i32, i32, i32, i32, f64, i32, i32, i32, i32],
    "dbs_decode_int": [i32, ip, ip, ip, ip, ip, ip, ip],
    "dbs_rk_evaluate_dev": [vp, i32, i32, i32, dp, ip, ip, ip, ip, f64, i32, vp, i32, vp, i32,
                            i32, vp],
    "dbs_rdata_34_dev": [vp, i32, i32, i32, i32, dp, ip, ip, ip, i32, vp, i32, ip, i32, vp, i32],
    "dbs_bspline_overlaps_dev": [vp, vp, vp, vp],
    "dbs_vnucl_dev": [vp, nucp, vp],
    "dbs_dhl_pq_dev": [vp, nucp, i32, i32, f64, f64, vp],
    "dbs_full_mat_sym_dev": [vp, vp, ch, vp],
    "dbs_update_hx_dev": [vp, mbp, i32, i32, vp, f64],
    "dbs_update_hw_dev": [vp, mbp, i32, i32, vp, vp, f64],
    "dbs_update_hv_dev": [vp, mbp, i32, i32, vp, f64],
    "dbs_update_hb_dev": [vp, mbp, i32, ip, ip, dp],
    "dbs_update_acf_dev": [vp, mbp, i32, ip, ip, ip, dp],
    "dbs_o_data_dev": [vp, mbp, i32, i32, ip, ip, dp, f64, i32, ip, KBORT_FN, vp, vp, i32, vp, i32,
                       ip, dp],
    "dbs_overlap_ci_dev": [vp, i32, ip, ip, dp, f64, i32, vp, i32],
    "dbs_ldata_dev": [vp, mbp, ldp, i32, i32, ip, ip, ip, ip, dp, f64, vp],
    "dbs_rdata_dev": [vp, mbp, auxp, i32, i32, i32, i32, dp, ip, ip, ip, ip, vp, i32],
    "dbs_sdata_dev": [vp, mbp, auxp, i32, i32, i32, dp, ip, ip, ip, ip, vp, i32],
    "dbs_bav_dev": [vp, vp, vp, ch, ch, vp],
    "dbs_hf_matrix_dev": [vp, i32, i32, vp, dp, dp, i32, f64, i32, i32, ip, ip, ip, ip, dp, vp],
    "dbs_mchf_matrix_dev": [vp, i32, i32, vp, vp, i32, ip, ip, dp, i32, ip, ip, ip, ip, ip, dp, i32, f64, f64,
                            i32, i32, vp, vp, vp, vp, dp, dp],
    "dbs_core_matrix_dev": [vp, i32, i32, vp, vp, dp, dp, i32, dp, i32, vp, vp],
    "dbs_hf_eiv_dev": [vp, vp, i32, vp, ip, i32, vp, vp, dp],
    "dbs_zero_cond": [i32, i32, i32, i32, i32, i32, ip, ip, i32, i32, i32, i32, i32, ip, ip],
    "dbs_orth_conditions": [i32, ip, i32, ip, ip, IBORT_FN, vp, i32, ip, ip],
    "dbs_orth_dev": [vp, mbp, i32, ip, ip, i32, vp, vp, i32],
    "dbs_diag_channels_dev": [vp, mbp, vp, i32, ip, f64, f64, f64, ip],
    "dbs_record_matrix_dev": [vp, mbp, vp, i32, ip, vp, i64, lp, vp, i64, lp, dp],
    "dbs_engine_create": [C.POINTER(vp), i32, C.POINTER(vp), ip, ip, ip, i32],
    "dbs_engine_destroy": [vp],
    "dbs_engine_buffers": [vp, i32, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp),
                           C.POINTER(vp)],
    "dbs_engine_forward": [vp],
    "dbs_engine_forward_adjoint": [vp],
    "dbs_engine_forward_plain": [vp],
    "dbs_engine_graph_capture": [vp, i32],
    "dbs_engine_graph_launch": [vp],
    "dbs_engine_set_tuning": [vp, i32, ip],
    "dbs_engine_set_memory_budget": [vp, sz],
    "dbs_engine_memory": [vp, ip, C.POINTER(sz)],
    "dbs_engine_get_tuning": [vp, i32, ip],
    "dbs_engine_autotune": [vp, i32, i32],
    "dbs_engine_work": [vp, i32, dp, dp],
    "dbs_engine_sync": [vp],
    "dbs_engine_profile": [vp, i32, i32, i32, C.c_char_p, dp, dp, dp, ip],
}

_lib = None


class DBSError(RuntimeError):
    pass


def lib():
    """Load the HIP library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DBSError(
                f"libdbsr3_amd.so not found at {LIB_PATH}; build it with "
                "`python -m dbsr3_amd.build` (hipcc, gfx950)")
        L = C.CDLL(LIB_PATH)
        L.dbs_last_error.restype = C.c_char_p
        L.dbs_last_error.argtypes = []
        L.dbs_quadr_pq_table.restype = C.c_double
        L.dbs_quadr_pq_table.argtypes = [vp, i32, i32, i32]
        for name, args in PROTOTYPES.items():
            f = getattr(L, name, None)
            if f is None:  # reported by tests/test_capi.py
                continue
            f.restype = C.c_int
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().dbs_last_error().decode(errors="replace")
        raise DBSError(f"status {rc}: {msg}")
    return rc


def exported_symbols():
    """Every function name the public header declares."""
    return ["dbs_last_error", "dbs_quadr_pq_table"] + list(PROTOTYPES)
