"""ctypes binding of ``libbraincell_amd.so`` (the C ABI in ``include/braincell_amd.h``).

The library is built in-tree by ``braincell_amd.build``.  Loading fails loudly
when it is missing: there is no CPU fallback for any kernel.
"""

from __future__ import annotations

import ctypes
import os

from .build import library_path

__all__ = ["lib", "check", "BcPack", "StaggerDesc", "AdjointDesc", "chan_const_type", "NativeError"]

BC_OK = 0
STATUS_NAMES = {1: "INVALID_ARG", 2: "INVALID_LAYOUT", 3: "HIP_ERROR", 4: "NONFINITE", 5: "UNSUPPORTED"}

class NativeError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"braincell_amd native error {STATUS_NAMES.get(code, code)}: {message}")
        self.code = code


class BcPack(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "cv0", "n_cv", "row0", "n_rows", "ch0", "n_ch", "lv0", "n_lv", "bnd0", "n_bnd",
        "cls0", "n_cls", "clamp0", "n_clamp", "probe0", "n_probe")]


class StaggerDesc(ctypes.Structure):
    _fields_ = [
        ("precision", ctypes.c_int32), ("channels", ctypes.c_uint32), ("n_packs", ctypes.c_int32),
        ("threads", ctypes.c_int32), ("cv_per_thread", ctypes.c_int32),
        ("max_rows", ctypes.c_int32), ("max_chains", ctypes.c_int32), ("max_cls", ctypes.c_int32),
        ("max_levels", ctypes.c_int32),
        ("n_cv_total", ctypes.c_int64),
        ("packs", ctypes.c_void_p), ("cv_info", ctypes.c_void_p), ("bnd_info", ctypes.c_void_p),
        ("chain_off", ctypes.c_void_p), ("chain_top", ctypes.c_void_p), ("kid_off", ctypes.c_void_p),
        ("kids", ctypes.c_void_p), ("level_off", ctypes.c_void_p),
        ("lo", ctypes.c_void_p), ("up", ctypes.c_void_p),
        ("cls_tab", ctypes.c_void_p), ("clamp_row", ctypes.c_void_p), ("clamp_on", ctypes.c_void_p),
        ("clamp_off", ctypes.c_void_p), ("clamp_amp", ctypes.c_void_p), ("probe_row", ctypes.c_void_p),
        ("probe_col", ctypes.c_void_p), ("chan_const", ctypes.c_void_p), ("V", ctypes.c_void_p), ("Vbnd", ctypes.c_void_p),
        ("state", ctypes.c_void_p), ("trace", ctypes.c_void_p), ("trace_ld", ctypes.c_int64),
        ("rec", ctypes.c_void_p), ("status", ctypes.c_void_p), ("step0", ctypes.c_int32), ("n_steps", ctypes.c_int32),
        ("dt", ctypes.c_double), ("flags", ctypes.c_uint32), ("syn", ctypes.c_void_p),
        ("chain_rec", ctypes.c_void_p), ("chain_par", ctypes.c_void_p), ("kid_pad", ctypes.c_void_p),
        ("slot_link", ctypes.c_void_p), ("slot_chain", ctypes.c_void_p), ("spk", ctypes.c_void_p),
        ("chain_work", ctypes.c_void_p), ("cv_aux", ctypes.c_void_p), ("bnd_aux", ctypes.c_void_p),
    ]


class SpikeDesc(ctypes.Structure):
    _fields_ = [("cap", ctypes.c_int32), ("v_th", ctypes.c_double), ("v_last", ctypes.c_void_p),
                ("count", ctypes.c_void_p), ("steps", ctypes.c_void_p)]


class SynapseDesc(ctypes.Structure):
    _fields_ = [
        ("n_cells", ctypes.c_int32), ("spk_len", ctypes.c_int32), ("min_delay", ctypes.c_int32),
        ("max_delay", ctypes.c_int32), ("v_th", ctypes.c_double),
        ("pack_off", ctypes.c_void_p), ("row", ctypes.c_void_p), ("row_off", ctypes.c_void_p),
        ("par", ctypes.c_void_p), ("state", ctypes.c_void_p), ("con_off", ctypes.c_void_p),
        ("con_pre", ctypes.c_void_p), ("con_delay", ctypes.c_void_p), ("con_w", ctypes.c_void_p),
        ("cv_cell", ctypes.c_void_p), ("spk", ctypes.c_void_p),
        ("n_ext", ctypes.c_int32), ("ext_step0", ctypes.c_int32), ("ext_steps", ctypes.c_int32),
        ("ext_slot", ctypes.c_void_p), ("ext_drive", ctypes.c_void_p), ("n_syn", ctypes.c_int32),
        ("syn_trace", ctypes.c_void_p), ("drive", ctypes.c_void_p), ("max_syn", ctypes.c_int32),
        ("max_syn_rows", ctypes.c_int32),
    ]


class CableDesc(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("n_cells", ctypes.c_int32), ("threads", ctypes.c_int32),
                ("max_cv", ctypes.c_int32)] + [(n, ctypes.c_void_p) for n in (
                    "cell_cv0", "k_off", "k_col", "k_val", "a", "b", "jac", "clamp_off", "clamp_cv", "clamp_on",
                    "clamp_end", "clamp_amp", "probe_off", "probe_cv", "probe_col")] + [
                ("trace_ld", ctypes.c_int32), ("trace", ctypes.c_void_p), ("V", ctypes.c_void_p),
                ("status", ctypes.c_void_p), ("step0", ctypes.c_int32), ("n_steps", ctypes.c_int32),
                ("dt", ctypes.c_double), ("max_nnz", ctypes.c_int32)]


class AdjointDesc(ctypes.Structure):
    _fields_ = [("layout", ctypes.c_void_p), ("n_steps", ctypes.c_int32), ("rec", ctypes.c_void_p),
                ("trace_bar", ctypes.c_void_p), ("Vbar", ctypes.c_void_p), ("Sbar", ctypes.c_void_p),
                ("cls_grad", ctypes.c_void_p)]


CHAN_CONST_FIELDS = (
    "E_Na", "E_K", "E_IL", "na_vsh", "na_phi", "kdr_vsh", "kdr_phi", "cal_vsh", "cal_phi_p", "cal_phi_q",
    "cat_vsh", "cat_phi_p", "cat_phi_q", "ahp_alpha", "ahp_beta", "ahp_n", "ahp_phi", "hcn_E", "hcn_phi",
    "ca_drive", "ca_tau_inv", "ca_rest", "ca_Co", "ca_rtzf", "E_Ca_fixed",
)


def chan_const_type(precision: int):
    ct = ctypes.c_float if precision == 32 else ctypes.c_double
    return type(f"ChanConst{precision}", (ctypes.Structure,), {"_fields_": [(n, ct) for n in CHAN_CONST_FIELDS]})


_LIB = None


def lib():
    """Load (once) and return the native library; raises when it is not built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"braincell_amd native library not built: {path} is missing "
                           "(run `python -m braincell_amd.build`)")
    L = ctypes.CDLL(path)
    L.bc_version.restype = ctypes.c_int
    L.bc_last_error.restype = ctypes.c_char_p
    L.bc_stagger_supported.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_int]
    L.bc_stagger_supported.restype = ctypes.c_int
    L.bc_stagger_run.argtypes = [ctypes.POINTER(StaggerDesc), ctypes.c_void_p]
    L.bc_stagger_run.restype = ctypes.c_int
    L.bc_stagger_graph_create.argtypes = [ctypes.POINTER(StaggerDesc), ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    L.bc_stagger_graph_create.restype = ctypes.c_int
    L.bc_graph_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.bc_graph_launch.restype = ctypes.c_int
    L.bc_graph_destroy.argtypes = [ctypes.c_void_p]
    L.bc_graph_destroy.restype = ctypes.c_int
    L.bc_adjoint_supported.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_int]
    L.bc_adjoint_supported.restype = ctypes.c_int
    L.bc_stagger_adjoint.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.bc_stagger_adjoint.restype = ctypes.c_int
    L.bc_point_supported.argtypes = [ctypes.c_int]
    L.bc_point_supported.restype = ctypes.c_int
    L.bc_point_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.bc_point_run.restype = ctypes.c_int
    L.bc_cable_be_run.argtypes = [ctypes.POINTER(CableDesc), ctypes.c_void_p]
    L.bc_cable_be_run.restype = ctypes.c_int
    _LIB = L
    return L


def check(rc: int):
    if rc != BC_OK:
        msg = lib().bc_last_error()
        raise NativeError(rc, msg.decode() if msg else "")
