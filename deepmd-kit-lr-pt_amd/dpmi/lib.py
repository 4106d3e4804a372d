"""ctypes binding of libdpmi.so (include/dpmi.h).

The library is loaded from this directory only (built in-tree by
``python -m dpmi._build``); a missing library raises instead of falling back
to any CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libdpmi.so")

_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_D = C.c_double
_LL = C.c_longlong

STATUS = {0: "OK", 1: "ARGUMENT", 2: "HIP", 3: "NLIST_CAPACITY", 4: "MEMORY", 5: "RANGE", 6: "UNSUPPORTED"}

DPMI_FP64 = 0
DPMI_MIXED = 1

# every entry point of include/dpmi.h with its argument types
SIGNATURES = {
    "dpmi_last_error": ([], C.c_char_p),
    "dpmi_version": ([], C.c_char_p),
    "dpmi_normalize_coord_f64": ([_P, _I, _P, _P], _I),
    "dpmi_copy_coord_f64": ([_P, _P, _P, _P, _P, _P, _I, _I, _F, _P, _P], _I),
    "dpmi_build_nlist_f64": ([_P, _P, _P, _P, _P, _I, _I, _I, _F, _I, _P, _P], _I),
    "dpmi_prod_env_mat_a_f64": ([_P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _I, _P, _P, _I, _I, _I, _F, _F, _P, _I,
                                 _P], _I),
    "dpmi_tabulate_fusion_se_a_f64": ([_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P], _I),
    "dpmi_tabulate_fusion_se_a_f32": ([_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P], _I),
    "dpmi_tabulate_fusion_se_a_grad_f64": ([_P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P], _I),
    "dpmi_tabulate_fusion_se_a_grad_f32": ([_P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P], _I),
    "dpmi_tabulate_fusion_se_a_grad_grad_f64": ([_P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P], _I),
    "dpmi_tabulate_fusion_se_a_grad_grad_f32": ([_P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _P], _I),
    "dpmi_prod_force_a_f64": ([_P, _P, _P, _P, _I, _I, _I, _I, _P], _I),
    "dpmi_prod_virial_a_f64": ([_P, _P, _P, _P, _P, _P, _I, _I, _I, _P], _I),
    "dpmi_prod_virial_a_f64_noalloc": ([_P, _P, _P, _P, _P, _P, _I, _I, _I, _P], _I),
    "dpmi_prod_force_grad_a_f64": ([_P, _P, _P, _P, _I, _I, _I, _P], _I),
    "dpmi_prod_virial_grad_a_f64": ([_P, _P, _P, _P, _P, _I, _I, _P], _I),
    "dpmi_prod_force_a_f32": ([_P, _P, _P, _P, _I, _I, _I, _I, _P], _I),
    "dpmi_prod_virial_a_f32": ([_P, _P, _P, _P, _P, _P, _I, _I, _I, _P], _I),
    "dpmi_prod_force_grad_a_f32": ([_P, _P, _P, _P, _I, _I, _I, _P], _I),
    "dpmi_prod_virial_grad_a_f32": ([_P, _P, _P, _P, _P, _I, _I, _P], _I),
    "dpmi_tabulate_build_se_a": ([_P, _P, _I, _P, _P, _P, _P, _D, _D, _D, _D, _D, _P], _I),
    "dpmi_sea_model_create": ([_P, _P], _I),
    "dpmi_sea_model_load": ([_P, C.c_char_p, _P], _I),
    "dpmi_sea_model_info": ([_P, _P, _P, _P, _P, _I], _I),
    "dpmi_sea_model_destroy": ([_P], None),
    "dpmi_sea_engine_create": ([_P, _P, _I, _P, _P, _I, _D], _I),
    "dpmi_sea_engine_destroy": ([_P], None),
    "dpmi_sea_engine_compute": ([_P, _P, _P, _P, _P, _P, _P, _P, _P], _I),
    "dpmi_sea_engine_set_growth": ([_P, _I], _I),
    "dpmi_sea_set_net_chunk_bytes": ([C.c_longlong], _I),
    "dpmi_sea_engine_compute_lower": ([_P, _P, _P, _I, _I, _P, _P, _P, _I, _P, _P, _P, _P, _P, _P], _I),
    "dpmi_sea_engine_graph_capture": ([_P, _P, _P, _P, _P, _P, _P, _P], _I),
    "dpmi_sea_engine_graph_capture_device_box": ([_P, _P, _P, _P, _P, _P, _P, _P], _I),
    "dpmi_sea_engine_graph_launch": ([_P, _P], _I),
    "dpmi_sea_engine_autotune": ([_P, _P, _P, _I, _P], _I),
    "dpmi_sea_engine_tuning": ([_P, _P], _I),
    "dpmi_sea_engine_set_tuning": ([_P, _P], _I),
    "dpmi_sea_engine_status": ([_P], _I),
    "dpmi_sea_engine_fetch": ([_P, _I, _I, _P, _P], _I),
    "dpmi_sea_engine_set_timing": ([_P, _I], _I),
    "dpmi_sea_engine_timing": ([_P, _P, _I, _P, _I], _I),
}


class ModelDesc(C.Structure):
    """dpmi_sea_model_desc"""

    _fields_ = [
        ("ntypes", _I),
        ("sel", C.POINTER(_I)),
        ("rcut", _F),
        ("rcut_smth", _F),
        ("last_layer_size", _I),
        ("axis_neuron", _I),
        ("type_one_side", _I),
        ("davg", C.POINTER(_D)),
        ("dstd", C.POINTER(_D)),
        ("nnets", _I),
        ("nspline", C.POINTER(_I)),
        ("tables", C.POINTER(C.POINTER(_D))),
        ("table_info", C.POINTER(_D)),
        ("fit_nlayers", _I),
        ("fit_width", C.POINTER(_I)),
        ("fit_w", C.POINTER(C.POINTER(_D))),
        ("fit_b", C.POINTER(C.POINTER(_D))),
        ("fit_idt", C.POINTER(C.POINTER(_D))),
        ("fit_act", C.POINTER(_I)),
        ("fit_resnet", C.POINTER(_I)),
        ("bias_atom_e", C.POINTER(_D)),
        ("embed_nlayers", _I),
        ("embed_width", C.POINTER(_I)),
        ("embed_w", C.POINTER(C.POINTER(_D))),
        ("embed_b", C.POINTER(C.POINTER(_D))),
        ("embed_idt", C.POINTER(C.POINTER(_D))),
        ("embed_act", C.POINTER(_I)),
        ("embed_resnet", C.POINTER(_I)),
    ]


class DpmiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dpmi {STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                f"{LIBPATH} is missing: build it with `python -m dpmi._build` (HIP extension required)")
        lib = C.CDLL(LIBPATH)
        for name, (argt, rest) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = argt
            fn.restype = rest
        _lib = lib
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise DpmiError(rc, load().dpmi_last_error().decode())
