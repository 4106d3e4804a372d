"""Compressed embedding tables built by the library's HIP builder
(dpmi_tabulate_build_se_a, include/dpmi.h): the same tables as
SeAModel.build_tables (deepmd/utils/tabulate.py:67-370), computed on the
device from the embedding weights."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .lib import check, load


def build_table(layers, lower: float, upper: float, extrapolate: float = 5.0, stride0: float = 0.01,
                stride1: float = 0.1) -> np.ndarray:
    """One net (list of dpmi.model.Layer, input width 1) -> [nspline, 6 M]."""
    lib = load()
    nl = len(layers)
    width = (C.c_int * (nl + 1))(1, *[int(l.w.shape[1]) for l in layers])
    keep = []

    def ptrs(arrs):
        out = (C.POINTER(C.c_double) * nl)()
        for k, a in enumerate(arrs):
            if a is None:
                out[k] = C.POINTER(C.c_double)()
            else:
                a = np.ascontiguousarray(a, dtype=np.float64)
                keep.append(a)
                out[k] = a.ctypes.data_as(C.POINTER(C.c_double))
        keep.append(out)
        return C.cast(out, C.c_void_p)

    w = ptrs([l.w for l in layers])
    b = ptrs([l.b for l in layers])
    idt = ptrs([l.idt for l in layers])
    ns = C.c_int(0)
    args = (nl, width, w, b, idt, float(lower), float(upper), float(extrapolate), float(stride0), float(stride1),
            None)
    check(lib.dpmi_tabulate_build_se_a(None, C.byref(ns), *args))
    M = int(layers[-1].w.shape[1])
    out = np.zeros((ns.value, 6 * M))
    check(lib.dpmi_tabulate_build_se_a(out.ctypes.data_as(C.c_void_p), C.byref(ns), *args))
    return out


def build_tables_native(model, min_nbor_dist: float, extrapolate: float = 5.0, stride0: float = 0.01,
                        stride1: float = 0.1, check_frequency: float = -1.0):
    """Every net of `model` -> (tables, table_info), as build_tables."""
    tables, infos = [], []
    for net, (ll, uu) in enumerate(model.table_ranges(min_nbor_dist)):
        tables.append(build_table(model.embed[net], ll, uu, extrapolate, stride0, stride1))
        infos.append([ll, uu, uu * extrapolate, stride0, stride1, check_frequency])
    return tables, np.asarray(infos, dtype=np.float64)
