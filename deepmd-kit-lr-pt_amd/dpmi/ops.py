"""torch-tensor front-end of the kernel-level C ABI.

One function per reference lib function on the se_e2_a path (names follow
deepmd::*_gpu in source/lib/include); arguments are CUDA(HIP) tensors with the
reference's layouts, small host parameters are Python values or numpy
arrays.  Every call runs on torch's current stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .lib import check, load


def _p(t):
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "device tensors must be contiguous HIP tensors"
    return C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _host(a, dtype):
    a = np.ascontiguousarray(np.asarray(a, dtype=dtype))
    return a, a.ctypes.data_as(C.c_void_p)


def normalize_coord(coord: torch.Tensor, box) -> torch.Tensor:
    """coord.h normalize_coord_gpu: wraps coord [natom, 3] in place."""
    b, bp = _host(box, np.float64)
    check(load().dpmi_normalize_coord_f64(_p(coord), coord.numel() // 3, bp, _stream()))
    return coord


def copy_coord(coord: torch.Tensor, atype: torch.Tensor, box, rcut: float, mem_nall: int | None = None):
    """coord.h copy_coord_gpu -> (coord_ext [nall,3], atype_ext [nall], mapping [nall])."""
    nloc = atype.numel()
    mem = mem_nall or max(64 * nloc, 1024)
    lib = load()
    b, bp = _host(box, np.float64)
    while True:
        oc = torch.empty((mem, 3), dtype=torch.float64, device=coord.device)
        ot = torch.empty(mem, dtype=torch.int32, device=coord.device)
        om = torch.empty(mem, dtype=torch.int32, device=coord.device)
        nall = C.c_int(0)
        rc = lib.dpmi_copy_coord_f64(_p(oc), _p(ot), _p(om), C.byref(nall), _p(coord), _p(atype), nloc, mem,
                                     float(rcut), bp, _stream())
        if rc == 4 and mem_nall is None:
            mem = max(2 * mem, nall.value)
            continue
        check(rc)
        n = nall.value
        return oc[:n], ot[:n], om[:n]


def build_nlist(coord_ext: torch.Tensor, nloc: int, rcut: float, atype_ext=None, mem_size: int = 256,
                nframes: int = 1):
    """neighbor_list.h build_nlist_gpu -> (ilist, numneigh, jlist [nrows, mem_size], max_list_size)."""
    lib = load()
    nall = coord_ext.numel() // 3 // nframes
    rows = nframes * nloc
    while True:
        il = torch.empty(rows, dtype=torch.int32, device=coord_ext.device)
        nn = torch.empty(rows, dtype=torch.int32, device=coord_ext.device)
        jl = torch.full((rows, mem_size), -1, dtype=torch.int32, device=coord_ext.device)
        mx = C.c_int(0)
        rc = lib.dpmi_build_nlist_f64(_p(il), _p(nn), _p(jl), C.byref(mx), _p(coord_ext), nloc, nall, mem_size,
                                      float(rcut), nframes, _p(atype_ext), _stream())
        if rc == 3:
            mem_size = max(2 * mem_size, mx.value)
            continue
        check(rc)
        return il, nn, jl, mx.value


def prod_env_mat_a(coord_ext, atype_ext, ilist, numneigh, jlist, avg, std, nloc, rcut, rcut_smth, sel, nframes=1):
    """prod_env_mat.h prod_env_mat_a_gpu -> (em, em_deriv, rij, nlist)."""
    sel_a, sp = _host(sel, np.int32)
    nnei = int(sel_a.sum())
    nall = atype_ext.numel() // nframes
    rows = nframes * nloc
    dev = coord_ext.device
    em = torch.empty((rows, nnei * 4), dtype=torch.float64, device=dev)
    ed = torch.empty((rows, nnei * 12), dtype=torch.float64, device=dev)
    rij = torch.empty((rows, nnei * 3), dtype=torch.float64, device=dev)
    nl = torch.empty((rows, nnei), dtype=torch.int32, device=dev)
    check(load().dpmi_prod_env_mat_a_f64(_p(em), _p(ed), _p(rij), _p(nl), _p(coord_ext), _p(atype_ext), len(ilist),
                                         _p(ilist), _p(numneigh), _p(jlist), jlist.shape[1], _p(avg), _p(std), nloc,
                                         nall, nframes, float(rcut), float(rcut_smth), sp, len(sel_a), _stream()))
    return em, ed, rij, nl


def _suf(t):
    return "f64" if t.dtype == torch.float64 else "f32"


def _info(table_info, dtype):
    return _host(table_info.detach().cpu().numpy() if torch.is_tensor(table_info) else table_info,
                 np.float64 if dtype == torch.float64 else np.float32)


def tabulate_fusion_se_a(table, table_info, em_x, em, last_layer_size: int, is_sorted: bool = True):
    """tabulate.h tabulate_fusion_se_a_gpu -> out [nloc, ndescrpt, M]."""
    nloc, nnei, nd = em.shape
    out = torch.empty((nloc, nd, last_layer_size), dtype=table.dtype, device=table.device)
    ti, tp = _info(table_info, table.dtype)
    fn = getattr(load(), "dpmi_tabulate_fusion_se_a_" + _suf(table))
    check(fn(_p(out), _p(table), tp, _p(em_x), _p(em), nloc, nnei, int(last_layer_size), int(is_sorted), nd,
             _stream()))
    return out


def tabulate_fusion_se_a_grad(table, table_info, em_x, em, dy, is_sorted: bool = True):
    nloc, nnei, nd = em.shape
    M = dy.shape[-1]
    dx = torch.zeros((nloc, nnei), dtype=table.dtype, device=table.device)
    dem = torch.zeros((nloc, nnei, nd), dtype=table.dtype, device=table.device)
    ti, tp = _info(table_info, table.dtype)
    fn = getattr(load(), "dpmi_tabulate_fusion_se_a_grad_" + _suf(table))
    check(fn(_p(dx), _p(dem), _p(table), tp, _p(em_x), _p(em), _p(dy), nloc, nnei, M, int(is_sorted), nd,
             _stream()))
    return dx, dem


def tabulate_fusion_se_a_grad_grad(table, table_info, em_x, em, dz_dy_dem_x, dz_dy_dem, last_layer_size: int,
                                   is_sorted: bool = True):
    nloc, nnei, nd = em.shape
    out = torch.empty((nloc, nd, last_layer_size), dtype=table.dtype, device=table.device)
    ti, tp = _info(table_info, table.dtype)
    fn = getattr(load(), "dpmi_tabulate_fusion_se_a_grad_grad_" + _suf(table))
    check(fn(_p(out), _p(table), tp, _p(em_x), _p(em), _p(dz_dy_dem_x), _p(dz_dy_dem), nloc, nnei,
             int(last_layer_size), int(is_sorted), nd, _stream()))
    return out


def prod_force_a(net_deriv, em_deriv, nlist, nloc, nall, nframes=1):
    """prod_force.h prod_force_a_gpu -> force [nframes, nall, 3] (dtype of net_deriv)."""
    nnei = nlist.shape[-1]
    f = torch.empty((nframes, nall, 3), dtype=net_deriv.dtype, device=net_deriv.device)
    fn = getattr(load(), "dpmi_prod_force_a_" + _suf(net_deriv))
    check(fn(_p(f), _p(net_deriv), _p(em_deriv), _p(nlist), nloc, nall, nnei, nframes, _stream()))
    return f


def prod_virial_a(net_deriv, em_deriv, rij, nlist, nloc, nall):
    """prod_virial.h prod_virial_a_gpu (one frame) -> (virial [9], atom_virial [nall, 9])."""
    nnei = nlist.shape[-1]
    v = torch.empty(9, dtype=net_deriv.dtype, device=net_deriv.device)
    av = torch.empty((nall, 9), dtype=net_deriv.dtype, device=net_deriv.device)
    fn = getattr(load(), "dpmi_prod_virial_a_" + _suf(net_deriv))
    check(fn(_p(v), _p(av), _p(net_deriv), _p(em_deriv), _p(rij), _p(nlist), nloc, nall, nnei, _stream()))
    return v, av


def prod_force_grad_a(grad, em_deriv, nlist, nloc, nframes=1):
    """prod_force_grad.h: d(loss)/d(net_deriv) [nframes*nloc, 4*nnei] from d(loss)/d(force) [nframes, nloc, 3]."""
    nnei = nlist.shape[-1]
    out = torch.empty((nframes * nloc, 4 * nnei), dtype=grad.dtype, device=grad.device)
    fn = getattr(load(), "dpmi_prod_force_grad_a_" + _suf(grad))
    check(fn(_p(out), _p(grad), _p(em_deriv), _p(nlist), nloc, nnei, nframes, _stream()))
    return out


def prod_virial_grad_a(grad, em_deriv, rij, nlist, nloc):
    """prod_virial_grad.h (one frame): d(loss)/d(net_deriv) [nloc, 4*nnei] from d(loss)/d(virial) [9]."""
    nnei = nlist.shape[-1]
    out = torch.empty((nloc, 4 * nnei), dtype=grad.dtype, device=grad.device)
    fn = getattr(load(), "dpmi_prod_virial_grad_a_" + _suf(grad))
    check(fn(_p(out), _p(grad), _p(em_deriv), _p(rij), _p(nlist), nloc, nnei, _stream()))
    return out
