"""ctypes front-end shared by the oracle restatement and the reference lib.

Both libraries export the same argument lists under the prefixes ``orc_``
(oracle/sea_oracle.cc) and ``ref_`` (oracle/ref_capi.cc over the reference
sources); ``CpuLib`` wraps either with numpy in/out.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libdeepmd_ref.so")

_P = C.c_void_p
_I = C.c_int
_F = C.c_float


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.c_void_p)


def build(ref: bool = False) -> None:
    target = "ref" if ref else "oracle"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


class CpuLib:
    def __init__(self, path: str, prefix: str):
        self.path = path
        self.prefix = prefix
        self.lib = C.CDLL(path)
        sig = {
            "normalize_coord": [_P, _I, _P],
            "compute_cell_info": [_P, _F, _P],
            "copy_coord": [_P, _P, _P, _P, _P, _P, _I, _I, _F, _P],
            "build_nlist": [_P, _P, _P, _P, _P, _I, _I, _I, _F, _I, _P],
            "format_nlist": None,
            "prod_env_mat_a": None,
            "prod_force_a": [_P, _P, _P, _P, _I, _I, _I, _I],
            "prod_virial_a": [_P, _P, _P, _P, _P, _P, _I, _I, _I],
            "prod_force_grad_a": [_P, _P, _P, _P, _I, _I, _I],
            "prod_virial_grad_a": [_P, _P, _P, _P, _P, _I, _I],
        }
        for suf in ("f64", "f32"):
            sig[f"tabulate_fusion_se_a_{suf}"] = [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I]
            sig[f"tabulate_fusion_se_a_grad_{suf}"] = [_P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I]
            sig[f"tabulate_fusion_se_a_grad_grad_{suf}"] = [_P, _P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I]
        for name, argt in sig.items():
            fn = getattr(self.lib, prefix + name)
            if argt is not None:
                fn.argtypes = argt
        # argument lists differ between the two libraries for these two
        if prefix == "ref_":
            self.lib.ref_format_nlist.argtypes = [_P, _I, _P, _P, _P, _I, _P, _P, _I, _I, _F, _P, _I]
            self.lib.ref_prod_env_mat_a.argtypes = [_P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _I, _I, _P, _P,
                                                    _I, _I, _I, _F, _F, _P, _I]
            self.lib.ref_last_error.restype = C.c_char_p
        else:
            self.lib.orc_format_nlist.argtypes = [_P, _I, _P, _P, _P, _I, _P, _P, _F, _P, _I]
            self.lib.orc_prod_env_mat_a.argtypes = [_P, _P, _P, _P, _P, _P, _I, _P, _P, _P, _I, _P, _P,
                                                    _I, _I, _I, _F, _F, _P, _I]

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    def _check(self, rc):
        if rc is not None and rc < 0 and self.prefix == "ref_":
            raise RuntimeError(self.lib.ref_last_error().decode())
        return rc

    # ------------------------------------------------------------------
    def normalize_coord(self, coord: np.ndarray, box: np.ndarray) -> np.ndarray:
        c = np.ascontiguousarray(coord, dtype=np.float64).copy()
        b = np.ascontiguousarray(box, dtype=np.float64)
        self._check(self._f("normalize_coord")(_ptr(c), c.size // 3, _ptr(b)))
        return c

    def compute_cell_info(self, rcut: float, box: np.ndarray) -> np.ndarray:
        ci = np.zeros(23, np.int32)
        b = np.ascontiguousarray(box, dtype=np.float64)
        self._check(self._f("compute_cell_info")(_ptr(ci), rcut, _ptr(b)))
        return ci

    def copy_coord(self, coord, atype, box, rcut, mem_nall=None):
        """-> (coord_ext [nall,3], atype_ext [nall], mapping [nall])."""
        c = np.ascontiguousarray(coord, dtype=np.float64).reshape(-1)
        t = np.ascontiguousarray(atype, dtype=np.int32)
        b = np.ascontiguousarray(box, dtype=np.float64)
        nloc = t.size
        mem = mem_nall or max(64 * nloc, 1024)
        while True:
            oc = np.zeros(mem * 3)
            ot = np.zeros(mem, np.int32)
            om = np.zeros(mem, np.int32)
            nall = C.c_int(0)
            rc = self._check(self._f("copy_coord")(_ptr(oc), _ptr(ot), _ptr(om), C.byref(nall), _ptr(c), _ptr(t),
                                                   nloc, mem, rcut, _ptr(b)))
            if rc == 0:
                n = nall.value
                return oc[: 3 * n].reshape(n, 3).copy(), ot[:n].copy(), om[:n].copy()
            mem *= 2

    def build_nlist(self, coord_ext, nloc, rcut, atype_ext=None, mem_size=None, nframes=1):
        """-> (ilist, numneigh, jlist [nrows, mem_size])."""
        c = np.ascontiguousarray(coord_ext, dtype=np.float64).reshape(-1)
        nall = c.size // 3 // nframes
        t = None if atype_ext is None else np.ascontiguousarray(atype_ext, dtype=np.int32)
        mem = mem_size or 256
        nrows = nframes * nloc
        while True:
            il = np.zeros(nrows, np.int32)
            nn = np.zeros(nrows, np.int32)
            jl = np.full((nrows, mem), -1, np.int32)
            mx = C.c_int(0)
            rc = self._check(self._f("build_nlist")(_ptr(il), _ptr(nn), _ptr(jl), C.byref(mx), _ptr(c), nloc,
                                                    nall, mem, rcut, nframes, _ptr(t)))
            if rc == 0:
                return il, nn, jl
            mem *= 2

    def format_nlist(self, ilist, numneigh, jlist, coord_ext, atype_ext, nloc, rcut, sel):
        coord = np.ascontiguousarray(coord_ext, dtype=np.float64).reshape(-1)
        t = np.ascontiguousarray(atype_ext, dtype=np.int32)
        nall = t.size
        sel_a = np.asarray(sel, np.int32)
        nnei = int(sel_a.sum())
        out = np.full((nloc, nnei), -1, np.int32)
        jl = np.ascontiguousarray(jlist, dtype=np.int32)
        if self.prefix == "ref_":
            self._check(self.lib.ref_format_nlist(_ptr(out), len(ilist), _ptr(ilist), _ptr(numneigh), _ptr(jl),
                                                  jl.shape[1], _ptr(coord), _ptr(t), nloc, nall, rcut, _ptr(sel_a),
                                                  len(sel_a)))
        else:
            self.lib.orc_format_nlist(_ptr(out), len(ilist), _ptr(ilist), _ptr(numneigh), _ptr(jl), jl.shape[1],
                                      _ptr(coord), _ptr(t), rcut, _ptr(sel_a), len(sel_a))
        return out

    def prod_env_mat_a(self, coord_ext, atype_ext, ilist, numneigh, jlist, avg, std, nloc, rcut, rcut_smth, sel,
                       nframes=1):
        """-> em [nf*nloc, nnei*4], em_deriv [nf*nloc, nnei*12], rij [nf*nloc, nnei*3], nlist [nf*nloc, nnei]."""
        coord = np.ascontiguousarray(coord_ext, dtype=np.float64).reshape(-1)
        t = np.ascontiguousarray(atype_ext, dtype=np.int32).reshape(-1)
        nall = t.size // nframes
        sel_a = np.asarray(sel, np.int32)
        nnei = int(sel_a.sum())
        nrows = nframes * nloc
        em = np.zeros((nrows, nnei * 4))
        ed = np.zeros((nrows, nnei * 12))
        rij = np.zeros((nrows, nnei * 3))
        nl = np.zeros((nrows, nnei), np.int32)
        jl = np.ascontiguousarray(jlist, dtype=np.int32)
        a = np.ascontiguousarray(avg, dtype=np.float64)
        s = np.ascontiguousarray(std, dtype=np.float64)
        if self.prefix == "ref_":
            self._check(self.lib.ref_prod_env_mat_a(_ptr(em), _ptr(ed), _ptr(rij), _ptr(nl), _ptr(coord), _ptr(t),
                                                    len(ilist), _ptr(ilist), _ptr(numneigh), _ptr(jl), jl.shape[1],
                                                    jl.shape[1], _ptr(a), _ptr(s), nloc, nall, nframes, rcut,
                                                    rcut_smth, _ptr(sel_a), len(sel_a)))
        else:
            self.lib.orc_prod_env_mat_a(_ptr(em), _ptr(ed), _ptr(rij), _ptr(nl), _ptr(coord), _ptr(t), len(ilist),
                                        _ptr(ilist), _ptr(numneigh), _ptr(jl), jl.shape[1], _ptr(a), _ptr(s), nloc,
                                        nall, nframes, rcut, rcut_smth, _ptr(sel_a), len(sel_a))
        return em, ed, rij, nl

    # ------------------------------------------------------------------
    @staticmethod
    def _suf(dtype):
        return "f64" if np.dtype(dtype) == np.float64 else "f32"

    def tabulate_fusion_se_a(self, table, info, em_x, em, M, is_sorted=True):
        dt = np.asarray(table).dtype
        table = np.ascontiguousarray(table, dtype=dt)
        info = np.ascontiguousarray(info, dtype=dt)
        em = np.ascontiguousarray(em, dtype=dt)
        em_x = np.ascontiguousarray(em_x, dtype=dt)
        nloc, nnei, nd = em.shape
        out = np.zeros((nloc, nd, M), dt)
        self._check(self._f("tabulate_fusion_se_a_" + self._suf(dt))(
            _ptr(out), _ptr(table), _ptr(info), _ptr(em_x), _ptr(em), nloc, nnei, M, int(is_sorted), nd))
        return out

    def tabulate_fusion_se_a_grad(self, table, info, em_x, em, dy, is_sorted=True):
        dt = np.asarray(table).dtype
        table = np.ascontiguousarray(table, dtype=dt)
        info = np.ascontiguousarray(info, dtype=dt)
        em = np.ascontiguousarray(em, dtype=dt)
        em_x = np.ascontiguousarray(em_x, dtype=dt)
        dy = np.ascontiguousarray(dy, dtype=dt)
        nloc, nnei, nd = em.shape
        M = dy.shape[-1]
        dx = np.zeros((nloc, nnei), dt)
        dem = np.zeros((nloc, nnei, nd), dt)
        self._check(self._f("tabulate_fusion_se_a_grad_" + self._suf(dt))(
            _ptr(dx), _ptr(dem), _ptr(table), _ptr(info), _ptr(em_x), _ptr(em), _ptr(dy), nloc, nnei, M,
            int(is_sorted), nd))
        return dx, dem

    def tabulate_fusion_se_a_grad_grad(self, table, info, em_x, em, dz_dx, dz_dem, M, is_sorted=True):
        dt = np.asarray(table).dtype
        args = [np.ascontiguousarray(a, dtype=dt) for a in (table, info, em_x, em, dz_dx, dz_dem)]
        nloc, nnei, nd = args[3].shape
        out = np.zeros((nloc, nd, M), dt)
        self._check(self._f("tabulate_fusion_se_a_grad_grad_" + self._suf(dt))(
            _ptr(out), *[_ptr(a) for a in args], nloc, nnei, M, int(is_sorted), nd))
        return out

    def prod_force_a(self, net_deriv, em_deriv, nlist, nloc, nall, nframes=1):
        nd = np.ascontiguousarray(net_deriv, dtype=np.float64)
        ed = np.ascontiguousarray(em_deriv, dtype=np.float64)
        nl = np.ascontiguousarray(nlist, dtype=np.int32)
        nnei = nl.shape[-1]
        f = np.zeros((nframes, nall, 3))
        self._check(self._f("prod_force_a")(_ptr(f), _ptr(nd), _ptr(ed), _ptr(nl), nloc, nall, nnei, nframes))
        return f

    def prod_virial_a(self, net_deriv, em_deriv, rij, nlist, nloc, nall):
        nd = np.ascontiguousarray(net_deriv, dtype=np.float64)
        ed = np.ascontiguousarray(em_deriv, dtype=np.float64)
        r = np.ascontiguousarray(rij, dtype=np.float64)
        nl = np.ascontiguousarray(nlist, dtype=np.int32)
        nnei = nl.shape[-1]
        v = np.zeros(9)
        av = np.zeros((nall, 9))
        self._check(self._f("prod_virial_a")(_ptr(v), _ptr(av), _ptr(nd), _ptr(ed), _ptr(r), _ptr(nl), nloc, nall,
                                             nnei))
        return v, av

    def prod_force_grad_a(self, grad, em_deriv, nlist, nloc, nframes=1):
        g = np.ascontiguousarray(grad, dtype=np.float64)
        ed = np.ascontiguousarray(em_deriv, dtype=np.float64)
        nl = np.ascontiguousarray(nlist, dtype=np.int32)
        nnei = nl.shape[-1]
        out = np.zeros((nframes * nloc, 4 * nnei))
        self._check(self._f("prod_force_grad_a")(_ptr(out), _ptr(g), _ptr(ed), _ptr(nl), nloc, nnei, nframes))
        return out

    def prod_virial_grad_a(self, grad, em_deriv, rij, nlist, nloc):
        g = np.ascontiguousarray(grad, dtype=np.float64)
        ed = np.ascontiguousarray(em_deriv, dtype=np.float64)
        r = np.ascontiguousarray(rij, dtype=np.float64)
        nl = np.ascontiguousarray(nlist, dtype=np.int32)
        nnei = nl.shape[-1]
        out = np.zeros((nloc, 4 * nnei))
        self._check(self._f("prod_virial_grad_a")(_ptr(out), _ptr(g), _ptr(ed), _ptr(r), _ptr(nl), nloc, nnei))
        return out


_cache = {}


def oracle_lib() -> CpuLib:
    if "orc" not in _cache:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        _cache["orc"] = CpuLib(ORACLE_SO, "orc_")
    return _cache["orc"]


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib() -> CpuLib | None:
    if not ref_available():
        return None
    if "ref" not in _cache:
        _cache["ref"] = CpuLib(REF_SO, "ref_")
    return _cache["ref"]
