"""Python handle of the whole-path engine (dpmi_sea_* in include/dpmi.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .lib import DPMI_FP64, DPMI_MIXED, ModelDesc, check, load

FETCH = {"nall": (0, np.int32), "coord_ext": (1, np.float64), "atype_ext": (2, np.int32),
         "mapping": (3, np.int32), "nlist": (4, np.int32), "em": (5, np.float64), "descriptor": (6, np.float64),
         "net_deriv": (7, np.float64), "atom_energy": (8, np.float64), "ext_capacity": (9, np.int32),
         "regrows": (10, np.int32), "row_capacity": (11, np.int32)}


def _arr(a, dtype):
    return np.ascontiguousarray(np.asarray(a, dtype=dtype))


def set_net_chunk_bytes(nbytes: int) -> None:
    """Row-buffer budget of the uncompressed embedding for engines created
    afterwards (dpmi_sea_set_net_chunk_bytes)."""
    check(load().dpmi_sea_set_net_chunk_bytes(int(nbytes)))


class DeviceModel:
    """An SeAModel (dpmi.model) uploaded to the GPU."""

    def __init__(self, model, compressed: bool = True):
        """compressed=False uploads the embedding nets instead of the tables
        (the uncompressed path; FP64 engines only)."""
        if compressed and model.tables is None:
            raise ValueError("the model has no compressed table: call build_tables() first")
        keep = []  # numpy buffers referenced by the descriptor until creation returns

        def ptr(a, ct):
            a = _arr(a, np.float64 if ct is C.c_double else np.int32)
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(ct))

        def ptr_array(arrs):
            ps = (C.POINTER(C.c_double) * len(arrs))()
            for k, a in enumerate(arrs):
                ps[k] = C.POINTER(C.c_double)() if a is None else ptr(a, C.c_double)
            keep.append(ps)
            return C.cast(ps, C.POINTER(C.POINTER(C.c_double)))

        d = ModelDesc()
        nt = model.ntypes
        d.ntypes = nt
        d.sel = ptr(model.sel, C.c_int)
        d.rcut = model.rcut
        d.rcut_smth = model.rcut_smth
        d.last_layer_size = model.M
        d.axis_neuron = model.axis_neuron
        d.type_one_side = int(model.type_one_side)
        d.davg = ptr(model.davg.reshape(nt, -1), C.c_double)
        d.dstd = ptr(model.dstd.reshape(nt, -1), C.c_double)
        d.nnets = model.nnets
        if compressed:
            d.nspline = ptr([t.shape[0] for t in model.tables], C.c_int)
            d.tables = ptr_array(model.tables)
            d.table_info = ptr(model.table_info, C.c_double)
        else:
            layers0 = model.embed[0]
            d.embed_nlayers = len(layers0)
            d.embed_width = ptr([1] + [ly.w.shape[1] for ly in layers0], C.c_int)
            d.embed_w = ptr_array([ly.w for net in model.embed for ly in net])
            d.embed_b = ptr_array([ly.b for net in model.embed for ly in net])
            d.embed_idt = ptr_array([ly.idt for net in model.embed for ly in net])
            d.embed_act = ptr([int(ly.act) for ly in layers0], C.c_int)
            d.embed_resnet = ptr([int(ly.resnet) for ly in layers0], C.c_int)
        fa = model.fitting_arrays()
        L = len(fa)
        d.fit_nlayers = L
        d.fit_width = ptr([model.dim_descrpt] + [W.shape[2] for (W, *_r) in fa], C.c_int)
        ws, bs, idts = [], [], []
        for t in range(nt):
            for (W, b, idt, act, res) in fa:
                ws.append(W[t])
                bs.append(b[t])
                idts.append(None if idt is None else idt[t])
        d.fit_w = ptr_array(ws)
        d.fit_b = ptr_array(bs)
        d.fit_idt = ptr_array(idts)
        d.fit_act = ptr([int(act) for (_W, _b, _i, act, _r) in fa], C.c_int)
        d.fit_resnet = ptr([int(res) for (_W, _b, _i, _a, res) in fa], C.c_int)
        d.bias_atom_e = ptr(model.bias_atom_e, C.c_double)
        h = C.c_void_p()
        check(load().dpmi_sea_model_create(C.byref(h), C.byref(d)))
        self.handle = h
        self.model = model

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            load().dpmi_sea_model_destroy(h)
            self.handle = None


class Engine:
    """E+F+V of a fixed set of frames (types fixed, coordinates and boxes per
    call).  ``natoms`` lists the atom count of each frame; ``atype`` holds the
    types of all frames back to back."""

    def __init__(self, dmodel: DeviceModel, natoms, atype, precision: str = "fp64",
                 nall_per_atom: float = 0.0):
        self.dmodel = dmodel
        self.natoms = _arr(natoms, np.int32).reshape(-1)
        self.atype = _arr(atype, np.int32).reshape(-1)
        assert self.atype.size == int(self.natoms.sum())
        self.nframes = len(self.natoms)
        prec = {"fp64": DPMI_FP64, "mixed": DPMI_MIXED}[precision]
        h = C.c_void_p()
        check(load().dpmi_sea_engine_create(C.byref(h), dmodel.handle, self.nframes,
                                            self.natoms.ctypes.data_as(C.c_void_p),
                                            self.atype.ctypes.data_as(C.c_void_p), prec, float(nall_per_atom)))
        self.handle = h
        self.ntot = int(self.natoms.sum())
        dev = torch.device("cuda")
        self.energy = torch.empty(self.nframes, dtype=torch.float64, device=dev)
        self.force = torch.empty((self.ntot, 3), dtype=torch.float64, device=dev)
        self.virial = torch.empty((self.nframes, 9), dtype=torch.float64, device=dev)
        self.atom_energy = torch.empty(self.ntot, dtype=torch.float64, device=dev)
        self.atom_virial = None  # allocated on the first request

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            load().dpmi_sea_engine_destroy(h)
            self.handle = None

    def _check_coord(self, coord: torch.Tensor) -> None:
        if not (coord.dtype == torch.float64 and coord.is_cuda and coord.is_contiguous()):
            raise ValueError("coord must be a contiguous float64 HIP tensor")
        if coord.numel() != 3 * self.ntot:
            raise ValueError(f"coord holds {coord.numel()} values; the engine's frames need {3 * self.ntot}")

    def compute(self, coord: torch.Tensor, box, atom_energy: bool = False, atom_virial: bool = False):
        """coord: [ntot, 3] float64 device tensor; box: [nframes, 9] host
        array.  Returns (energy [nf], force [ntot, 3], virial [nf, 9]) views
        of engine-owned buffers (overwritten by the next call).  With
        atom_virial, the per-atom virials [ntot, 9] are left in
        self.atom_virial."""
        self._check_coord(coord)
        b = _arr(box, np.float64).reshape(self.nframes, 9)
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        if atom_virial and self.atom_virial is None:
            self.atom_virial = torch.empty((self.ntot, 9), dtype=torch.float64, device=self.force.device)
        check(load().dpmi_sea_engine_compute(self.handle, C.c_void_p(coord.data_ptr()), b.ctypes.data_as(C.c_void_p),
                                             C.c_void_p(self.energy.data_ptr()), C.c_void_p(self.force.data_ptr()),
                                             C.c_void_p(self.virial.data_ptr()),
                                             C.c_void_p(self.atom_energy.data_ptr()) if atom_energy else None,
                                             C.c_void_p(self.atom_virial.data_ptr()) if atom_virial else None, st))
        return self.energy, self.force, self.virial

    def compute_lower(self, coord_ext: torch.Tensor, atype_ext: torch.Tensor, ilist: torch.Tensor,
                      numneigh: torch.Tensor, jlist: torch.Tensor, atom_virial: bool = False):
        """forward_lower with the caller's ghosts and candidate rows (one
        frame): coord_ext [nall, 3] float64, atype_ext [nall] int32 (the
        engine's local types first), ilist / numneigh [nloc] int32, jlist
        [nloc, jstride] int32 rows (any order, rcut + skin allowed), all HIP
        tensors.  Returns (energy [1], force_ext [nall, 3], virial [9],
        atom_energy [nloc], atom_virial_ext [nall, 9] or None) as new
        tensors."""
        nall = coord_ext.shape[0]
        if not (coord_ext.dtype == torch.float64 and coord_ext.is_cuda and coord_ext.is_contiguous()
                and coord_ext.numel() == 3 * nall):
            raise ValueError("coord_ext must be a contiguous [nall, 3] float64 HIP tensor")
        for name, t in (("atype_ext", atype_ext), ("ilist", ilist), ("numneigh", numneigh), ("jlist", jlist)):
            if not (t.dtype == torch.int32 and t.is_cuda and t.is_contiguous()):
                raise ValueError(f"{name} must be a contiguous int32 HIP tensor")
        if atype_ext.numel() != nall or jlist.dim() != 2 or jlist.shape[0] != ilist.numel():
            raise ValueError("atype_ext must hold nall types and jlist one row per ilist entry")
        dev = coord_ext.device
        e = torch.empty(1, dtype=torch.float64, device=dev)
        f = torch.empty((nall, 3), dtype=torch.float64, device=dev)
        v = torch.empty(9, dtype=torch.float64, device=dev)
        ae = torch.empty(self.ntot, dtype=torch.float64, device=dev)
        av = torch.empty((nall, 9), dtype=torch.float64, device=dev) if atom_virial else None
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
        check(load().dpmi_sea_engine_compute_lower(
            self.handle, p(coord_ext), p(atype_ext), nall, ilist.numel(), p(ilist), p(numneigh), p(jlist),
            int(jlist.shape[1]), p(e), p(f), p(v), p(ae), p(av) if av is not None else None, st))
        return e, f, v, ae, av

    def set_growth(self, on: bool = True) -> None:
        """Capacity growth of eager steps (on by default): an overflowing
        ghost or neighbour-row buffer is grown and the step recomputed."""
        check(load().dpmi_sea_engine_set_growth(self.handle, int(on)))

    def capture(self, coord: torch.Tensor, box, atom_energy: bool = False):
        """One eager step, then the same step captured into a hipGraph on the
        current (non-default) stream.  replay() re-runs it on the values
        stored in `coord` at that time (same tensor, same boxes)."""
        self._check_coord(coord)
        b = _arr(box, np.float64).reshape(self.nframes, 9)
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(load().dpmi_sea_engine_graph_capture(
            self.handle, C.c_void_p(coord.data_ptr()), b.ctypes.data_as(C.c_void_p),
            C.c_void_p(self.energy.data_ptr()), C.c_void_p(self.force.data_ptr()),
            C.c_void_p(self.virial.data_ptr()), C.c_void_p(self.atom_energy.data_ptr()) if atom_energy else None, st))
        self._graph_coord = coord  # the captured graph reads this storage
        return self.energy, self.force, self.virial

    def capture_device_box(self, coord: torch.Tensor, box: torch.Tensor, atom_energy: bool = False):
        """As capture(), with the boxes in a device tensor [nframes, 9]
        (float64) that every replay re-reads (rescale coordinates and boxes
        in place between replays; a box that needs another cell grid is
        reported by status())."""
        self._check_coord(coord)
        if box.device.type != "cuda" or box.dtype != torch.float64 or not box.is_contiguous() \
                or box.numel() != 9 * self.nframes:
            raise ValueError("box: contiguous float64 device tensor of nframes * 9 values")
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(load().dpmi_sea_engine_graph_capture_device_box(
            self.handle, C.c_void_p(coord.data_ptr()), C.c_void_p(box.data_ptr()),
            C.c_void_p(self.energy.data_ptr()), C.c_void_p(self.force.data_ptr()),
            C.c_void_p(self.virial.data_ptr()), C.c_void_p(self.atom_energy.data_ptr()) if atom_energy else None, st))
        self._graph_coord = coord
        self._graph_box = box  # the captured graph reads this storage
        return self.energy, self.force, self.virial

    def replay(self):
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(load().dpmi_sea_engine_graph_launch(self.handle, st))
        return self.energy, self.force, self.virial

    def autotune(self, coord: torch.Tensor, box, reps: int = 5) -> dict:
        """Picks the stage launch shapes for this size bucket (cached for
        later engines of the same bucket); returns the chosen knobs."""
        self._check_coord(coord)
        b = _arr(box, np.float64).reshape(self.nframes, 9)
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(load().dpmi_sea_engine_autotune(self.handle, C.c_void_p(coord.data_ptr()),
                                              b.ctypes.data_as(C.c_void_p), int(reps), st))
        return self.tuning()

    TUNING_KEYS = ("env_threads_per_block", "embed_fwd_waves", "embed_bwd_waves", "force_rows_per_workgroup",
                   "embed_matrix_cores", "fitting_rows_per_tile")

    def tuning(self) -> dict:
        """Current launch knobs; embed_matrix_cores is a mask (bit 0 forward,
        bit 1 adjoint of the table embedding on the matrix cores, mixed
        engine only)."""
        k = (C.c_int * 6)()
        check(load().dpmi_sea_engine_tuning(self.handle, k))
        return dict(zip(self.TUNING_KEYS, list(k)))

    def set_tuning(self, **knobs) -> None:
        """Sets launch knobs by name (keys of tuning()); the others keep
        their values."""
        bad = set(knobs) - set(self.TUNING_KEYS)
        if bad:
            raise KeyError(f"unknown knobs {sorted(bad)}")
        k = (C.c_int * 6)(0, 0, 0, 0, -1, 0)
        for i, name in enumerate(self.TUNING_KEYS):
            if name in knobs:
                k[i] = int(knobs[name])
        check(load().dpmi_sea_engine_set_tuning(self.handle, k))

    def status(self) -> None:
        check(load().dpmi_sea_engine_status(self.handle))

    def fetch(self, what: str, frame: int = 0) -> np.ndarray:
        code, dt = FETCH[what]
        n = C.c_longlong(0)
        check(load().dpmi_sea_engine_fetch(self.handle, frame, code, None, C.byref(n)))
        out = np.zeros(max(n.value, 1), dtype=dt)
        check(load().dpmi_sea_engine_fetch(self.handle, frame, code, out.ctypes.data_as(C.c_void_p), C.byref(n)))
        return out[: n.value]

    def set_timing(self, on: bool = True) -> None:
        load().dpmi_sea_engine_set_timing(self.handle, int(on))

    def timing(self) -> dict:
        ms = np.zeros(32)
        names = C.create_string_buffer(1024)
        k = load().dpmi_sea_engine_timing(self.handle, ms.ctypes.data_as(C.c_void_p), 32, names, 1024)
        labels = names.value.decode().split(";")
        return {labels[i] if i < len(labels) else f"stage{i}": float(ms[i]) for i in range(max(k, 0))}
