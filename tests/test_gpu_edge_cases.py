"""Edge cases of the reference path, on the GPU, against the oracle.

* sel smaller than the neighbour count: rows truncated per type
  (fmt_nlist.cc:130-144);
* atoms of negative type: no neighbours, zero environment rows, no energy
  or force (prod_env_mat.cc:108-123), i.e. the same E/F/V as the system
  without them;
* types and indices outside the sort-key range: DPMI_ERR_RANGE from the
  kernel-level entry, DPMI_ERR_ARGUMENT from the engine;
* a triclinic box through the whole path;
* lattice planes on the cell faces and a neighbour shell at the cutoff: the
  neighbour walk's per-run windows keep every pair the distance test keeps;
* more candidates per centre than the initial key capacity: the capacity
  grows and the step is retried;
* capacity growth: an engine whose ghost capacity is too small recovers
  and returns the oracle's result; a captured graph is dropped when a later
  step changes the box.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from dpmi import model as dmodel
from dpmi import systems
from dpmi.engine import DeviceModel, Engine
from dpmi.lib import DpmiError, load
from dpmi.model import SeAModel
from oracle import oracle_lib
from oracle.sea import evaluate

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, np.float64).reshape(-1), np.asarray(b, np.float64).reshape(-1)
    assert a.shape == b.shape
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _run(m, coord, atype, box, precision="fp64", mask=0, **kw):
    eng = Engine(DeviceModel(m), [len(atype)], atype, precision=precision, **kw)
    if mask:
        eng.set_tuning(embed_matrix_cores=mask)
    e, f, v = eng.compute(torch.as_tensor(coord, device="cuda"), box, atom_energy=True)
    eng.status()
    return eng, e.cpu().numpy().copy(), f.cpu().numpy().copy(), v.cpu().numpy().copy()


def _vs_oracle(m, coord, atype, box, eng, e, f, v):
    ref = evaluate(m, coord, atype, box, oracle_lib(), keep=True)
    assert np.array_equal(eng.fetch("mapping"), ref.extra["mapping"])
    assert np.array_equal(eng.fetch("nlist").reshape(ref.nlist.shape), ref.nlist)
    assert abs(e[0] - ref.energy) / abs(ref.energy) < 1e-10
    assert rel(f, ref.force) < 1e-10
    assert rel(v[0], ref.virial) < 1e-10
    return ref


def test_sel_truncation():
    print("case sel [10, 20] truncation", flush=True)
    m = SeAModel.random(2, [10, 20], 6.0, 0.5, [25, 50, 100], 16, [240, 240, 240], seed=3)
    m.build_tables(0.5)
    coord, atype, box = systems.water(64, seed=8)
    eng, e, f, v = _run(m, coord, atype, box)
    ref = _vs_oracle(m, coord, atype, box, eng, e, f, v)
    # every row is full in both type sections: the cut at sel was applied
    assert np.all(ref.nlist >= 0)
    em, e_m, f_m, v_m = _run(m, coord, atype, box, "mixed")
    assert np.array_equal(em.fetch("nlist"), eng.fetch("nlist"))
    assert abs(e_m[0] - e[0]) / len(atype) * 1e3 < 0.1
    assert np.sqrt(np.mean((f_m - f) ** 2)) * 1e3 < 1.0


@pytest.mark.parametrize("precision,mask", [("fp64", 0), ("mixed", 0), ("mixed", 12)])
def test_negative_types_are_absent(precision, mask):
    """Atoms of type -1 own no neighbours and appear in no list; mask 12
    runs the mixed engine's row-run embedding kernels (absent rows zeroed
    there too)."""
    print(f"case type -1 atoms precision={precision} mask={mask}", flush=True)
    m = dmodel.water_model(seed=1)
    coord, atype, box = systems.water(64, seed=3)
    neg = np.random.default_rng(1).choice(len(atype), 12, replace=False)
    a2 = atype.copy()
    a2[neg] = -1
    eng, e, f, v = _run(m, coord, a2, box, precision, mask)
    keep = np.setdiff1d(np.arange(len(atype)), neg)
    e2, f2, v2 = _run(m, coord[keep], atype[keep], box, precision, mask)[1:]
    if mask:
        assert np.all(eng.fetch("net_deriv").reshape(len(atype), -1)[neg] == 0.0)
    assert np.all(f[neg] == 0.0)
    assert np.all(eng.fetch("atom_energy")[neg] == 0.0)
    if precision == "fp64":
        _vs_oracle(m, coord, a2, box, eng, e, f, v)
        assert abs(e[0] - e2[0]) <= 1e-12 * abs(e2[0])
        assert rel(f[keep], f2) < 1e-10
        assert rel(v[0], v2[0]) < 1e-10
    else:
        assert abs(e[0] - e2[0]) / len(keep) * 1e3 < 1e-3
        assert np.sqrt(np.mean((f[keep] - f2) ** 2)) * 1e3 < 1e-2


def test_type_out_of_range_errors():
    print("case type / index range errors", flush=True)
    m = dmodel.water_model(seed=1)
    coord, atype, box = systems.water(8, seed=3)
    bad = atype.copy()
    bad[0] = 2  # ntypes = 2
    with pytest.raises(DpmiError, match="ARGUMENT"):
        Engine(DeviceModel(m), [len(bad)], bad)
    # kernel level: a neighbour type >= ntypes, then an index >= nall
    lib = load()
    nloc = len(atype)
    dev = torch.device("cuda")
    c = torch.as_tensor(coord, device=dev).contiguous()
    for case in ("type", "index"):
        t = torch.as_tensor(atype, device=dev).clone()
        ilist = torch.arange(nloc, dtype=torch.int32, device=dev)
        jl = torch.stack([torch.tensor([(i + k) % nloc for k in range(1, 4)], dtype=torch.int32)
                          for i in range(nloc)]).to(dev).contiguous()
        nn = torch.full((nloc,), 3, dtype=torch.int32, device=dev)
        if case == "type":
            t[1] = 200
        else:
            jl[0, 0] = nloc + 5
        nnei = 138
        em = torch.empty(nloc * nnei * 4, dtype=torch.float64, device=dev)
        ed = torch.empty(nloc * nnei * 12, dtype=torch.float64, device=dev)
        rij = torch.empty(nloc * nnei * 3, dtype=torch.float64, device=dev)
        nl = torch.empty(nloc * nnei, dtype=torch.int32, device=dev)
        avg = torch.zeros(2 * nnei * 4, dtype=torch.float64, device=dev)
        std = torch.ones(2 * nnei * 4, dtype=torch.float64, device=dev)
        sel = (C.c_int * 2)(46, 92)
        p = lambda x: C.c_void_p(x.data_ptr())  # noqa: E731
        rc = lib.dpmi_prod_env_mat_a_f64(p(em), p(ed), p(rij), p(nl), p(c), p(t), nloc, p(ilist), p(nn), p(jl), 3,
                                         p(avg), p(std), nloc, nloc, 1, 6.0, 0.5, sel, 2, None)
        torch.cuda.synchronize()
        assert rc == 5, (case, rc, lib.dpmi_last_error())  # DPMI_ERR_RANGE


def _triclinic_water(nmol, seed):
    coord, atype, box = systems.water(nmol, seed=seed)
    L = box[0]
    frac = coord / L
    h = np.array([[L, 0.0, 0.0], [0.21 * L, 0.97 * L, 0.0], [-0.13 * L, 0.17 * L, 1.02 * L]])
    return frac @ h, atype, h.reshape(9)


@pytest.mark.parametrize("nmol", [64, 512])
def test_triclinic_full_path(nmol):
    print(f"case triclinic water nmol={nmol}", flush=True)
    m = dmodel.water_model(seed=2)
    coord, atype, box = _triclinic_water(nmol, seed=6)
    eng, e, f, v = _run(m, coord, atype, box)
    _vs_oracle(m, coord, atype, box, eng, e, f, v)
    em, e_m, f_m, v_m = _run(m, coord, atype, box, "mixed")
    assert abs(e_m[0] - e[0]) / len(atype) * 1e3 < 0.1
    assert np.sqrt(np.mean((f_m - f) ** 2)) * 1e3 < 1.0


@pytest.mark.parametrize("jitter", [0.0, 1e-12])
def test_lattice_on_cell_faces_and_cutoff_shell(jitter):
    """fcc lattice with a = rc / sqrt(5): its tenth neighbour shell lies at the
    cutoff, and with 10 unit cells per edge the cells are 2.5 lattice
    constants wide, so lattice planes lie on the cell faces.  Rows equal the
    oracle's bit for bit (many exactly equal distances and positions)."""
    print(f"case fcc lattice on cell faces jitter={jitter}", flush=True)
    m = dmodel.copper_model(seed=2)
    a = 8.0 / np.sqrt(5.0)
    coord, atype, box = systems.fcc_copper(10, seed=3, sigma=jitter, a=a)
    eng, e, f, v = _run(m, coord, atype, box)
    ref = evaluate(m, coord, atype, box, oracle_lib(), keep=True)
    assert np.array_equal(eng.fetch("nlist").reshape(ref.nlist.shape), ref.nlist)
    assert abs(e[0] - ref.energy) / abs(ref.energy) < 1e-10
    scale = max(float(np.max(np.abs(ref.force))), 1.0)
    assert float(np.max(np.abs(f - ref.force))) < 1e-10 * scale
    em, e_m, f_m, v_m = _run(m, coord, atype, box, "mixed")
    assert np.array_equal(em.fetch("nlist"), eng.fetch("nlist"))


def test_candidate_capacity_growth():
    """sel [10, 20] and rcut 8: about 210 candidates per centre against an
    initial key capacity of 128.  The step grows the capacity, is retried and
    returns the oracle's truncated rows and E/F/V."""
    print("case key capacity growth", flush=True)
    m = SeAModel.random(2, [10, 20], 8.0, 0.5, [25, 50, 100], 16, [240, 240, 240], seed=3)
    m.build_tables(0.5)
    coord, atype, box = systems.water(216, seed=8)
    eng, e, f, v = _run(m, coord, atype, box)
    assert int(eng.fetch("row_capacity")[0]) >= 256
    assert int(eng.fetch("regrows")[0]) >= 1
    _vs_oracle(m, coord, atype, box, eng, e, f, v)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_capacity_growth_recovers(precision):
    """Ghost capacity nloc + 64 for 192 atoms that need 1,536 extended atoms:
    the step grows the buffer, recomputes, and matches the oracle; a later
    step needs no further growth."""
    print(f"case capacity growth precision={precision}", flush=True)
    m = dmodel.water_model(seed=1)
    coord, atype, box = systems.water(64, seed=1)
    eng, e, f, v = _run(m, coord, atype, box, precision, nall_per_atom=1.0)
    assert int(eng.fetch("regrows")[0]) >= 1
    assert int(eng.fetch("ext_capacity")[0]) >= int(eng.fetch("nall")[0])
    ref = evaluate(m, coord, atype, box, oracle_lib())
    if precision == "fp64":
        assert abs(e[0] - ref.energy) / abs(ref.energy) < 1e-10
        assert rel(f, ref.force) < 1e-10
    else:
        assert np.sqrt(np.mean((f - ref.force) ** 2)) * 1e3 < 1.0
    n0 = int(eng.fetch("regrows")[0])
    eng.compute(torch.as_tensor(coord + 0.05, device="cuda"), box)
    eng.status()
    assert int(eng.fetch("regrows")[0]) == n0


def test_growth_off_reports_until_read():
    """With growth off the overflow is only reported (sticky until read)."""
    m = dmodel.water_model(seed=1)
    coord, atype, box = systems.water(64, seed=1)
    eng = Engine(DeviceModel(m), [len(atype)], atype, nall_per_atom=1.0)
    eng.set_growth(False)
    x = torch.as_tensor(coord, device="cuda")
    eng.compute(x, box)
    eng.compute(x, box)
    with pytest.raises(DpmiError, match="MEMORY"):
        eng.status()
    eng.status()


def test_graph_dropped_after_box_change():
    print("case graph invalidation", flush=True)
    m = dmodel.water_model(seed=1)
    coord, atype, box = systems.water(64, seed=2)
    eng = Engine(DeviceModel(m), [len(atype)], atype, precision="mixed")
    x = torch.as_tensor(coord, device="cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        eng.capture(x, box)
        eng.replay()
        s.synchronize()
        # same grid, other box: the graph holds the old box and is dropped
        eng.compute(x, box * 1.01)
        with pytest.raises(DpmiError, match="capture again"):
            eng.replay()
        eng.capture(x, box)
        eng.replay()
        # a box that changes the cell grid reallocates: dropped as well
        eng.compute(x * 1.6, box * 1.6)
        with pytest.raises(DpmiError, match="capture again"):
            eng.replay()
        s.synchronize()
    eng.status()
