"""forward_lower through dpmi_sea_engine_compute_lower: the caller's ghosts
and an unsorted rcut + skin candidate list (as an MD code hands them over).

* the reference model's goldens (source/tests/infer/deeppot-testcase.yaml):
  extended forces and per-atom virials folded onto their owners equal the
  golden forces and atomic virials per atom, atomic energies per atom; the
  non-periodic block with no ghosts at all;
* the same system through the periodic engine (its own ghosts and cell
  list): energy, folded forces, virial and per-atom virials agree;
* the mixed engine's lower path within the accuracy budget of FP64;
* local types that differ from the engine's are reported.
"""
import numpy as np
import pytest
import torch

from _fixtures import deeppot_sea_cases, deeppot_sea_model
from dpmi import model as dmodel
from dpmi import systems
from dpmi.engine import DeviceModel, Engine
from dpmi.lib import DpmiError
from oracle import oracle_lib

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, np.float64).reshape(-1), np.asarray(b, np.float64).reshape(-1)
    assert a.shape == b.shape
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def ext_and_rows(coord, atype, box, rc, skin=0.5, seed=1):
    """Extended atoms (ghost shell of rc + skin in a periodic box, none
    without a box) and shuffled candidate rows within rc + skin."""
    O = oracle_lib()
    nloc = len(atype)
    if box is None:
        cext, text, mapping = np.asarray(coord, np.float64).reshape(-1, 3), np.asarray(atype, np.int32), \
            np.arange(nloc, dtype=np.int32)
    else:
        cn = O.normalize_coord(coord, box)
        cext, text, mapping = O.copy_coord(cn, atype, box, rc + skin)
    il, nn, jl = O.build_nlist(cext, nloc, rc + skin, text)
    rng = np.random.default_rng(seed)
    jl = jl.copy()
    for i in range(nloc):
        jl[i, : nn[i]] = rng.permutation(jl[i, : nn[i]])
    return cext, text, mapping, il, nn, jl


def run_lower(eng, cext, text, il, nn, jl, atom_virial=True):
    d = "cuda"
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=d)  # noqa: E731
    e, f, v, ae, av = eng.compute_lower(t(cext, torch.float64), t(text, torch.int32), t(il, torch.int32),
                                        t(nn, torch.int32), t(jl, torch.int32), atom_virial=atom_virial)
    eng.status()
    return (e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy(), ae.cpu().numpy(),
            None if av is None else av.cpu().numpy())


def fold(x, mapping, nloc):
    out = np.zeros((nloc,) + x.shape[1:])
    np.add.at(out, mapping, x)
    return out


def test_lower_reference_model_goldens():
    m = deeppot_sea_model()
    m.build_tables(0.5)
    dm = DeviceModel(m)
    for k, c in enumerate(deeppot_sea_cases()):
        print(f"case golden block {k} (box {'none' if c['box'] is None else 'periodic'})", flush=True)
        coord = np.array(c["coord"]).reshape(-1, 3)
        at = np.array(c["atype"], np.int32)
        box = None if c["box"] is None else np.array(c["box"])
        cext, text, mapping, il, nn, jl = ext_and_rows(coord, at, box, m.rcut)
        eng = Engine(dm, [len(at)], at)
        e, f, v, ae, av = run_lower(eng, cext, text, il, nn, jl)
        nloc = len(at)
        np.testing.assert_allclose(ae, c["atomic_energy"], atol=1e-9, rtol=0)
        np.testing.assert_allclose(fold(f, mapping, nloc).reshape(-1), c["force"], atol=1e-9, rtol=0)
        np.testing.assert_allclose(fold(av, mapping, nloc).reshape(-1), c["atomic_virial"], atol=1e-9, rtol=0)
        np.testing.assert_allclose(v, np.array(c["atomic_virial"]).reshape(-1, 9).sum(0), atol=1e-9, rtol=0)
        assert abs(e[0] - np.sum(c["atomic_energy"])) < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_lower_equals_periodic_engine(precision):
    print(f"case water 512 H2O lower vs periodic engine precision={precision}", flush=True)
    m = dmodel.water_model(seed=3)
    coord, atype, box = systems.water(512, seed=4)
    dm = DeviceModel(m)
    eng = Engine(dm, [len(atype)], atype, precision=precision)
    e0, f0, v0 = [t.cpu().numpy().copy() for t in eng.compute(torch.as_tensor(coord, device="cuda"), box,
                                                                 atom_virial=True)]
    av0 = eng.atom_virial.cpu().numpy().copy()
    eng.status()
    cext, text, mapping, il, nn, jl = ext_and_rows(coord, atype, box, m.rcut, skin=1.0, seed=2)
    e, f, v, ae, av = run_lower(eng, cext, text, il, nn, jl)
    nloc = len(atype)
    tol = 1e-10 if precision == "fp64" else 1e-5
    assert abs(e[0] - e0[0]) <= tol * abs(e0[0])
    assert rel(fold(f, mapping, nloc), f0) < tol
    assert rel(v, v0[0]) < tol
    assert rel(fold(av, mapping, nloc), av0) < tol
    if precision == "mixed":
        ed = Engine(dm, [nloc], atype, precision="fp64")
        e_d, f_d, v_d, _, _ = run_lower(ed, cext, text, il, nn, jl, atom_virial=False)
        assert abs(e[0] - e_d[0]) / nloc * 1e3 < 0.1
        assert np.sqrt(np.mean((fold(f, mapping, nloc) - fold(f_d, mapping, nloc)) ** 2)) * 1e3 < 1.0


def test_lower_reports_type_mismatch():
    m = dmodel.water_model(seed=3)
    coord, atype, box = systems.water(8, seed=4)
    eng = Engine(DeviceModel(m), [len(atype)], atype)
    cext, text, mapping, il, nn, jl = ext_and_rows(coord, atype, box, m.rcut)
    text = text.copy()
    text[0] = 1 - text[0]
    with pytest.raises(DpmiError, match="ARGUMENT"):
        run_lower(eng, cext, text, il, nn, jl)
