"""Whole-path parity: the FP64 engine (cell list .. force/virial) against the
oracle on the same seeded systems.  Ghost order, neighbour lists: bit-exact.
Energy, force, virial: within 1e-10 relative."""
import numpy as np
import pytest
import torch

from _fixtures import deeppot_sea_cases, deeppot_sea_model
from dpmi import model as dmodel
from dpmi import systems
from dpmi.engine import DeviceModel, Engine
from oracle import oracle_lib
from oracle.sea import evaluate

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, np.float64).reshape(-1), np.asarray(b, np.float64).reshape(-1)
    assert a.shape == b.shape
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def water():
    return dmodel.water_model(seed=1)


def _run(m, coord, atype, box, **kw):
    dm = DeviceModel(m)
    eng = Engine(dm, [len(atype)], atype, **kw)
    e, f, v = eng.compute(torch.as_tensor(coord, device="cuda"), box, atom_energy=True)
    eng.status()
    return eng, e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy()


@pytest.mark.parametrize("nmol,seed", [(64, 1), (64, 2), (343, 5)])
def test_engine_fp64_water(water, nmol, seed):
    print(f"case water nmol={nmol} seed={seed}", flush=True)
    coord, atype, box = systems.water(nmol, seed=seed)
    eng, e, f, v = _run(water, coord, atype, box)
    ref = evaluate(water, coord, atype, box, oracle_lib(), keep=True)
    x = ref.extra
    assert int(eng.fetch("nall")[0]) == len(x["atype_ext"])
    assert np.array_equal(eng.fetch("coord_ext").reshape(-1, 3), x["coord_ext"])
    assert np.array_equal(eng.fetch("mapping"), x["mapping"])
    assert np.array_equal(eng.fetch("nlist").reshape(ref.nlist.shape), ref.nlist)
    assert rel(eng.fetch("em"), x["em"]) < 1e-12
    assert rel(eng.fetch("descriptor").reshape(ref.descriptor.shape), ref.descriptor) < 1e-10
    assert rel(eng.fetch("atom_energy"), ref.atom_energy) < 1e-10
    assert rel(eng.fetch("net_deriv"), x["net_deriv"]) < 1e-10
    assert abs(e[0] - ref.energy) / abs(ref.energy) < 1e-10
    assert rel(f, ref.force) < 1e-10
    assert rel(v[0], ref.virial) < 1e-10


def test_engine_reference_model_golden():
    """The reference's se_e2_a test model (type_one_side, resnet_dt, biases),
    compressed, against the model-level goldens of
    source/tests/infer/deeppot-testcase.yaml (PBC blocks)."""
    m = deeppot_sea_model()
    m.build_tables(0.5)
    for c in deeppot_sea_cases():
        if c["box"] is None:
            continue
        coord = np.array(c["coord"]).reshape(-1, 3)
        at = np.array(c["atype"], np.int32)
        eng, e, f, v = _run(m, coord, at, np.array(c["box"]))
        np.testing.assert_allclose(eng.fetch("atom_energy"), c["atomic_energy"], atol=1e-9, rtol=0)
        np.testing.assert_allclose(f.reshape(-1), c["force"], atol=1e-9, rtol=0)
        np.testing.assert_allclose(v[0], np.array(c["atomic_virial"]).reshape(-1, 9).sum(0), atol=1e-9)


def test_engine_repeat_and_move(water):
    """Buffers are reused across calls; a displaced configuration recomputes
    every stage (neighbour list included)."""
    coord, atype, box = systems.water(64, seed=9)
    dm = DeviceModel(water)
    eng = Engine(dm, [len(atype)], atype)
    for shift in (0.0, 0.37, 0.0):
        c2 = coord + shift
        e, f, v = eng.compute(torch.as_tensor(c2, device="cuda"), box)
        ref = evaluate(water, c2, atype, box, oracle_lib())
        assert abs(e.item() - ref.energy) / abs(ref.energy) < 1e-10
        assert rel(f.cpu().numpy(), ref.force) < 1e-10


def test_engine_batched_uneven_frames(water):
    frames = systems.uneven_water_frames(5, seed=3, lo=96, hi=600)
    natoms = [len(a) for (_, a, _) in frames]
    coord = np.concatenate([c for (c, _, _) in frames])
    atype = np.concatenate([a for (_, a, _) in frames])
    box = np.stack([b for (_, _, b) in frames])
    dm = DeviceModel(water)
    eng = Engine(dm, natoms, atype)
    e, f, v = eng.compute(torch.as_tensor(coord, device="cuda"), box)
    eng.status()
    e, f, v = e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy()
    off = 0
    for k, (c, a, b) in enumerate(frames):
        ref = evaluate(water, c, a, b, oracle_lib())
        assert abs(e[k] - ref.energy) / abs(ref.energy) < 1e-10
        assert rel(f[off:off + len(a)], ref.force) < 1e-10
        assert rel(v[k], ref.virial) < 1e-10
        off += len(a)


@pytest.mark.parametrize("ncell", [5, 6])
def test_engine_copper_small(ncell):
    """fcc Cu, rcut 8: 5^3 cells hold about 62 atoms per search cell, 6^3
    about 108 (more than a wave: the ghost emission ranks by counting
    instead of the in-register sort)."""
    print(f"case copper ncell={ncell}", flush=True)
    m = dmodel.copper_model(seed=2)
    coord, atype, box = systems.fcc_copper(ncell, seed=4)
    eng, e, f, v = _run(m, coord, atype, box)
    ref = evaluate(m, coord, atype, box, oracle_lib(), keep=True)
    x = ref.extra
    assert np.array_equal(eng.fetch("coord_ext").reshape(-1, 3), x["coord_ext"])
    assert np.array_equal(eng.fetch("mapping"), x["mapping"])
    assert np.array_equal(eng.fetch("nlist").reshape(ref.nlist.shape), ref.nlist)
    assert abs(e[0] - ref.energy) / abs(ref.energy) < 1e-10
    assert rel(f, ref.force) < 1e-10
    assert rel(v[0], ref.virial) < 1e-10


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_status_reports_ghost_capacity_until_read(water, precision):
    """A ghost shell larger than the engine's capacity (192 atoms in a 12.4 A
    box need 1,536 extended atoms; capacity nloc + 64) is flagged, the
    kernels stay inside the buffers, the status stays set over later steps
    until it is read, and reading clears it."""
    from dpmi.lib import DpmiError

    print(f"case ghost capacity precision={precision}", flush=True)
    coord, atype, box = systems.water(64, seed=1)
    dm = DeviceModel(water)
    eng = Engine(dm, [len(atype)], atype, precision=precision, nall_per_atom=1.0)
    eng.set_growth(False)  # report only (growth is tested in test_gpu_edge_cases.py)
    x = torch.as_tensor(coord, device="cuda")
    eng.compute(x, box)
    eng.compute(x, box)  # a second step does not reset the status
    with pytest.raises(DpmiError, match="MEMORY"):
        eng.status()
    eng.status()  # read once: cleared
    # an engine with room evaluates the same system normally afterwards
    ok = Engine(dm, [len(atype)], atype, precision=precision)
    ok.compute(x, box)
    ok.status()
