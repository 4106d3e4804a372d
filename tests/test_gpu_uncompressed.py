"""The uncompressed se_e2_a embedding (se_a.py:785-839 without compression)
in the FP64 engine: the embedding net evaluated for every neighbour slot on
the matrix cores (fit64.hip layer products over rows laid out net by net),
contracted with the environment matrix, and its adjoint back through the net
to em[.., 0].  Checked against the oracle's uncompressed evaluation
(oracle.sea.evaluate(embedding="net")) and the reference model's goldens.
"""
import numpy as np
import pytest
import torch

from _fixtures import deeppot_sea_cases, deeppot_sea_model
from dpmi import model as dmodel
from dpmi import systems
from dpmi.engine import DeviceModel, Engine, set_net_chunk_bytes
from dpmi.lib import DpmiError
from oracle import oracle_lib
from oracle.sea import evaluate

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a, np.float64).reshape(-1), np.asarray(b, np.float64).reshape(-1)
    assert a.shape == b.shape
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _run(m, coord, atype, box, natoms=None, **kw):
    dm = DeviceModel(m, compressed=False)
    natoms = natoms or [len(atype)]
    eng = Engine(dm, natoms, atype, precision="fp64", **kw)
    e, f, v = eng.compute(torch.as_tensor(coord, device="cuda"), box, atom_energy=True)
    torch.cuda.synchronize()
    eng.status()
    return eng, e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy()


def _check(m, coord, atype, box, e, f, v, ae=None, tol=1e-10):
    ref = evaluate(m, coord, atype, box, oracle_lib(), embedding="net")
    assert abs(e[0] - ref.energy) <= tol * abs(ref.energy)
    assert rel(f, ref.force) < tol
    assert rel(v[0], ref.virial) < tol
    if ae is not None:
        assert rel(ae, ref.atom_energy) < tol


@pytest.fixture(scope="module")
def water_net():
    # data statistics: order-one normalised inputs, so the net is not near-linear
    m = dmodel.water_model(seed=3, env_stat=True)
    return m


@pytest.mark.parametrize("nmol", [27, 128])
def test_water_uncompressed_matches_oracle(water_net, nmol):
    print(f"case uncompressed water nmol={nmol}", flush=True)
    coord, atype, box = systems.water(nmol, seed=11)
    eng, e, f, v = _run(water_net, coord, atype, box)
    _check(water_net, coord, atype, box, e, f, v, eng.fetch("atom_energy"))


def test_reference_model_uncompressed_goldens():
    """The reference's test model (type_one_side, resnet_dt embedding and
    fitting) without compression, against deeppot-testcase.yaml."""
    m = deeppot_sea_model()
    for c in deeppot_sea_cases():
        if c["box"] is None:
            continue
        print("case reference model uncompressed", flush=True)
        coord = np.array(c["coord"]).reshape(-1, 3)
        at = np.array(c["atype"], np.int32)
        eng, e, f, v = _run(m, coord, at, np.array(c["box"]))
        np.testing.assert_allclose(eng.fetch("atom_energy"), c["atomic_energy"], atol=1e-10, rtol=0)
        np.testing.assert_allclose(f.reshape(-1), c["force"], atol=1e-10, rtol=0)
        np.testing.assert_allclose(v[0], np.array(c["atomic_virial"]).reshape(-1, 9).sum(0), atol=1e-10)


def test_chunked_rows_match_single_chunk(water_net):
    """Row buffers too small for one chunk: the net runs chunk by chunk and is
    recomputed in the backward; results equal the one-chunk engine."""
    coord, atype, box = systems.water(64, seed=5)
    _, e1, f1, v1 = _run(water_net, coord, atype, box)
    # one row of the water net holds (2 + 2*175 + 100 + 4*100) doubles: about
    # 40 atoms of 138 slots per chunk here, so 192 atoms make 5 chunks
    set_net_chunk_bytes(40 * 138 * 8 * 852)
    try:
        print("case uncompressed chunked", flush=True)
        _, e2, f2, v2 = _run(water_net, coord, atype, box)
    finally:
        set_net_chunk_bytes(16 << 30)
    assert abs(e2[0] - e1[0]) <= 1e-12 * abs(e1[0])
    assert rel(f2, f1) < 1e-12 and rel(v2, v1) < 1e-12
    _check(water_net, coord, atype, box, e2, f2, v2)


def test_uncompressed_batch_negative_types_and_one_side():
    """Two frames, atoms of type -1 (absent), and a type_one_side model."""
    m = dmodel.SeAModel.random(2, [46, 92], 6.0, 0.5, [16, 32, 64], 8, [64, 64], seed=4, type_one_side=True)
    m.set_env_stat([systems.water(64, seed=1)])
    frames = [systems.water(27, seed=s) for s in (2, 3)]
    coord = np.concatenate([c for (c, _, _) in frames])
    atype = np.concatenate([a for (_, a, _) in frames]).copy()
    atype[[3, 40, 100]] = -1
    box = np.stack([b for (_, _, b) in frames])
    natoms = [len(a) for (_, a, _) in frames]
    print("case uncompressed one-side batch", flush=True)
    eng, e, f, v = _run(m, coord, atype, box, natoms=natoms)
    off = 0
    for k, (c, a, b) in enumerate(frames):
        n = len(a)
        at = atype[off:off + n]
        keep = at >= 0
        ref = evaluate(m, c[keep], at[keep], b, oracle_lib(), embedding="net")
        assert abs(e[k] - ref.energy) <= 1e-10 * abs(ref.energy)
        assert rel(f[off:off + n][keep], ref.force) < 1e-10
        assert np.all(f[off:off + n][~keep] == 0)
        assert rel(v[k], ref.virial) < 1e-10
        off += n


def test_uncompressed_graph_replay(water_net):
    coord, atype, box = systems.water(64, seed=8)
    dm = DeviceModel(water_net, compressed=False)
    eg = Engine(dm, [len(atype)], atype, precision="fp64")
    x = torch.as_tensor(coord, device="cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        eg.capture(x, box)
    s.synchronize()
    for it in range(2):
        print(f"case uncompressed replay {it}", flush=True)
        with torch.cuda.stream(s):
            x.add_(0.01)
            e, f, v = [t.clone() for t in eg.replay()]
        s.synchronize()
        c2 = x.cpu().numpy()
        _check(water_net, c2, atype, box, e.cpu().numpy(), f.cpu().numpy(), v.cpu().numpy())


def test_uncompressed_mixed_is_unsupported(water_net):
    coord, atype, box = systems.water(27, seed=1)
    with pytest.raises(DpmiError, match="UNSUPPORTED"):
        Engine(DeviceModel(water_net, compressed=False), [len(atype)], atype, precision="mixed")
