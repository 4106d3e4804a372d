"""The HIP table builder (dpmi_tabulate_build_se_a) against the numpy builder
(SeAModel.build_tables), and both against the embedding net itself.

The coefficients a3..a5 amplify last-bit differences of the knot values by
1/tt^3..1/tt^5, so two correct builders agree on them only to a relative
1e-5 or so; what a table must reproduce is checked instead:
* a0, a1, 2 a2 are the net's value, slope and curvature at the knots:
  equal to the numpy builder's within 1e-13 of each column's scale;
* the quintic of every interval, evaluated inside it, equals the numpy
  table's within 1e-12 and the net within the O(h^6) interpolation error;
* the row count, knots and table_info are identical;
* the whole path (FP64 engine) with native tables equals the one with
  numpy tables within 1e-11.
"""
import copy

import numpy as np
import pytest
import torch

from _fixtures import deeppot_sea_model
from dpmi import model as dmodel
from dpmi import systems
from dpmi.engine import DeviceModel, Engine

pytestmark = pytest.mark.gpu


def quintic(table, M, t_frac, info):
    """Evaluates every interval's polynomial at t = t_frac * (its width)."""
    ns = table.shape[0]
    c = table.reshape(ns, M, 6)
    n0 = int((info[1] - info[0]) / info[3])
    tt = np.where(np.arange(ns) < n0, info[3], info[4])[:, None] * t_frac
    return sum(c[:, :, p] * tt ** p for p in range(6))


MODELS = {
    "water": lambda: dmodel.water_model(seed=1),
    "water_env_stat": lambda: dmodel.water_model(seed=2, env_stat=True),
    "water_one_side": lambda: dmodel.water_model(seed=3, type_one_side=True),
    "copper": lambda: dmodel.copper_model(seed=1),
    "reference_model": lambda: _ref_model(),
}


def _ref_model():
    m = deeppot_sea_model()
    m.build_tables(0.5)
    return m


@pytest.mark.parametrize("name", sorted(MODELS))
def test_native_tables_match_numpy(name):
    print(f"case native tables {name}", flush=True)
    m = MODELS[name]()
    ref_tables, ref_info = m.tables, m.table_info
    n = copy.deepcopy(m)
    d_min = 0.5 if name != "copper" else 1.0
    n.build_tables(d_min, native=True)
    np.testing.assert_array_equal(n.table_info, ref_info)
    M = m.M
    for k, (a, b) in enumerate(zip(n.tables, ref_tables)):
        assert a.shape == b.shape
        ca, cb = a.reshape(-1, M, 6), b.reshape(-1, M, 6)
        for p, fac in ((0, 1.0), (1, 1.0), (2, 2.0)):
            scale = np.max(np.abs(cb[:, :, p]), axis=0) + 1e-300
            assert np.max(np.abs(ca[:, :, p] - cb[:, :, p]) / scale) < 1e-13, (k, p)
        for t in (0.25, 0.5, 0.75, 1.0):
            qa, qb = quintic(a, M, t, ref_info[k]), quintic(b, M, t, ref_info[k])
            assert np.max(np.abs(qa - qb)) < 1e-12 * np.max(np.abs(qb)), (k, t)
        # against the net itself at interval midpoints of the fine part
        info = ref_info[k]
        n0 = int((info[1] - info[0]) / info[3])
        xs = info[0] + (np.arange(min(n0, 200)) + 0.5) * info[3]
        direct = m.embed_eval(k, xs)[0]
        mid = quintic(a, M, 0.5, info)[: len(xs)]
        assert np.max(np.abs(mid - direct)) < 1e-9 * max(1.0, np.max(np.abs(direct)))


def test_native_tables_whole_path():
    print("case native tables through the FP64 engine", flush=True)
    m = dmodel.water_model(seed=4, env_stat=True)
    n = copy.deepcopy(m)
    n.build_tables(0.5, native=True)
    coord, atype, box = systems.water(128, seed=5)
    out = []
    for mm in (m, n):
        eng = Engine(DeviceModel(mm), [len(atype)], atype)
        e, f, v = eng.compute(torch.as_tensor(coord, device="cuda"), box)
        eng.status()
        out.append((e.cpu().numpy().copy(), f.cpu().numpy().copy(), v.cpu().numpy().copy()))
    (e0, f0, v0), (e1, f1, v1) = out
    assert abs(e1[0] - e0[0]) <= 1e-11 * abs(e0[0])
    assert np.max(np.abs(f1 - f0)) <= 1e-11 * np.max(np.abs(f0))
    assert np.max(np.abs(v1 - v0)) <= 1e-11 * np.max(np.abs(v0))
