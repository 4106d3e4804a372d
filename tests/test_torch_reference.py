"""The plain-torch restatement used as the autograd reference of the
training tests (tests/_torch_sea.py) agrees with the oracle on energy and
forces (1e-12), on the oracle's own ghosts and neighbour list (CPU only)."""
import numpy as np
import torch

import _torch_sea as ref
from dpmi import model as dmodel
from dpmi import systems
from oracle import oracle_lib
from oracle.sea import evaluate


def test_torch_restatement_matches_oracle():
    m = dmodel.SeAModel.random(2, [46, 92], 6.0, 0.5, [25, 50, 100], 16, [240, 240, 240], seed=3, resnet_dt=True)
    rng = np.random.default_rng(4)
    for layers in m.fitting:
        for layer in layers:
            layer.b = rng.normal(0.0, 0.1, size=layer.b.shape)
    m.bias_atom_e = np.array([-0.3, 0.2])
    m.build_tables(0.5)
    coord, atype, box = systems.water(64, seed=31)
    r = evaluate(m, coord, atype, box, oracle_lib(), keep=True)
    nloc = len(atype)
    ce = torch.as_tensor(r.extra["coord_ext"]).reshape(-1, 3)
    mp = torch.as_tensor(r.extra["mapping"]).long()
    x0 = ce[:nloc].clone()
    topo = {"x0": x0, "owner": mp, "shift": ce - x0[mp], "nlist": torch.as_tensor(r.nlist).long().reshape(nloc, -1),
            "atype": torch.as_tensor(atype)}
    params = [(torch.as_tensor(W), torch.as_tensor(b), None if idt is None else torch.as_tensor(idt))
              for (W, b, idt, _a, _r) in m.fitting_arrays()]
    x = x0.clone().requires_grad_(True)
    E = ref.energy(m, topo, x, params)
    (g,) = torch.autograd.grad(E, x)
    assert abs(E.item() - r.energy) <= 1e-12 * abs(r.energy)
    assert np.linalg.norm(-g.numpy() - r.force) <= 1e-12 * np.linalg.norm(r.force)
