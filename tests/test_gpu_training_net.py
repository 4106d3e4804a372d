"""Training with the uncompressed embedding (SeaTrainer(embedding="net")):
the embedding nets' parameters train with the fitting's; forces and virials
run through the hand-written adjoint kernels as in the compressed trainer.

* Energy and forces against the oracle's uncompressed evaluation (FP64).
* Force-loss gradients of the embedding and fitting parameters against torch
  autograd of the plain-torch restatement (tests/_torch_sea.py with the
  embedding nets), FP64 1e-8 and FP32.
* A few Adam steps lower an energy + force loss.
"""
import numpy as np
import pytest
import torch

import _torch_sea as ref
from dpmi import model as dmodel
from dpmi import systems
from dpmi.training import SeaTrainer, env_batch, loss_fn, train_step
from oracle import oracle_lib
from oracle.sea import evaluate

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def water():
    m = dmodel.SeAModel.random(2, [46, 92], 6.0, 0.5, [16, 32, 64], 8, [64, 64], seed=5, resnet_dt=True)
    m.set_env_stat([systems.water(64, seed=1)])
    rng = np.random.default_rng(6)
    for net in m.embed:
        for layer in net:
            layer.b = rng.normal(0.0, 0.1, size=layer.b.shape)
    m.bias_atom_e = np.array([-0.3, 0.2])
    return m


def _frames(nf, nmol, seed):
    cs, bs = [], []
    for f in range(nf):
        c, a, b = systems.water(nmol, seed=seed + f)
        cs.append(c)
        bs.append(b)
    return np.stack(cs), a, np.stack(bs)


def test_net_trainer_matches_oracle(water):
    coords, atype, boxes = _frames(1, 27, 3)
    tr = SeaTrainer(water, torch.float64, embedding="net")
    e, f, v = tr(env_batch(water, coords, atype, boxes, torch.float64))
    o = evaluate(water, coords[0], atype, boxes[0], oracle_lib(), embedding="net")
    assert abs(e.item() - o.energy) <= 1e-10 * abs(o.energy)
    assert rel(f.detach().cpu().numpy()[0], o.force) < 1e-10
    assert rel(v.detach().cpu().numpy()[0], o.virial) < 1e-10


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-8), (torch.float32, 3e-3)])
def test_net_force_loss_gradients_match_torch_autograd(water, dtype, tol):
    print(f"case net dtype={dtype}", flush=True)
    coords, atype, boxes = _frames(2, 27, 11)
    tr = SeaTrainer(water, dtype, embedding="net")
    batch = env_batch(water, coords, atype, boxes, dtype)
    e0, f0, _ = tr(batch, need_virial=False)
    rng = np.random.default_rng(8)
    e_ref = e0.detach() + torch.as_tensor(rng.normal(0, 0.05, e0.shape), dtype=dtype, device=e0.device)
    f_ref = f0.detach() + torch.as_tensor(rng.normal(0, 0.05, f0.shape), dtype=dtype, device=f0.device)
    tr.zero_grad()
    energy, force, _ = tr(batch, need_virial=False)
    loss = loss_fn(energy, force, e_ref, f_ref, batch.nloc)
    loss.backward()
    # restatement: FP64 on the CPU from the trainer's current parameters
    cpu = lambda t: t.detach().double().cpu().requires_grad_(True)  # noqa: E731
    fit = []
    for li in range(tr.nlayers):
        fit.append((cpu(tr.W[li]), cpu(tr.b[li]), cpu(tr.idt[li]) if tr.has_idt[li] else None))
    emb, k = [], 0
    for net in water.embed:
        layers = []
        for ly in net:
            layers.append((cpu(tr.eW[k]), cpu(tr.eb[k]), cpu(tr.eidt[k]) if ly.idt is not None else None, ly.act,
                           ly.resnet))
            k += 1
        emb.append(layers)
    ers, frs = [], []
    for fr in range(coords.shape[0]):
        topo = ref.frame_topology(water, coords[fr], atype, boxes[fr])
        x = topo["x0"].clone().requires_grad_(True)
        E = ref.energy(water, topo, x, fit, embed=emb)
        (g,) = torch.autograd.grad(E, x, create_graph=True)
        ers.append(E)
        frs.append(-g)
    E, F = torch.stack(ers), torch.stack(frs)
    assert rel(energy.detach().cpu().numpy(), E.detach().numpy()) < max(tol, 1e-10)
    assert rel(force.detach().cpu().numpy(), F.detach().numpy()) < max(tol, 1e-10)
    loss_fn(E, F, e_ref.double().cpu(), f_ref.double().cpu(), batch.nloc).backward()
    mine = [p.grad for p in tr.eW] + [p.grad for p in tr.eb] + [p.grad for p in tr.W] + [p.grad for p in tr.b]
    theirs = ([ly[0].grad for net in emb for ly in net] + [ly[1].grad for net in emb for ly in net] +
              [p[0].grad for p in fit] + [p[1].grad for p in fit])
    for a, b in zip(mine, theirs):
        assert rel(a.detach().double().cpu().numpy(), b.numpy()) < tol


def test_net_training_steps_lower_the_loss(water):
    """Labels from another random model of the same architecture; Adam on the
    embedding and fitting parameters, energy + force loss, FP32."""
    coords, atype, boxes = _frames(4, 27, 21)
    teacher = dmodel.SeAModel.random(2, [46, 92], 6.0, 0.5, [16, 32, 64], 8, [64, 64], seed=77, resnet_dt=True)
    teacher.davg, teacher.dstd = water.davg, water.dstd
    batch = env_batch(water, coords, atype, boxes, torch.float32)
    e_ref, f_ref, _ = SeaTrainer(teacher, torch.float32, embedding="net")(batch, need_virial=False)
    e_ref, f_ref = e_ref.detach(), f_ref.detach()
    tr = SeaTrainer(water, torch.float32, embedding="net")
    w0 = [p.detach().clone() for p in tr.eW]
    opt = torch.optim.Adam(tr.parameters(), lr=1e-3)
    losses = [train_step(tr, opt, batch, e_ref, f_ref) for _ in range(20)]
    print("net training losses", losses[0], losses[-1], flush=True)
    assert all(np.isfinite(losses))
    assert losses[-1] < 0.5 * losses[0]
    # the embedding nets moved
    assert any(float((p.detach() - q).abs().max()) > 0 for p, q in zip(tr.eW, w0))
