"""Training step of an se_e2_a model through hand-written adjoints.

The energy is differentiated twice: forces need dE/d(em), and a force loss
needs the gradient of those forces with respect to the fitting parameters.
Every non-library derivative on that path is a hand-written GPU kernel:

* table embedding: ``torch.ops.deepmd.tabulate_fusion_se_a`` with its
  grad and grad-grad kernels (source/op/pt/tabulate_multi_device.cc:764-996);
* forces: :class:`ProdForce`, forward ``prod_force_a`` and backward
  ``prod_force_grad_a`` (source/lib/src/prod_force.cc:24-87,
  prod_force_grad.cc:23-75);
* virials: :class:`ProdVirial`, forward ``prod_virial_a`` and backward
  ``prod_virial_grad_a`` (prod_virial.cc:23-69, prod_virial_grad.cc:22-66).

The GRRG contraction and the fitting net are plain tensor algebra (GEMMs and
tanh) on the device, as in deepmd/pt/model/descriptor/se_a.py:844-851 and
deepmd/pt/model/task/fitting.py:1079-1104.  With ``embedding="net"`` the
table is replaced by the embedding nets themselves (se_a.py:785-839 without
compression: one net per type pair or per neighbour type, tanh layers with
timestep and x / [x, x] residuals, network/mlp.py:188-219), whose parameters
train with the fitting's; their first and second derivatives are torch
autograd of those tensor ops (GEMMs on the ROCm BLAS libraries).  The environment matrix comes from
the kernel-level ``prod_env_mat_a``; neighbour indices are mapped to their
owning local atoms, so the force kernels fold ghost contributions directly
and the adjoint's local-index convention (prod_force_grad.cc:57) is exact.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import ops
from .torch_ops import load as _load_op


# ------------------------------------------------------------ autograd ops --
class ProdForce(torch.autograd.Function):
    """force[f, i] = sum of -net_deriv . em_deriv over i's slots plus the
    terms of slots that point at i (local indices)."""

    @staticmethod
    def forward(ctx, net_deriv, em_deriv, nlist, nloc: int, nframes: int):
        ctx.save_for_backward(em_deriv, nlist)
        ctx.dims = (nloc, nframes)
        return ops.prod_force_a(net_deriv.contiguous(), em_deriv, nlist, nloc, nloc, nframes)

    @staticmethod
    def backward(ctx, grad_force):
        em_deriv, nlist = ctx.saved_tensors
        nloc, nframes = ctx.dims
        g = ops.prod_force_grad_a(grad_force.contiguous(), em_deriv, nlist, nloc, nframes)
        return g, None, None, None, None


class ProdVirial(torch.autograd.Function):
    """virial[f] (lib convention virial[3p+q] = sum nd ed_p rij_q), one
    prod_virial_a call per frame."""

    @staticmethod
    def forward(ctx, net_deriv, em_deriv, rij, nlist, nloc: int, nframes: int):
        ctx.save_for_backward(em_deriv, rij, nlist)
        ctx.dims = (nloc, nframes)
        nd = net_deriv.contiguous()
        out = []
        for f in range(nframes):
            s = slice(f * nloc, (f + 1) * nloc)
            v, _ = ops.prod_virial_a(nd[s], em_deriv[s], rij[s], nlist[s], nloc, nloc)
            out.append(v)
        return torch.stack(out)

    @staticmethod
    def backward(ctx, grad_virial):
        em_deriv, rij, nlist = ctx.saved_tensors
        nloc, nframes = ctx.dims
        gv = grad_virial.contiguous()
        parts = [ops.prod_virial_grad_a(gv[f], em_deriv[f * nloc:(f + 1) * nloc], rij[f * nloc:(f + 1) * nloc],
                                        nlist[f * nloc:(f + 1) * nloc], nloc) for f in range(nframes)]
        return torch.cat(parts), None, None, None, None, None


# ------------------------------------------------------------------ batch --
@dataclass
class EnvBatch:
    """Environment matrix of nframes frames of nloc atoms each (rows packed
    frame by frame)."""
    em: torch.Tensor        # [nf*nloc, nnei, 4]
    em_deriv: torch.Tensor  # [nf*nloc, nnei*12]
    rij: torch.Tensor       # [nf*nloc, nnei*3]
    nlist: torch.Tensor     # [nf*nloc, nnei] local owner indices, -1 padding
    atype: torch.Tensor     # [nf*nloc] int64
    nloc: int
    nframes: int


def env_batch(model, coords, atype, boxes, dtype=torch.float32) -> EnvBatch:
    """Kernel-level normalize -> copy_coord -> build_nlist -> prod_env_mat_a
    per frame (FP64), stored in `dtype`.  coords [nf, nloc, 3], atype [nloc],
    boxes [nf, 9] (numpy)."""
    dev = torch.device("cuda")
    nf, nloc = coords.shape[0], coords.shape[1]
    at = torch.as_tensor(np.asarray(atype, np.int32), device=dev)
    avg = torch.as_tensor(model.davg.reshape(model.ntypes, -1), dtype=torch.float64, device=dev)
    std = torch.as_tensor(model.dstd.reshape(model.ntypes, -1), dtype=torch.float64, device=dev)
    ems, eds, rijs, nls = [], [], [], []
    for f in range(nf):
        x = torch.as_tensor(np.asarray(coords[f], np.float64).copy(), device=dev)
        ops.normalize_coord(x, boxes[f])
        cx, ct, mp = ops.copy_coord(x, at, boxes[f], model.rcut)
        il, nn, jl, _ = ops.build_nlist(cx, nloc, model.rcut, ct)
        em, ed, rij, nl = ops.prod_env_mat_a(cx, ct, il, nn, jl, avg, std, nloc, model.rcut, model.rcut_smth,
                                             model.sel)
        owner = torch.where(nl >= 0, mp.long()[nl.long().clamp(min=0)], torch.full_like(nl.long(), -1))
        ems.append(em)
        eds.append(ed)
        rijs.append(rij)
        nls.append(owner.to(torch.int32))
    nnei = int(sum(model.sel))
    return EnvBatch(torch.cat(ems).to(dtype).reshape(nf * nloc, nnei, 4), torch.cat(eds).to(dtype),
                    torch.cat(rijs).to(dtype), torch.cat(nls).contiguous(), at.long().repeat(nf), nloc, nf)


# ------------------------------------------------------------------ model --
class SeaTrainer(torch.nn.Module):
    """se_e2_a energy with trainable fitting parameters, forces and virials
    through the hand-written adjoints.  embedding="table": the compressed
    table (fixed) through the tabulate op; embedding="net": the embedding
    nets, trainable."""

    def __init__(self, model, dtype=torch.float32, embedding: str = "table"):
        super().__init__()
        if embedding not in ("table", "net"):
            raise ValueError("embedding is 'table' or 'net'")
        dev = torch.device("cuda")
        self.m = model
        self.dtype = dtype
        self.embedding = embedding
        if embedding == "table":
            _load_op()
            self.tables = [torch.as_tensor(t, dtype=dtype, device=dev) for t in model.tables]
            self.infos = [torch.as_tensor(model.table_info[n], dtype=dtype) for n in range(len(model.tables))]
        else:
            # per net and layer: W [n_in, n_out], b, idt (absent: not trained)
            self.eW = torch.nn.ParameterList()
            self.eb = torch.nn.ParameterList()
            self.eidt = torch.nn.ParameterList()
            self.e_shape = []
            for net in model.embed:
                for ly in net:
                    self.eW.append(torch.nn.Parameter(torch.as_tensor(ly.w, dtype=dtype, device=dev)))
                    self.eb.append(torch.nn.Parameter(torch.as_tensor(ly.b, dtype=dtype, device=dev)))
                    self.eidt.append(torch.nn.Parameter(
                        torch.as_tensor(ly.idt if ly.idt is not None else np.zeros_like(ly.b), dtype=dtype,
                                        device=dev), requires_grad=ly.idt is not None))
                    self.e_shape.append((ly.act, ly.resnet, ly.idt is not None))
            self.e_nl = len(model.embed[0])
        fa = model.fitting_arrays()
        self.nlayers = len(fa)
        self.act = [bool(a) for (_w, _b, _i, a, _r) in fa]
        self.res = []
        for (W, _b, _i, _a, r) in fa:
            self.res.append(bool(r) and W.shape[1] == W.shape[2])
        self.W = torch.nn.ParameterList()
        self.b = torch.nn.ParameterList()
        self.idt = torch.nn.ParameterList()
        self.has_idt = []
        for (W, b, idt, _a, _r) in fa:
            self.W.append(torch.nn.Parameter(torch.as_tensor(W, dtype=dtype, device=dev)))
            self.b.append(torch.nn.Parameter(torch.as_tensor(b, dtype=dtype, device=dev)))
            self.has_idt.append(idt is not None)
            self.idt.append(torch.nn.Parameter(torch.as_tensor(idt if idt is not None else np.zeros_like(b),
                                                               dtype=dtype, device=dev),
                                               requires_grad=idt is not None))
        self.bias = torch.as_tensor(model.bias_atom_e, dtype=dtype, device=dev)

    def net_index(self, ti: int, tj: int) -> int:
        return tj if self.m.type_one_side else tj * self.m.ntypes + ti

    def embed_net(self, net: int, x: torch.Tensor) -> torch.Tensor:
        """Embedding net `net` on inputs x [n] -> [n, M] (mlp.py:188-219)."""
        h = x.reshape(-1, 1)
        for li in range(self.e_nl):
            k = net * self.e_nl + li
            act, resnet, has_idt = self.e_shape[k]
            y = h @ self.eW[k] + self.eb[k]
            if act:
                y = torch.tanh(y)
            if has_idt:
                y = y * self.eidt[k]
            if resnet and y.shape[1] == h.shape[1]:
                y = y + h
            elif resnet and y.shape[1] == 2 * h.shape[1]:
                y = y + torch.cat([h, h], dim=1)
            h = y
        return h

    def contract(self, net: int, blk: torch.Tensor) -> torch.Tensor:
        """xyz contribution [r, 4, M] of one neighbour block blk [r, nb, 4]."""
        M = self.m.M
        if self.embedding == "table":
            return torch.ops.deepmd.tabulate_fusion_se_a(self.tables[net], self.infos[net],
                                                         blk[:, :, 0].reshape(-1, 1).contiguous(), blk, M)[0]
        r, nb = blk.shape[0], blk.shape[1]
        G = self.embed_net(net, blk[:, :, 0].reshape(-1)).reshape(r, nb, M)
        return torch.bmm(blk.transpose(1, 2).contiguous(), G)

    def atom_energy(self, em: torch.Tensor, atype: torch.Tensor) -> torch.Tensor:
        """em [n, nnei, 4] -> per-atom energy [n]."""
        m = self.m
        n, nnei = em.shape[0], em.shape[1]
        M, A = m.M, m.axis_neuron
        out = torch.zeros(n, dtype=self.dtype, device=em.device)
        sec = np.concatenate([[0], np.cumsum(m.sel)])
        for ti in range(m.ntypes):
            rows = torch.nonzero(atype == ti).flatten()
            if rows.numel() == 0:
                continue
            e = em[rows]
            xyz = None
            for tj in range(m.ntypes):
                blk = e[:, sec[tj]:sec[tj + 1], :].contiguous()
                o = self.contract(self.net_index(ti, tj), blk)
                xyz = o if xyz is None else xyz + o
            T = xyz / nnei                                   # [r, 4, M]
            D = torch.matmul(T.transpose(1, 2).contiguous(), T[:, :, :A].contiguous())  # [r, M, A]
            x = D.reshape(D.shape[0], M * A)
            for li in range(self.nlayers - 1):
                y = x @ self.W[li][ti] + self.b[li][ti]
                if self.act[li]:
                    y = torch.tanh(y)
                if self.has_idt[li]:
                    y = y * self.idt[li][ti]
                x = y + x if self.res[li] else y
            L = self.nlayers - 1
            ae = (x @ self.W[L][ti])[:, 0] + self.b[L][ti][0] + self.bias[ti]
            out = out.index_copy(0, rows, ae)
        return out

    def forward(self, batch: EnvBatch, need_virial: bool = True):
        """Returns (energy [nf], force [nf, nloc, 3], virial [nf, 9] or None);
        all differentiable with respect to the trainable parameters."""
        em = batch.em.detach().requires_grad_(True)
        ae = self.atom_energy(em, batch.atype)
        energy = ae.reshape(batch.nframes, batch.nloc).sum(1)
        (nd,) = torch.autograd.grad(energy.sum(), em, create_graph=True)
        nd = nd.reshape(em.shape[0], -1)
        force = ProdForce.apply(nd, batch.em_deriv, batch.nlist, batch.nloc, batch.nframes)
        virial = (ProdVirial.apply(nd, batch.em_deriv, batch.rij, batch.nlist, batch.nloc, batch.nframes)
                  if need_virial else None)
        return energy, force, virial


def loss_fn(energy, force, e_ref, f_ref, natoms: int, pref_e: float = 1.0, pref_f: float = 1.0):
    """Energy + force loss as in deepmd/pt/loss/ener.py (MSE per atom for
    energy, per component for force)."""
    le = torch.mean(((energy - e_ref) / natoms) ** 2)
    lf = torch.mean((force - f_ref) ** 2)
    return pref_e * le + pref_f * lf


def train_step(trainer: SeaTrainer, opt, batch: EnvBatch, e_ref, f_ref, pref_e=1.0, pref_f=1.0):
    """One optimiser step; returns the loss before the step."""
    opt.zero_grad(set_to_none=True)
    energy, force, _ = trainer(batch, need_virial=False)
    loss = loss_fn(energy, force, e_ref, f_ref, batch.nloc, pref_e, pref_f)
    loss.backward()
    opt.step()
    return float(loss.detach())
