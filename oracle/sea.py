"""Full se_e2_a E/F/V on the CPU -- TEST INFRASTRUCTURE ONLY.

Restates the reference's glue around the lib kernels, for one periodic frame:

  normalize_coord -> copy_coord -> build_nlist -> prod_env_mat_a
    (deepmd/tf/descriptor/se_a.py:578-666, source/op/tf/prod_env_mat_multi_device.cc:2224-2311)
  per (centre type, neighbour type) block: tabulate_fusion_se_a
    (deepmd/pt/model/descriptor/se_a.py:785-839)
  xyz /= nnei; D = xyz^T . xyz[:, :, :A]                (se_a.py:844-851)
  per-type fitting net + bias_atom_e                     (fitting.py:1079-1104)
  net_deriv = dE/d(em) through fitting, D and tabulate_grad (tf se_a.py:753)
  prod_force_a / prod_virial_a; ghost forces folded onto their owners
    (se_a.py:763-779, make_model.py:216)

``lib`` is either the oracle restatement or the compiled reference library;
with ``embedding="direct"`` the embedding net is evaluated directly instead of
through the compressed table (the uncompressed model, used to pin the glue to
the reference's model-level golden values).
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass
class EFV:
    energy: float
    atom_energy: np.ndarray  # [nloc]
    force: np.ndarray  # [nloc, 3]
    virial: np.ndarray  # [9]
    atom_virial_ext: np.ndarray | None = None  # [nall, 9] (lib convention)
    descriptor: np.ndarray | None = None  # [nloc, M*A]
    nlist: np.ndarray | None = None
    extra: dict | None = None


def _matmul(a, b):
    """FP64 product of the fitting layers: numpy, or libtorch on the CPU when
    the caller sets USE_TORCH (the reference PT CPU path runs these GEMMs in
    libtorch)."""
    if USE_TORCH:
        import torch

        return (torch.from_numpy(np.ascontiguousarray(a)) @ torch.from_numpy(np.ascontiguousarray(b))).numpy()
    return a @ b


USE_TORCH = False


def _fitting_fwd(layers, x):
    """Returns output and the per-layer tape for the backward pass."""
    tape = []
    h = x
    for l in layers:
        z = _matmul(h, l.w) + l.b
        y = np.tanh(z) if l.act else z
        yo = y * l.idt if l.idt is not None else y
        res = 0
        if l.resnet:
            if h.shape[-1] == yo.shape[-1]:
                res = 1
            elif 2 * h.shape[-1] == yo.shape[-1]:
                res = 2
        if res == 1:
            yo = yo + h
        elif res == 2:
            yo = yo + np.concatenate([h, h], axis=-1)
        tape.append((h, y, res))
        h = yo
    return h, tape


def _fitting_bwd(layers, tape, gout):
    g = gout
    for l, (h, y, res) in zip(reversed(layers), reversed(tape)):
        gy = g * l.idt if l.idt is not None else g
        if l.act:
            gy = gy * (1.0 - y * y)
        gh = _matmul(gy, l.w.T)
        if res == 1:
            gh = gh + g
        elif res == 2:
            n = h.shape[-1]
            gh = gh + g[..., :n] + g[..., n:]
        g = gh
    return g


class _Clock:
    """Wall time per stage, accumulated into `out` (seconds) when given."""

    def __init__(self, out):
        import time

        self.out, self.time = out, time.perf_counter
        self.t = self.time()

    def __call__(self, name):
        if self.out is None:
            return
        t = self.time()
        self.out[name] = self.out.get(name, 0.0) + (t - self.t)
        self.t = t


def evaluate(model, coord, atype, box, lib, embedding: str = "table", keep: bool = False,
             stage_seconds: dict | None = None) -> EFV:
    clock = _Clock(stage_seconds)
    nloc = len(atype)
    nnei = model.nnei
    sec = model.sec
    M, A = model.M, model.axis_neuron
    rc, rs = float(model.rcut), float(model.rcut_smth)
    cn = lib.normalize_coord(coord, box)
    cext, text, mapping = lib.copy_coord(cn, atype, box, rc)
    nall = len(text)
    clock("normalize_copy_coord")
    il, nn, jl = lib.build_nlist(cext, nloc, rc, text)
    clock("build_nlist")
    avg = model.davg.reshape(model.ntypes, -1)
    std = model.dstd.reshape(model.ntypes, -1)
    em, ed, rij, nlist = lib.prod_env_mat_a(cext, text, il, nn, jl, avg, std, nloc, rc, rs, model.sel)
    clock("prod_env_mat_a")
    em3 = em.reshape(nloc, nnei, 4)
    xyz = np.zeros((nloc, 4, M))
    blocks = []
    for ti in range(model.ntypes):
        rows = np.nonzero(atype == ti)[0]
        if len(rows) == 0:
            continue
        for tj in range(model.ntypes):
            net = model.net_index(ti, tj)
            rr = np.ascontiguousarray(em3[rows, sec[tj]:sec[tj + 1], :])
            ss = np.ascontiguousarray(rr[:, :, 0])
            if embedding == "table":
                gr = lib.tabulate_fusion_se_a(model.tables[net], model.table_info[net], ss, rr, M)
            else:
                gg = model.embed_eval(net, ss.reshape(-1))[0].reshape(len(rows), -1, M)
                gr = np.einsum("njc,njm->ncm", rr, gg)
            xyz[rows] += gr
            blocks.append((ti, tj, rows, rr, ss, net))
    clock("tabulate_fusion_se_a")
    T = xyz / nnei
    D = np.einsum("icm,ica->ima", T, T[:, :, :A]).reshape(nloc, M * A)
    e = np.zeros(nloc)
    tapes = {}
    for ti in range(model.ntypes):
        rows = np.nonzero(atype == ti)[0]
        if len(rows) == 0:
            continue
        out, tape = _fitting_fwd(model.fitting[ti], D[rows])
        e[rows] = out[:, 0] + model.bias_atom_e[ti]
        tapes[ti] = (rows, tape)
    energy = float(np.sum(e))
    # backward: dE/dD
    gD = np.zeros_like(D)
    for ti, (rows, tape) in tapes.items():
        gD[rows] = _fitting_bwd(model.fitting[ti], tape, np.ones((len(rows), 1)))
    clock("descriptor_fitting_fwd_bwd")
    G = gD.reshape(nloc, M, A)
    gT = np.einsum("ima,ica->icm", G, T[:, :, :A])
    gT[:, :, :A] += np.einsum("ima,icm->ica", G, T)
    gxyz = gT / nnei
    net_deriv = np.zeros((nloc, nnei, 4))
    for ti, tj, rows, rr, ss, net in blocks:
        dy = np.ascontiguousarray(gxyz[rows])
        if embedding == "table":
            dx, dem = lib.tabulate_fusion_se_a_grad(model.tables[net], model.table_info[net], ss, rr, dy)
        else:
            gg, dg, _ = model.embed_eval(net, ss.reshape(-1))
            gg = gg.reshape(len(rows), -1, M)
            dg = dg.reshape(len(rows), -1, M)
            dem = np.einsum("njm,ncm->njc", gg, dy)
            dx = np.einsum("njm,njc,ncm->nj", dg, rr, dy)
        nd = dem.copy()
        nd[:, :, 0] += dx
        net_deriv[rows, sec[tj]:sec[tj + 1], :] = nd
    nd2 = net_deriv.reshape(nloc, nnei * 4)
    clock("tabulate_fusion_se_a_grad")
    fext = lib.prod_force_a(nd2, ed, nlist, nloc, nall)[0]
    vir, avir = lib.prod_virial_a(nd2, ed, rij, nlist, nloc, nall)
    force = np.zeros((nloc, 3))
    np.add.at(force, mapping, fext)
    clock("prod_force_virial")
    extra = None
    if keep:
        extra = dict(em=em, em_deriv=ed, rij=rij, coord_ext=cext, atype_ext=text, mapping=mapping,
                     net_deriv=nd2, force_ext=fext, xyz=xyz, gD=gD)
    return EFV(energy, e, force, vir, avir, D, nlist, extra)
