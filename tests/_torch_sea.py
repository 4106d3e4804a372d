"""Test-only differentiable restatement of the se_e2_a energy (compressed
table, or the embedding nets themselves) in plain torch: E(local coordinates, fitting parameters), so that forces and
force-loss parameter gradients come from torch autograd (the role the
reference's PT autograd plays, deepmd/pt/model/model/transform_output.py
:70-101).  Ghosts and neighbour lists are taken from the kernel-level path
as fixed topology; everything numeric is recomputed here in FP64:
  rij = x_ext[j] - x[i], x_ext = x[owner] + fixed image shift;
  em = [s/r, s x/r^2, s y/r^2, s z/r^2] normalised (env_mat.cc:124-300,
  switcher.h spline5_switch), padding em = -avg/std;
  table lookup + quintic (tabulate.cc:46-73, 122-158) summed over all slots
  (the sorted-padding fold is the same sum);
  GRRG, fitting and bias as in se_a.py:844-851 and fitting.py:1079-1104.
"""
import numpy as np
import torch

from dpmi import ops


def frame_topology(model, coord, atype, box):
    """Ghost owners/shifts and the sorted neighbour list of one frame, from
    the kernel-level GPU path.  Returns dict of CPU tensors."""
    dev = torch.device("cuda")
    nloc = len(atype)
    at = torch.as_tensor(np.asarray(atype, np.int32), device=dev)
    x = torch.as_tensor(np.asarray(coord, np.float64).copy(), device=dev)
    ops.normalize_coord(x, box)
    cx, ct, mp = ops.copy_coord(x, at, box, model.rcut)
    il, nn, jl, _ = ops.build_nlist(cx, nloc, model.rcut, ct)
    avg = torch.as_tensor(model.davg.reshape(model.ntypes, -1), device=dev)
    std = torch.as_tensor(model.dstd.reshape(model.ntypes, -1), device=dev)
    _, _, _, nl = ops.prod_env_mat_a(cx, ct, il, nn, jl, avg, std, nloc, model.rcut, model.rcut_smth, model.sel)
    xw = x.cpu()
    owner = mp.long().cpu()
    shift = cx.cpu() - xw[owner]
    return {"x0": xw, "owner": owner, "shift": shift, "nlist": nl.long().cpu(), "atype": torch.as_tensor(atype)}


def _switch(r, rs, rc):
    u = (r - rs) / (rc - rs)
    s = u * u * u * (-6 * u * u + 15 * u - 10) + 1
    return torch.where(r < rs, torch.ones_like(r), torch.where(r < rc, s, torch.zeros_like(r)))


def _table_eval(table, info, x):
    """G [n, M] of the quintic table at em_x values x [n] (differentiable in x)."""
    lower, upper, mx, s0, s1 = (float(v) for v in info[:5])
    n0 = int((upper - lower) / s0)
    ntail = n0 + int((np.nextafter(mx, lower) - upper) / s1)
    xd = x.detach()
    idx = torch.where(xd < lower, torch.zeros_like(xd),
                      torch.where(xd < upper, torch.floor((xd - lower) / s0),
                                  torch.where(xd < mx, torch.clamp(n0 + torch.floor((xd - upper) / s1), max=ntail),
                                              torch.full_like(xd, ntail)))).long()
    fi = idx.to(x.dtype)  # an integer tensor times a Python float would round to FP32
    base = torch.where(idx < n0, fi * s0 + lower, (fi - n0) * s1 + upper)
    xx = torch.where(xd < lower, torch.zeros_like(x), torch.where(xd < mx, x - base, (mx - base).expand_as(x)))
    delta = torch.where(xd < lower, x - lower, torch.where(xd >= mx, x - mx, torch.zeros_like(x)))
    M = table.shape[1] // 6
    a = table.reshape(table.shape[0], M, 6)[idx]  # [n, M, 6]
    t = xx[:, None]
    v = a[..., 0] + (a[..., 1] + (a[..., 2] + (a[..., 3] + (a[..., 4] + a[..., 5] * t) * t) * t) * t) * t
    d = a[..., 1] + (2 * a[..., 2] + (3 * a[..., 3] + (4 * a[..., 4] + 5 * a[..., 5] * t) * t) * t) * t
    return v + d * delta[:, None]


def _net_eval(layers, x):
    """Embedding net on inputs x [n] -> [n, M]; layers: list of (W, b, idt or
    None, act, resnet) torch tensors / flags (network/mlp.py:188-219)."""
    h = x.reshape(-1, 1)
    for (W, b, idt, act, resnet) in layers:
        y = h @ W + b
        if act:
            y = torch.tanh(y)
        if idt is not None:
            y = y * idt
        if resnet and y.shape[1] == h.shape[1]:
            y = y + h
        elif resnet and y.shape[1] == 2 * h.shape[1]:
            y = y + torch.cat([h, h], dim=1)
        h = y
    return h


def energy(model, topo, x, params, embed=None):
    """Total energy of one frame.  x [nloc, 3] FP64 (may require grad);
    params: list over fitting layers of (W [nt, in, out], b [nt, out],
    idt [nt, out] or None) torch tensors; embed: None (the compressed table)
    or, per net, the layers of the embedding net as for _net_eval."""
    nloc = x.shape[0]
    nl = topo["nlist"]
    nnei = nl.shape[1]
    xe = x[topo["owner"]] + topo["shift"]  # locals first (owner = self, shift = 0)
    valid = nl >= 0
    j = nl.clamp(min=0)
    # padding slots get a unit displacement before the square root so that
    # no zero-length gradient reaches the coordinates
    rij = torch.where(valid[..., None], xe[j] - xe[:nloc, None, :], torch.ones(3, dtype=x.dtype))
    r_safe = torch.sqrt((rij * rij).sum(-1))
    s = _switch(r_safe, model.rcut_smth, model.rcut)
    em = torch.stack([s / r_safe, s * rij[..., 0] / r_safe ** 2, s * rij[..., 1] / r_safe ** 2,
                      s * rij[..., 2] / r_safe ** 2], -1)
    at = topo["atype"].long()
    avg = torch.as_tensor(model.davg)[at]
    std = torch.as_tensor(model.dstd)[at]
    em = torch.where(valid[..., None], (em - avg) / std, -avg / std)
    sec = np.concatenate([[0], np.cumsum(model.sel)])
    M, A = model.M, model.axis_neuron
    etot = 0.0
    for ti in range(model.ntypes):
        rows = torch.nonzero(at == ti).flatten()
        if rows.numel() == 0:
            continue
        xyz = 0.0
        for tj in range(model.ntypes):
            net = tj if model.type_one_side else tj * model.ntypes + ti
            blk = em[rows, sec[tj]:sec[tj + 1], :]
            if embed is None:
                G = _table_eval(torch.as_tensor(model.tables[net]), model.table_info[net],
                                blk[..., 0].reshape(-1)).reshape(blk.shape[0], blk.shape[1], M)
            else:
                G = _net_eval(embed[net], blk[..., 0].reshape(-1)).reshape(blk.shape[0], blk.shape[1], M)
            xyz = xyz + torch.einsum("rjc,rjk->rck", blk, G)
        T = xyz / nnei
        D = torch.einsum("rck,rca->rka", T, T[:, :, :A]).reshape(rows.numel(), M * A)
        h = D
        L = len(params) - 1
        for li in range(L):
            W, b, idt = params[li]
            y = torch.tanh(h @ W[ti] + b[ti])
            if idt is not None:
                y = y * idt[ti]
            h = y + h if W.shape[1] == W.shape[2] else y
        W, b, _ = params[L]
        etot = etot + ((h @ W[ti])[:, 0] + b[ti][0] + float(model.bias_atom_e[ti])).sum()
    return etot
