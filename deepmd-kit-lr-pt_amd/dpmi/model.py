"""se_e2_a model parameters and the compressed embedding table.

A model here is the data the se_e2_a energy path consumes:

* descriptor: ``rcut``, ``rcut_smth``, ``sel``, embedding widths
  ``neuron`` (1 -> n0 -> n1 -> n2 = M), ``axis_neuron`` (A), normalisation
  ``davg``/``dstd`` of shape [ntypes, nnei, 4], and one embedding net per
  (centre type, neighbour type) pair, or one per neighbour type when
  ``type_one_side`` (deepmd/pt/model/descriptor/se_a.py:706-740, 785-839);
* fitting: one net per centre type, D = M*A -> fit[0] -> fit[1] -> ... -> 1
  with tanh, residual on equal widths, optional ``idt`` (resnet_dt), and a
  per-type ``bias_atom_e`` (deepmd/pt/model/task/fitting.py:1079-1104,
  deepmd/pt/model/network/mlp.py:188-219, deepmd/dpmodel/utils/network.py:850-1030);
* the compressed embedding table: per net a ``[nspline, 6*M]`` array of
  quintic Hermite coefficients and ``table_info = [lower, upper,
  upper*extrapolate, stride0, stride1, check_frequency]``
  (deepmd/utils/tabulate.py:67-370, 468-496).

Embedding-net index convention: net ``e = tj * ntypes + ti`` serves centre
type ti and neighbour type tj (se_a.py:715-719); with ``type_one_side`` net
``tj`` serves every centre type.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np


@dataclasses.dataclass
class Layer:
    w: np.ndarray  # [n_in, n_out]
    b: np.ndarray  # [n_out]
    idt: np.ndarray | None = None  # [n_out] (resnet_dt timestep)
    act: bool = True  # tanh or linear
    resnet: bool = True


def layer_forward(layer: Layer, x: np.ndarray) -> np.ndarray:
    """mlp.py:188-219: linear -> tanh -> *idt -> residual (x or [x, x])."""
    y = x @ layer.w + layer.b
    if layer.act:
        y = np.tanh(y)
    if layer.idt is not None:
        y = y * layer.idt
    if layer.resnet:
        if x.shape[-1] == y.shape[-1]:
            y = y + x
        elif 2 * x.shape[-1] == y.shape[-1]:
            y = y + np.concatenate([x, x], axis=-1)
    return y


@dataclasses.dataclass
class SeAModel:
    ntypes: int
    sel: list
    rcut: float
    rcut_smth: float
    neuron: list
    axis_neuron: int
    fit_neuron: list
    type_one_side: bool
    davg: np.ndarray  # [ntypes, nnei, 4]
    dstd: np.ndarray  # [ntypes, nnei, 4]
    embed: list  # list (per net) of list[Layer]
    fitting: list  # list (per centre type) of list[Layer]
    bias_atom_e: np.ndarray  # [ntypes]
    tables: list | None = None  # per net [nspline, 6*M] float64
    table_info: np.ndarray | None = None  # [nnets, 6] float64

    @property
    def nnei(self) -> int:
        return int(sum(self.sel))

    @property
    def sec(self) -> list:
        out = [0]
        for s in self.sel:
            out.append(out[-1] + int(s))
        return out

    @property
    def M(self) -> int:
        return int(self.neuron[-1])

    @property
    def dim_descrpt(self) -> int:
        return self.M * self.axis_neuron

    @property
    def nnets(self) -> int:
        return self.ntypes if self.type_one_side else self.ntypes * self.ntypes

    def net_index(self, ti: int, tj: int) -> int:
        return tj if self.type_one_side else tj * self.ntypes + ti

    # ------------------------------------------------------------ builders
    @classmethod
    def random(
        cls,
        ntypes: int,
        sel,
        rcut: float,
        rcut_smth: float,
        neuron,
        axis_neuron: int,
        fit_neuron,
        seed: int = 1,
        type_one_side: bool = False,
        resnet_dt: bool = False,
    ) -> "SeAModel":
        """Seeded random weights N(0, 1/fan_in), zero biases, davg 0, dstd 1."""
        rng = np.random.default_rng(seed)
        nnei = int(sum(sel))
        nnets = ntypes if type_one_side else ntypes * ntypes

        def dense(n_in, n_out, act=True, resnet=True, dt=False):
            w = rng.normal(0.0, 1.0 / math.sqrt(n_in), size=(n_in, n_out))
            b = np.zeros(n_out)
            idt = np.full(n_out, 0.1) if dt else None
            return Layer(w, b, idt, act, resnet)

        embed = []
        for _ in range(nnets):
            layers, n_in = [], 1
            for n in neuron:
                layers.append(dense(n_in, n))
                n_in = n
            embed.append(layers)
        D = int(neuron[-1]) * axis_neuron
        fitting = []
        for _ in range(ntypes):
            layers, n_in = [], D
            for n in fit_neuron:
                layers.append(dense(n_in, n, dt=resnet_dt))
                n_in = n
            layers.append(dense(n_in, 1, act=False, resnet=False))
            fitting.append(layers)
        return cls(
            ntypes=ntypes,
            sel=list(sel),
            rcut=float(rcut),
            rcut_smth=float(rcut_smth),
            neuron=list(neuron),
            axis_neuron=int(axis_neuron),
            fit_neuron=list(fit_neuron),
            type_one_side=type_one_side,
            davg=np.zeros((ntypes, nnei, 4)),
            dstd=np.ones((ntypes, nnei, 4)),
            embed=embed,
            fitting=fitting,
            bias_atom_e=np.zeros(ntypes),
        )

    # ------------------------------------------------- input statistics
    def set_env_stat(self, frames, protection: float = 1e-2) -> None:
        """davg / dstd from the environment matrices of `frames` (list of
        (coord [n, 3], atype [n], box [9]) with orthorhombic boxes longer
        than 2 rcut), as deepmd/dpmodel/utils/env_mat_stat.py:300-423 does:
        per centre type, over every slot (empty slots count as zeros) of the
        unnormalised matrix, avg_r = mean of s/r, std_r and std_a (pooled
        over the three s x/r^2 components) with `protection` as the floor;
        davg = [avg_r, 0, 0, 0], dstd = [std_r, std_a, std_a, std_a] in
        every slot."""
        from scipy.spatial import cKDTree

        nt, nnei, sec = self.ntypes, self.nnei, self.sec
        acc = np.zeros((nt, 5))  # count, sum r, sum r^2, sum a, sum a^2
        for coord, atype, box in frames:
            box = np.asarray(box, np.float64).reshape(3, 3)
            if np.any(np.abs(box - np.diag(np.diag(box))) > 0) or np.any(np.diag(box) <= 2 * self.rcut):
                raise ValueError("set_env_stat needs orthorhombic boxes longer than 2 rcut")
            L = np.diag(box)
            x = np.mod(np.asarray(coord, np.float64), L)
            atype = np.asarray(atype)
            tree = cKDTree(x, boxsize=L)
            for i, nb in enumerate(tree.query_ball_point(x, self.rcut)):
                nb = np.array([j for j in nb if j != i], dtype=np.int64)
                rij = x[nb] - x[i]
                rij -= L * np.round(rij / L)
                r = np.linalg.norm(rij, axis=1)
                env = np.zeros((nnei, 4))
                for tj in range(nt):
                    sel = np.nonzero(atype[nb] == tj)[0]
                    sel = sel[np.argsort(r[sel], kind="stable")][: sec[tj + 1] - sec[tj]]
                    rr, dd = r[sel], rij[sel]
                    sw = np.array([_switch(v, self.rcut_smth, self.rcut) for v in rr])
                    blk = env[sec[tj]: sec[tj] + len(sel)]
                    blk[:, 0] = sw / rr
                    blk[:, 1:] = (sw / rr ** 2)[:, None] * dd
                ti = atype[i]
                acc[ti] += [nnei, env[:, 0].sum(), (env[:, 0] ** 2).sum(), env[:, 1:].sum(), (env[:, 1:] ** 2).sum()]
        davg = np.zeros((nt, nnei, 4))
        dstd = np.ones((nt, nnei, 4))
        for t in range(nt):
            n, sr, sr2, sa, sa2 = acc[t]
            if n == 0:
                dstd[t] = 0.1
                continue
            avg_r = sr / n
            std_r = max(np.sqrt(max(sr2 / n - avg_r * avg_r, 0.0)), protection)
            na = 3 * n
            std_a = max(np.sqrt(max(sa2 / na - (sa / na) ** 2, 0.0)), protection)
            davg[t, :, 0] = avg_r
            dstd[t, :, 0] = std_r
            dstd[t, :, 1:] = std_a
        self.davg, self.dstd = davg, dstd

    # ------------------------------------------------------ embedding net
    def embed_eval(self, net: int, x: np.ndarray):
        """Embedding net output and its first and second derivative w.r.t.
        the scalar input, all [len(x), M] (tanh' = 1 - y^2, tanh'' =
        -2 y (1 - y^2); residual concat carried through the derivatives).
        Same quantities as deepmd/utils/tabulate_math.py:353-470."""
        h = x.reshape(-1, 1).astype(np.float64)
        dh = np.ones_like(h)
        d2h = np.zeros_like(h)
        for layer in self.embed[net]:
            z = h @ layer.w + layer.b
            dz = dh @ layer.w
            d2z = d2h @ layer.w
            y = np.tanh(z)
            g1 = 1.0 - y * y
            g2 = -2.0 * y * g1
            yo = y
            dyo = g1 * dz
            d2yo = g2 * dz * dz + g1 * d2z
            if layer.idt is not None:
                yo, dyo, d2yo = yo * layer.idt, dyo * layer.idt, d2yo * layer.idt
            if h.shape[1] == yo.shape[1]:
                yo, dyo, d2yo = yo + h, dyo + dh, d2yo + d2h
            elif 2 * h.shape[1] == yo.shape[1]:
                cat = lambda a: np.concatenate([a, a], axis=1)  # noqa: E731
                yo, dyo, d2yo = yo + cat(h), dyo + cat(dh), d2yo + cat(d2h)
            h, dh, d2h = yo, dyo, d2yo
        return h, dh, d2h

    # -------------------------------------------------- compressed table
    def build_tables(
        self,
        min_nbor_dist: float,
        extrapolate: float = 5.0,
        stride0: float = 0.01,
        stride1: float = 0.1,
        check_frequency: float = -1.0,
        native: bool = False,
    ) -> None:
        """Quintic Hermite table per embedding net (tabulate.py:67-370);
        with native, built by the library's HIP builder
        (dpmi_tabulate_build_se_a) instead of numpy.

        Knots: arange(lower, upper, stride0) ++ arange(upper, ext*upper,
        stride1) ++ [ext*upper], with lower = floor(-davg/dstd) and upper =
        ceil((s(d_min)/d_min - davg)/dstd) per centre type (tabulate.py:468-486).
        On each interval of width t the coefficients a0..a5 are the unique
        quintic matching value, slope and curvature at both knots.
        """
        if native:
            from .tables import build_tables_native

            self.tables, self.table_info = build_tables_native(self, min_nbor_dist, extrapolate, stride0, stride1,
                                                               check_frequency)
            return
        tables, infos = [], []
        for net, (ll, uu) in enumerate(self.table_ranges(min_nbor_dist)):
            xx = np.arange(ll, uu, stride0, dtype=np.float64)
            xx = np.append(xx, np.arange(uu, extrapolate * uu, stride1, dtype=np.float64))
            xx = np.append(xx, np.array([extrapolate * uu], dtype=np.float64))
            nspline = int((uu - ll) / stride0 + (extrapolate * uu - uu) / stride1)
            vv, dd, d2 = self.embed_eval(net, xx)
            M = self.M
            tt = np.full((nspline, M), stride1)
            tt[: int((uu - ll) / stride0), :] = stride0
            y0, y1 = vv[:nspline], vv[1 : nspline + 1]
            g0, g1 = dd[:nspline], dd[1 : nspline + 1]
            c0, c1 = d2[:nspline], d2[1 : nspline + 1]
            hh = y1 - y0
            coef = np.empty((nspline, M, 6))
            coef[:, :, 0] = y0
            coef[:, :, 1] = g0
            coef[:, :, 2] = 0.5 * c0
            # the reference's expression order (tabulate.py:305-363)
            coef[:, :, 3] = (1 / (2 * tt * tt * tt)) * (20 * hh - (8 * g1 + 12 * g0) * tt - (3 * c0 - c1) * tt * tt)
            coef[:, :, 4] = (1 / (2 * tt * tt * tt * tt)) * (
                -30 * hh + (14 * g1 + 16 * g0) * tt + (3 * c0 - 2 * c1) * tt * tt)
            coef[:, :, 5] = (1 / (2 * tt * tt * tt * tt * tt)) * (
                12 * hh - 6 * (g1 + g0) * tt + (c1 - c0) * tt * tt)
            tables.append(np.ascontiguousarray(coef.reshape(nspline, 6 * M)))
            infos.append([ll, uu, uu * extrapolate, stride0, stride1, check_frequency])
        self.tables = tables
        self.table_info = np.asarray(infos, dtype=np.float64)

    def table_ranges(self, min_nbor_dist: float) -> list:
        """(lower, upper) of every net's table (tabulate.py:67-140, 468-486):
        per centre type floor(-davg/dstd) and ceil((s(d_min)/d_min - davg)/dstd)
        of the first column; one-sided nets take the extremes over the types
        with sel > 0."""
        sw = _switch(min_nbor_dist, self.rcut_smth, self.rcut)
        davg0 = self.davg[:, 0, 0]
        dstd0 = self.dstd[:, 0, 0]
        lower_t = np.floor(-davg0 / dstd0)
        upper_t = np.ceil(((1.0 / min_nbor_dist) * sw - davg0) / dstd0)
        out = []
        for net in range(self.nnets):
            if self.type_one_side:
                idx = np.asarray(self.sel) > 0
                out.append((float(np.min(lower_t[idx])), float(np.max(upper_t[idx]))))
            else:
                ti = net % self.ntypes
                out.append((float(lower_t[ti]), float(upper_t[ti])))
        return out

    def table_stack(self) -> tuple[np.ndarray, np.ndarray]:
        """All tables padded to a common row count -> [nnets, nspline, 6M]."""
        assert self.tables is not None, "build_tables() first"
        ns = max(t.shape[0] for t in self.tables)
        out = np.zeros((len(self.tables), ns, self.tables[0].shape[1]))
        for k, t in enumerate(self.tables):
            out[k, : t.shape[0]] = t
        return out, self.table_info.copy()

    # ---------------------------------------------------------- fitting
    def fitting_arrays(self):
        """Per-type fitting weights as dense arrays for the device path:
        list over layers of (W [ntypes, n_in, n_out], b [ntypes, n_out],
        idt [ntypes, n_out] or None, act, resnet)."""
        out = []
        for li in range(len(self.fitting[0])):
            ls = [self.fitting[t][li] for t in range(self.ntypes)]
            W = np.stack([l.w for l in ls])
            b = np.stack([l.b for l in ls])
            idt = None if ls[0].idt is None else np.stack([l.idt for l in ls])
            out.append((W, b, idt, ls[0].act, ls[0].resnet))
        return out


MODEL_MAGIC = b"DPMISEA1"


def save_model_file(model: SeAModel, path: str, min_nbor_dist: float, extrapolate: float = 5.0,
                    stride0: float = 0.01, stride1: float = 0.1, check_frequency: float = -1.0,
                    type_map=None) -> None:
    """Writes the model as a dpmi model file, read natively by
    dpmi_sea_model_load (which builds the compressed tables on the GPU with
    the given compression settings).  Little-endian; see include/dpmi.h."""
    import struct

    nt = model.ntypes
    emb = model.embed
    fit = model.fitting
    n_e, n_f = len(emb[0]), len(fit[0])
    out = bytearray(MODEL_MAGIC)
    out += struct.pack("<5i", nt, int(model.type_one_side), int(model.axis_neuron), n_e, n_f)
    out += struct.pack("<2d", float(model.rcut), float(model.rcut_smth))
    out += struct.pack(f"<{nt}i", *[int(v) for v in model.sel])
    out += struct.pack(f"<{n_e + 1}i", 1, *[int(l.w.shape[1]) for l in emb[0]])
    out += struct.pack(f"<{n_f + 1}i", int(fit[0][0].w.shape[0]), *[int(l.w.shape[1]) for l in fit[0]])
    out += struct.pack(f"<{n_e}i", *[int(l.idt is not None) for l in emb[0]])
    out += struct.pack(f"<{n_f}i", *[int(l.act) | (int(l.resnet) << 1) | (int(l.idt is not None) << 2)
                                     for l in fit[0]])
    out += np.ascontiguousarray(model.davg, np.float64).tobytes()
    out += np.ascontiguousarray(model.dstd, np.float64).tobytes()
    for net in list(emb) + list(fit):
        for l in net:
            out += np.ascontiguousarray(l.w, np.float64).tobytes() + np.ascontiguousarray(l.b, np.float64).tobytes()
            if l.idt is not None:
                out += np.ascontiguousarray(l.idt, np.float64).tobytes()
    out += np.ascontiguousarray(model.bias_atom_e, np.float64).tobytes()
    out += struct.pack("<5d", float(min_nbor_dist), float(extrapolate), float(stride0), float(stride1),
                       float(check_frequency))
    names = list(type_map) if type_map is not None else [f"T{t}" for t in range(nt)]
    out += struct.pack("<i", len(names))
    for nm in names:
        b = nm.encode()
        out += struct.pack("<i", len(b)) + b
    with open(path, "wb") as f:
        f.write(bytes(out))


def _switch(x: float, rmin: float, rmax: float) -> float:
    if x < rmin:
        return 1.0
    if x < rmax:
        u = (x - rmin) / (rmax - rmin)
        return u * u * u * (-6 * u * u + 15 * u - 10) + 1
    return 0.0


def water_model(seed: int = 1, min_nbor_dist: float = 0.5, type_one_side: bool = False,
                env_stat: bool = False) -> SeAModel:
    """The water se_e2_a model of the benchmark: rcut 6.0, rcut_smth 0.5,
    sel [46, 92], embedding [25, 50, 100], axis 16, fitting [240]*3, tanh,
    N(0, 1/fan_in) weights, zero biases, davg 0, dstd 1, compressed at
    strides 0.01/0.1.  With env_stat, davg/dstd come from a 768-atom water
    frame instead (as training sets them): the normalised environment
    matrix and the energies are then of order one, and the tables span
    about 2000 rows."""
    from .systems import water

    m = SeAModel.random(2, [46, 92], 6.0, 0.5, [25, 50, 100], 16, [240, 240, 240],
                        seed=seed, type_one_side=type_one_side)
    if env_stat:
        m.set_env_stat([water(256, seed=90001)])
    m.build_tables(min_nbor_dist)
    return m


def copper_model(seed: int = 1, min_nbor_dist: float = 1.0, env_stat: bool = False) -> SeAModel:
    """Single-type fcc Cu model: rcut 8.0, rcut_smth 2.0, sel [200],
    embedding [32, 64, 128], axis 16, fitting [240]*3, davg 0, dstd 1; with
    env_stat, davg/dstd from a 500-atom fcc frame."""
    from .systems import fcc_copper

    m = SeAModel.random(1, [200], 8.0, 2.0, [32, 64, 128], 16, [240, 240, 240], seed=seed)
    if env_stat:
        m.set_env_stat([fcc_copper(5, seed=90002)])
    m.build_tables(min_nbor_dist)
    return m
