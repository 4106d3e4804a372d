"""Mixed-precision engine (DPMI_MIXED).

* Shared FP64 geometry: ghosts and neighbour lists equal the FP64 engine's
  (which the whole-path tests pin to the oracle) bit for bit.
* Environment matrix: evaluated in FP64 and stored in FP32, so it equals the
  FP64 engine's values rounded to FP32 (within one FP32 ulp).
* Fitting: the atomic energies equal a numpy evaluation of the same
  arithmetic (FP16 descriptor and weights, FP32/FP64 accumulate, FP16
  hidden activations) on the engine's own descriptor, which separates the
  FP16 MFMA GEMMs from the table/contraction stages.
* End to end against FP64: energy error <= 0.1 meV/atom, force RMSE <= 1
  meV/A, virial error <= 1 meV/atom (per-component max |dV| / natoms).
"""
import numpy as np
import pytest
import torch

from dpmi import model as dmodel
from dpmi import systems
from dpmi.engine import DeviceModel, Engine

pytestmark = pytest.mark.gpu

E_TOL_MEV_ATOM = 0.1
F_RMSE_TOL_MEV_A = 1.0
V_TOL_MEV_ATOM = 1.0


@pytest.fixture(scope="module")
def water():
    return dmodel.water_model(seed=1)


@pytest.fixture(scope="module")
def water_dev(water):
    return DeviceModel(water)


def _run(dm, natoms, coord, atype, box, precision):
    eng = Engine(dm, natoms, atype, precision=precision)
    e, f, v = eng.compute(torch.as_tensor(coord, device="cuda"), box, atom_energy=True)
    eng.status()
    return eng, e.cpu().numpy().copy(), f.cpu().numpy().copy(), v.cpu().numpy().copy()


def errors(e_m, f_m, v_m, e_d, f_d, v_d, natoms):
    return {
        "energy_meV_atom": float(abs(e_m - e_d) / natoms * 1e3),
        "force_rmse_meV_A": float(np.sqrt(np.mean((f_m - f_d) ** 2)) * 1e3),
        "virial_meV_atom": float(np.max(np.abs(v_m - v_d)) / natoms * 1e3),
    }


def _fp16(a):
    return np.asarray(a, np.float64).astype(np.float16).astype(np.float64)


def _split16(a):
    """FP16 hi + lo parts of the weights, as the mixed fitting multiplies them."""
    a = np.asarray(a, np.float64)
    hi = _fp16(a)
    return hi + _fp16(a - hi)


def fitting_reference(m, desc16, atype):
    """Per-atom energy of the mixed fitting arithmetic: x_0 = FP16
    descriptor, hidden layer y = tanh(x (W16hi + W16lo) + b), x' = y * idt + x
    (equal widths), stored in FP16 except the last hidden output, which feeds
    the output layer in FP32."""
    fa = m.fitting_arrays()
    H = len(fa) - 1
    out = np.zeros(len(atype))
    for t in range(m.ntypes):
        rows = np.nonzero(atype == t)[0]
        if rows.size == 0:
            continue
        x = desc16[rows]
        for li in range(H):
            W, b, idt, act, res = fa[li]
            z = x @ _split16(W[t]) + b[t].astype(np.float32)
            y = np.tanh(z) if act else z
            o = y * idt[t].astype(np.float32) if idt is not None else y
            if res and W.shape[1] == W.shape[2]:
                o = o + x
            x = o if li == H - 1 else _fp16(o)
        W, b, _idt, _act, _res = fa[H]
        out[rows] = x @ W[t][:, 0].astype(np.float32) + b[t][0] + m.bias_atom_e[t]
    return out


@pytest.mark.parametrize("nmol,seed", [(64, 1), (512, 7)])
def test_mixed_water_against_fp64(water, water_dev, nmol, seed):
    _check_water_against_fp64(water, water_dev, nmol, seed)


def test_mixed_water_with_data_statistics_against_fp64():
    """davg/dstd from the environment statistics of a water frame: energies
    of order one (the random model with davg 0 / dstd 1 gives ~1e-4
    eV/atom), tables of ~2000 rows."""
    m = dmodel.water_model(seed=1, env_stat=True)
    _check_water_against_fp64(m, DeviceModel(m), 512, 7)


def _check_water_against_fp64(water, water_dev, nmol, seed):
    print(f"case mixed water nmol={nmol} seed={seed}", flush=True)
    coord, atype, box = systems.water(nmol, seed=seed)
    n = len(atype)
    ed, e_d, f_d, v_d = _run(water_dev, [n], coord, atype, box, "fp64")
    em, e_m, f_m, v_m = _run(water_dev, [n], coord, atype, box, "mixed")
    # shared FP64 geometry
    assert np.array_equal(em.fetch("coord_ext"), ed.fetch("coord_ext"))
    assert np.array_equal(em.fetch("nlist"), ed.fetch("nlist"))
    # env matrix: FP64 values rounded once to FP32
    a, b = em.fetch("em"), ed.fetch("em")
    assert np.all(np.abs(a - b.astype(np.float32).astype(np.float64)) <= 2.0 ** -23 * np.abs(b) + 1e-30)
    # descriptor in FP16 (FP32 table and contraction)
    dm, dd = em.fetch("descriptor"), ed.fetch("descriptor")
    assert np.max(np.abs(dm - dd)) <= 2e-3 * np.max(np.abs(dd))
    # fitting GEMMs against the same arithmetic in numpy
    ref_ae = fitting_reference(water, dm.reshape(n, -1), np.asarray(atype))
    ae = em.fetch("atom_energy")
    assert np.max(np.abs(ae - ref_ae)) <= 1e-3 * np.max(np.abs(ref_ae - np.asarray(water.bias_atom_e)[atype])) + 1e-6
    err = errors(e_m[0], f_m, v_m[0], e_d[0], f_d, v_d[0], n)
    print("errors vs fp64:", err, flush=True)
    assert err["energy_meV_atom"] <= E_TOL_MEV_ATOM
    assert err["force_rmse_meV_A"] <= F_RMSE_TOL_MEV_A
    assert err["virial_meV_atom"] <= V_TOL_MEV_ATOM


def test_mixed_batched_uneven_frames(water, water_dev):
    frames = systems.uneven_water_frames(6, seed=11, lo=96, hi=900)
    natoms = [len(a) for (_, a, _) in frames]
    coord = np.concatenate([c for (c, _, _) in frames])
    atype = np.concatenate([a for (_, a, _) in frames])
    box = np.stack([b for (_, _, b) in frames])
    _, e_d, f_d, v_d = _run(water_dev, natoms, coord, atype, box, "fp64")
    _, e_m, f_m, v_m = _run(water_dev, natoms, coord, atype, box, "mixed")
    off = 0
    for k, n in enumerate(natoms):
        err = errors(e_m[k], f_m[off:off + n], v_m[k], e_d[k], f_d[off:off + n], v_d[k], n)
        assert err["energy_meV_atom"] <= E_TOL_MEV_ATOM, (k, err)
        assert err["force_rmse_meV_A"] <= F_RMSE_TOL_MEV_A, (k, err)
        assert err["virial_meV_atom"] <= V_TOL_MEV_ATOM, (k, err)
        off += n


@pytest.mark.parametrize("env_stat", [False, True])
def test_mixed_copper_small(env_stat):
    m = dmodel.copper_model(seed=2, env_stat=env_stat)
    coord, atype, box = systems.fcc_copper(5, seed=4)
    dm = DeviceModel(m)
    n = len(atype)
    _, e_d, f_d, v_d = _run(dm, [n], coord, atype, box, "fp64")
    _, e_m, f_m, v_m = _run(dm, [n], coord, atype, box, "mixed")
    err = errors(e_m[0], f_m, v_m[0], e_d[0], f_d, v_d[0], n)
    print("copper errors vs fp64:", err, flush=True)
    assert err["energy_meV_atom"] <= E_TOL_MEV_ATOM
    assert err["force_rmse_meV_A"] <= F_RMSE_TOL_MEV_A
    assert err["virial_meV_atom"] <= V_TOL_MEV_ATOM


def test_mixed_repeat_is_stable(water, water_dev):
    """Buffers reused across calls: a second call on the same input gives the
    same result up to the order of FP32 atomic accumulation."""
    coord, atype, box = systems.water(64, seed=3)
    eng = Engine(water_dev, [len(atype)], atype, precision="mixed")
    x = torch.as_tensor(coord, device="cuda")
    e1, f1, v1 = [t.cpu().numpy().copy() for t in eng.compute(x, box)]
    e2, f2, v2 = [t.cpu().numpy().copy() for t in eng.compute(x, box)]
    assert abs(e1[0] - e2[0]) <= 1e-5 * abs(e1[0])
    assert np.max(np.abs(f1 - f2)) <= 1e-5 * np.max(np.abs(f1))


def _perturbed(m, seed):
    """Non-zero fitting biases and atom-energy biases (the random models
    start with zeros)."""
    rng = np.random.default_rng(seed)
    for layers in m.fitting:
        for layer in layers:
            layer.b = rng.normal(0.0, 0.1, size=layer.b.shape)
    m.bias_atom_e = np.array([-93.5, -187.2][: m.ntypes])
    return m


@pytest.mark.parametrize("fit_neuron,resnet_dt,one_side", [
    ([240, 240, 240], True, False),   # fused kernel, timestep layers
    ([64, 96], False, True),          # fused kernel, unequal widths (no residual)
    ([320, 320], True, False),        # wider than the fused kernel: per-layer GEMMs
    ([100, 100, 100, 100], False, False),  # fused kernel, four hidden layers, padded widths
])
def test_mixed_fitting_variants(fit_neuron, resnet_dt, one_side):
    print(f"case fit={fit_neuron} resnet_dt={resnet_dt} one_side={one_side}", flush=True)
    m = dmodel.SeAModel.random(2, [46, 92], 6.0, 0.5, [25, 50, 100], 16, fit_neuron, seed=5,
                               type_one_side=one_side, resnet_dt=resnet_dt)
    m = _perturbed(m, 6)
    m.build_tables(0.5)
    coord, atype, box = systems.water(64, seed=12)
    n = len(atype)
    dm = DeviceModel(m)
    ed, e_d, f_d, v_d = _run(dm, [n], coord, atype, box, "fp64")
    em, e_m, f_m, v_m = _run(dm, [n], coord, atype, box, "mixed")
    ref_ae = fitting_reference(m, em.fetch("descriptor").reshape(n, -1), np.asarray(atype))
    ae = em.fetch("atom_energy")
    scale = np.max(np.abs(ref_ae - np.asarray(m.bias_atom_e)[atype]))
    assert np.max(np.abs(ae - ref_ae)) <= 1e-3 * scale + 1e-6
    err = errors(e_m[0], f_m, v_m[0], e_d[0], f_d, v_d[0], n)
    print("errors vs fp64:", err, flush=True)
    assert err["energy_meV_atom"] <= E_TOL_MEV_ATOM
    assert err["force_rmse_meV_A"] <= F_RMSE_TOL_MEV_A
    assert err["virial_meV_atom"] <= V_TOL_MEV_ATOM


@pytest.mark.parametrize("neuron,axis", [
    ([16, 32, 48], 8),     # M = 48: one 64-lane column group, 3 columns per lane of the 16-lane adjoint
    ([32, 64, 64], 16),    # M = 64: column groups exactly full
    ([25, 50, 100], 4),    # narrow descriptor axis
    ([40, 80, 160], 16),   # M = 160: beyond the 16-lane adjoint and the matrix-core kernels
])
def test_mixed_embedding_widths(neuron, axis):
    """Embedding widths and descriptor axes other than the benchmark's cover
    the other column layouts of the table kernels (lanes over 1, 2 or 3
    column groups; 3, 4, 7 columns per lane of the 16-lane adjoint; the
    lanes-over-columns adjoint for M > 128)."""
    print(f"case neuron={neuron} axis={axis}", flush=True)
    m = dmodel.SeAModel.random(2, [46, 92], 6.0, 0.5, neuron, axis, [64, 64], seed=7)
    m = _perturbed(m, 8)
    m.build_tables(0.5)
    coord, atype, box = systems.water(64, seed=13)
    n = len(atype)
    dm = DeviceModel(m)
    ed, e_d, f_d, v_d = _run(dm, [n], coord, atype, box, "fp64")
    # per-neighbour kernels, and (M <= 128) the row-run forward and adjoint
    # on the matrix cores (7 - 3 column blocks of 16, A < 16 included)
    for mask in ((0, 12) if neuron[-1] <= 128 else (0,)):
        em = Engine(dm, [n], atype, precision="mixed")
        em.set_tuning(embed_matrix_cores=mask)
        e_m, f_m, v_m = [t.cpu().numpy().copy() for t in em.compute(torch.as_tensor(coord, device="cuda"), box)]
        em.status()
        for what, tol in (("descriptor", 2e-3), ("net_deriv", 5e-3)):
            a, b = em.fetch(what), ed.fetch(what)
            rel = float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
            print("mask", mask, what, "relative deviation from fp64", rel, flush=True)
            assert rel < tol, (mask, what, rel)
        err = errors(e_m[0], f_m, v_m[0], e_d[0], f_d, v_d[0], n)
        print("mask", mask, "errors vs fp64:", err, flush=True)
        assert err["energy_meV_atom"] <= E_TOL_MEV_ATOM
        assert err["force_rmse_meV_A"] <= F_RMSE_TOL_MEV_A
        assert err["virial_meV_atom"] <= V_TOL_MEV_ATOM


def test_mixed_rejects_width_doubling_residual():
    """A residual layer of doubled width (concatenated input) is an FP64-only
    feature: the mixed engine refuses the model instead of mis-evaluating."""
    from dpmi.lib import DpmiError

    m = dmodel.SeAModel.random(2, [46, 92], 6.0, 0.5, [25, 50, 100], 16, [64, 128], seed=5)
    m.build_tables(0.5)
    coord, atype, box = systems.water(8, seed=1)
    dm = DeviceModel(m)
    _run(dm, [len(atype)], coord, atype, box, "fp64")
    with pytest.raises(DpmiError, match="UNSUPPORTED"):
        Engine(dm, [len(atype)], atype, precision="mixed")


@pytest.mark.parametrize("nmol", [100, 1500])
def test_fitting_tile_rows_give_identical_rows(water_dev, nmol):
    """Fused fitting tiles of 32 and 64 rows, and layer 0 in its own kernels
    (+256), compute every row with the same arithmetic in the same order, so
    atomic energies, net_deriv, energy and virial agree bit for bit; 16 waves
    per 64-row tile (+512) only regroup the energy's FP32 sums."""
    coord, atype, box = systems.water(nmol, seed=13)
    n = len(atype)
    out = {}
    for rows in (32, 64, 32 + 256, 64 + 256, 64 + 512):
        eng = Engine(water_dev, [n], atype, precision="mixed")
        eng.set_tuning(fitting_rows_per_tile=rows)
        assert eng.tuning()["fitting_rows_per_tile"] == rows
        e, f, v = eng.compute(torch.as_tensor(coord, device="cuda"), box, atom_energy=True)
        eng.status()
        out[rows] = (eng.fetch("atom_energy"), eng.fetch("net_deriv"), e.cpu().numpy(), v.cpu().numpy())
    for rows in (64, 32 + 256, 64 + 256):
        for a, b in zip(out[32], out[rows]):
            assert np.array_equal(a, b), rows
    # 16 waves per 64-row tile: the same products (net_deriv and the virial,
    # which follows from it, bit for bit); the energy's FP32 partial sums over
    # the column tiles are grouped by wave, so atom energies agree to FP32
    # rounding
    ae, nd, e, v = out[64 + 512]
    assert np.array_equal(nd, out[32][1])
    assert np.max(np.abs(ae - out[32][0])) <= 1e-6 * np.max(np.abs(out[32][0]))
    assert abs(e[0] - out[32][2][0]) <= 1e-6 * abs(out[32][2][0])
