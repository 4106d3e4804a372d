"""Table embedding of the mixed engine on the matrix cores.

The table-space kernels (embedding mask 3: forward and adjoint on the
matrix cores) and the per-neighbour FP32 kernels (mask 0) evaluate the same
sums in a different order and operand precision (split FP16 hi + lo
products with FP32 accumulation against FP32 quintics).  On the same
geometry their descriptor, net_deriv, energies, forces and virials agree to
FP32-level tolerances, and every mask stays within the mixed-precision
error budget of the FP64 engine.

Cases: water (two types, padding tails in both neighbour blocks, tiles of
fewer than 4 atoms), fcc copper (one type, long runs of equal table rows),
both with davg 0 / dstd 1 and with data-derived davg/dstd (energies of
order one, fine tables), and a water model with non-zero davg (non-zero
padding rows, so the sorted-padding tail carries weight) whose tables end
below the closest neighbours (extrapolated neighbours above the table
maximum).
"""
import numpy as np
import pytest
import torch

from dpmi import model as dmodel
from dpmi import systems
from dpmi.engine import DeviceModel, Engine

pytestmark = pytest.mark.gpu


def _eval(dm, natoms, coord, atype, box, precision="mixed", mask=None):
    eng = Engine(dm, natoms, atype, precision=precision)
    if mask is not None:
        eng.set_tuning(embed_matrix_cores=mask)
        assert eng.tuning()["embed_matrix_cores"] == mask
    e, f, v = eng.compute(torch.as_tensor(coord, device="cuda"), box, atom_energy=True)
    eng.status()
    out = {"e": e.cpu().numpy().copy(), "f": f.cpu().numpy().copy(), "v": v.cpu().numpy().copy(),
           "ae": eng.fetch("atom_energy"), "nd": eng.fetch("net_deriv"), "desc": eng.fetch("descriptor")}
    del eng
    return out


def _rel(a, b):
    a, b = np.asarray(a, np.float64).ravel(), np.asarray(b, np.float64).ravel()
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _special_water():
    """Non-zero davg/dstd and tables that end at 1.2 x upper, with upper set
    by a 1.2 A minimum distance, so O-H neighbours lie beyond the table."""
    m = dmodel.SeAModel.random(2, [46, 92], 6.0, 0.5, [25, 50, 100], 16, [240, 240, 240], seed=5)
    rng = np.random.default_rng(11)
    m.davg = np.broadcast_to(np.array([0.3, 0.05, -0.04, 0.02]), m.davg.shape).copy()
    m.davg[1] *= 1.5
    m.dstd = np.broadcast_to(np.array([0.25, 0.4, 0.4, 0.4]), m.dstd.shape).copy()
    m.dstd *= rng.uniform(0.9, 1.1, size=(m.ntypes, 1, 1))
    m.build_tables(1.2, extrapolate=1.2)
    return m


CASES = {
    "water63": lambda: (dmodel.water_model(seed=1), systems.water(63, seed=3)),
    "water63_stat": lambda: (dmodel.water_model(seed=1, env_stat=True), systems.water(63, seed=3)),
    "copper5": lambda: (dmodel.copper_model(seed=2), systems.fcc_copper(5, seed=4)),
    "copper5_stat": lambda: (dmodel.copper_model(seed=2, env_stat=True), systems.fcc_copper(5, seed=4)),
    "water_tail_extrapolated": lambda: (_special_water(), systems.water(40, seed=8)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_matrix_core_embedding_matches_per_neighbour_kernels(case):
    """Both masks against the FP64 engine: the table-space kernels are no
    less accurate than the per-neighbour FP32 kernels (the largest net_deriv
    deviation of either sits at a sorted-padding tail slot, whose weight
    nb - jt multiplies the FP32 rounding of the slot's gradient), and the two
    agree with each other to that level."""
    print(f"case {case}", flush=True)
    m, (coord, atype, box) = CASES[case]()
    dm = DeviceModel(m)
    n = len(atype)
    d = _eval(dm, [n], coord, atype, box, precision="fp64")
    ref = _eval(dm, [n], coord, atype, box, mask=0)
    for mask in (1, 2, 3, 4, 5, 8, 12):
        got = _eval(dm, [n], coord, atype, box, mask=mask)
        e_ref = {k: _rel(ref[k], d[k]) for k in ("e", "f", "v", "nd", "desc")}
        e_got = {k: _rel(got[k], d[k]) for k in ("e", "f", "v", "nd", "desc")}
        print(case, mask, "vs fp64: per-neighbour", e_ref, "table-space", e_got, flush=True)
        # both errors are dominated by FP16 rounding of the descriptor, so
        # they are compared as magnitudes, not digit for digit; the floor is
        # a tenth of that rounding (2^-11 relative), below which either
        # error can cancel by chance
        for k in e_ref:
            assert e_got[k] <= 3.0 * e_ref[k] + 5e-5, (mask, k, e_got[k], e_ref[k])
        assert _rel(got["nd"], ref["nd"]) < 5e-3
        assert _rel(got["desc"], ref["desc"]) < 2e-3  # FP16 descriptor
        assert _rel(got["f"], ref["f"]) < 2e-2 and _rel(got["e"], ref["e"]) < 1e-3


@pytest.mark.parametrize("case", list(CASES))
def test_matrix_core_embedding_within_mixed_budget(case):
    m, (coord, atype, box) = CASES[case]()
    dm = DeviceModel(m)
    n = len(atype)
    d = _eval(dm, [n], coord, atype, box, precision="fp64")
    g = _eval(dm, [n], coord, atype, box, mask=3)
    e_err = abs(g["e"][0] - d["e"][0]) / n * 1e3
    f_rmse = float(np.sqrt(np.mean((g["f"] - d["f"]) ** 2)) * 1e3)
    v_err = float(np.max(np.abs(g["v"][0] - d["v"][0])) / n * 1e3)
    print(case, "vs fp64: energy meV/atom", e_err, "force rmse meV/A", f_rmse, "virial meV/atom", v_err, flush=True)
    assert e_err <= 0.1 and f_rmse <= 1.0 and v_err <= 1.0


def test_tuning_knob_rejects_bad_mask():
    from dpmi.lib import DpmiError

    m = dmodel.water_model(seed=1)
    coord, atype, box = systems.water(8, seed=1)
    eng = Engine(DeviceModel(m), [len(atype)], atype, precision="mixed")
    with pytest.raises(DpmiError, match="ARGUMENT"):
        eng.set_tuning(embed_matrix_cores=16)
