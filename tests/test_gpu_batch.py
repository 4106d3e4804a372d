"""Batched uneven frames: hipGraph capture/replay and the autotuner.

* A captured step replayed on changed coordinates equals an eager step of a
  second engine on the same coordinates, replay after replay (eager work and
  synchronisations in between).
* Every tuned launch shape gives the same physics; the tuned knobs come from
  the candidate sets and are reused by a later engine of the same size
  bucket.
"""
import numpy as np
import pytest
import torch

from dpmi import model as dmodel
from dpmi import systems
from dpmi.engine import DeviceModel, Engine
from dpmi.lib import DpmiError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def water():
    return dmodel.water_model(seed=1)


@pytest.fixture(scope="module")
def batch():
    frames = systems.uneven_water_frames(8, seed=5, lo=96, hi=900)
    natoms = [len(a) for (_, a, _) in frames]
    coord = np.concatenate([c for (c, _, _) in frames])
    atype = np.concatenate([a for (_, a, _) in frames])
    box = np.stack([b for (_, _, b) in frames])
    return natoms, coord, atype, box


def _close(a, b, tol):
    a, b = np.asarray(a).ravel(), np.asarray(b).ravel()
    return np.max(np.abs(a - b)) <= tol * max(np.max(np.abs(b)), 1e-30)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_graph_replay_matches_eager(water, batch, precision):
    natoms, coord, atype, box = batch
    dm = DeviceModel(water)
    eg = Engine(dm, natoms, atype, precision=precision)
    ee = Engine(dm, natoms, atype, precision=precision)
    x = torch.as_tensor(coord, device="cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        eg.capture(x, box)
    s.synchronize()
    gen = torch.Generator(device="cuda").manual_seed(3)
    # replays and eager steps differ only in the order of FP64 atomic adds
    tol_e, tol_fv = (1e-12, 1e-11) if precision == "fp64" else (1e-10, 1e-9)
    for it in range(3):
        print(f"replay {precision} {it}", flush=True)
        with torch.cuda.stream(s):
            x.add_(0.02 * torch.randn(x.shape, dtype=x.dtype, device=x.device, generator=gen))
            e, f, v = [t.clone() for t in eg.replay()]
        s.synchronize()
        eg.status()
        e1, f1, v1 = ee.compute(x.clone(), box)
        torch.cuda.synchronize()
        ee.status()
        assert _close(e.cpu().numpy(), e1.cpu().numpy(), tol_e)
        assert _close(f.cpu().numpy(), f1.cpu().numpy(), tol_fv)
        assert _close(v.cpu().numpy(), v1.cpu().numpy(), tol_fv)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_autotune_keeps_results_and_caches(water, batch, precision):
    natoms, coord, atype, box = batch
    dm = DeviceModel(water)
    x = torch.as_tensor(coord, device="cuda")
    ref = Engine(dm, natoms, atype, precision=precision)
    e0, f0, v0 = [t.cpu().numpy().copy() for t in ref.compute(x, box)]
    eng = Engine(dm, natoms, atype, precision=precision)
    knobs = eng.autotune(x, box, reps=2)
    print(precision, knobs, flush=True)
    assert knobs["env_threads_per_block"] in (128, 256, 512)
    assert knobs["embed_fwd_waves"] in (1, 2, 4) and knobs["embed_bwd_waves"] in (1, 2, 4)
    assert knobs["force_rows_per_workgroup"] in (8, 32, 64)
    # forward: per-neighbour (0), table space (1) or row runs (8); adjoint:
    # per-neighbour (0), table space (2) or row runs (4)
    mixed_masks = {f | b for f in (0, 1, 8) for b in (0, 2, 4)}
    assert knobs["embed_matrix_cores"] in (mixed_masks if precision == "mixed" else {0})
    # fused fitting tiles of 32 rows, 64 rows with 16 waves (+512), or 32-row
    # tiles with layer 0 in its own kernels (+256)
    assert knobs["fitting_rows_per_tile"] in ((32, 64 + 512, 32 + 256) if precision == "mixed" else (0,))
    e, f, v = [t.cpu().numpy().copy() for t in eng.compute(x, box)]
    if precision == "fp64" or knobs["embed_matrix_cores"] == 0:
        # launch shapes alone do not change the arithmetic
        tol = 1e-12 if precision == "fp64" else 1e-9
        assert _close(e, e0, tol) and _close(f, f0, tol) and _close(v, v0, tol)
    else:
        # other embedding kernels round differently: FP32-level agreement
        # with the default kernels, and bit-equal to an untuned engine given
        # the same knobs
        assert _close(e, e0, 1e-6) and _close(f, f0, 1e-4) and _close(v, v0, 1e-4)
        same = Engine(dm, natoms, atype, precision=precision)
        same.set_tuning(**knobs)
        e2, f2, v2 = [t.cpu().numpy().copy() for t in same.compute(x, box)]
        assert _close(e, e2, 1e-12) and _close(f, f2, 1e-12) and _close(v, v2, 1e-12)
    later = Engine(dm, natoms, atype, precision=precision)
    assert later.tuning() == knobs


def _grid(box, rc=6.0):
    """Cells per dimension of cubic boxes [nf, 9] (face distance / rcut)."""
    return np.floor(box[:, [0, 4, 8]] / rc).astype(int)


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_graph_device_box_follows_rescaled_boxes(water, batch, precision):
    """NPT-style replays: coordinates and boxes rescaled in device memory
    between replays; each replay equals an eager step on the new box.  A box
    that needs another cell grid is reported (the step keeps the old box)."""
    natoms, coord, atype, box = batch
    dm = DeviceModel(water)
    eg = Engine(dm, natoms, atype, precision=precision)
    ee = Engine(dm, natoms, atype, precision=precision)
    x = torch.as_tensor(coord, device="cuda")
    b = torch.as_tensor(np.asarray(box, np.float64).reshape(-1, 9), device="cuda").contiguous()
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        eg.capture_device_box(x, b)
    s.synchronize()
    tol_e, tol_fv = (1e-12, 1e-11) if precision == "fp64" else (1e-10, 1e-9)
    g0 = _grid(b.cpu().numpy())
    done = 0
    for scale in (1.0015, 0.9985, 1.003, 0.997, 1.001):
        nb = b.cpu().numpy() * scale
        if not np.array_equal(_grid(nb), g0):
            continue  # this scale would cross a cell-count boundary of some frame
        print(f"device-box replay {precision} scale {scale}", flush=True)
        with torch.cuda.stream(s):
            x.mul_(scale)
            b.mul_(scale)
            e, f, v = [t.clone() for t in eg.replay()]
        s.synchronize()
        eg.status()
        e1, f1, v1 = ee.compute(x.clone(), b.cpu().numpy())
        torch.cuda.synchronize()
        ee.status()
        assert _close(e.cpu().numpy(), e1.cpu().numpy(), tol_e)
        assert _close(f.cpu().numpy(), f1.cpu().numpy(), tol_fv)
        assert _close(v.cpu().numpy(), v1.cpu().numpy(), tol_fv)
        g0 = _grid(b.cpu().numpy())
        done += 1
    assert done >= 2
    # a box with more cells per dimension: reported, nothing out of bounds
    with torch.cuda.stream(s):
        x.mul_(1.6)
        b.mul_(1.6)
        eg.replay()
    s.synchronize()
    with pytest.raises(DpmiError, match="ARGUMENT"):
        eg.status()
