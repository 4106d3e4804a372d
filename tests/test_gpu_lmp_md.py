"""The LAMMPS caller: NVE dynamics through the pair style
dpmi::lmp::PairDeepMD (include/dpmi_pair_deepmd.h), driven by
tests/cpp/md_driver.cpp, a single-process LAMMPS stand-in (ghost images
within cutoff + skin, full rcut + skin neighbour lists in arbitrary order,
Neighbor::decide re-neighbouring with ago, newton-on reverse communication,
velocity Verlet in metal units).  Reference: source/lmp/pair_deepmd.cpp:253-632.

Checks:
* at dump steps, both right after a re-neighbouring (ago = 0) and on a reused
  list (ago > 0), the owner-folded forces, the energy, the 6-component virial,
  the per-atom energies and the centroid per-atom virials equal the oracle's
  evaluation of the same local coordinates (FP64 1e-10; mixed within the
  accuracy budget);
* NVE: the total energy's deviation scales as dt^2 (velocity Verlet on forces
  that are the exact gradient of the energy) and stays far below the kinetic
  energy's own fluctuation;
* shuffled rows and special-bond bits above NEIGHMASK change nothing;
* a triclinic cell with boxlo != 0, and LAMMPS types mapped to NULL (absent
  atoms);
* the pair style's error contract.
"""
import os
import subprocess

import numpy as np
import pytest

from dpmi import model as dmodel
from dpmi import systems
from dpmi._build import MD_DRIVER
from oracle import oracle_lib
from oracle.sea import evaluate

pytestmark = pytest.mark.gpu

KB = 8.617343e-5
MVV2E = 1.0364269e-4
MASS = {0: 15.9994, 1: 1.00794}
ORDER = [0, 4, 8, 3, 6, 7, 1, 2, 5]  # LAMMPS centroid order of a row-major 3x3


def _w(d, name, a, dt):
    np.ascontiguousarray(a, dtype=dt).tofile(os.path.join(d, name))


def _r(d, name, dt=np.float64):
    return np.fromfile(os.path.join(d, name), dtype=dt)


@pytest.fixture(scope="module")
def water_md_model(tmp_path_factory):
    """The benchmark water architecture with davg/dstd from environment
    statistics (order-one energies and forces of ~0.1 eV/A)."""
    d = tmp_path_factory.mktemp("mdmodel")
    m = dmodel.water_model(seed=7, env_stat=True)
    path = str(d / "water.dpmi")
    dmodel.save_model_file(m, path, 0.5, type_map=["O", "H"])
    return m, path


def _system(d, coord, ltype, box, boxlo, temp, seed, masses):
    rng = np.random.default_rng(seed)
    m = np.array([masses[t - 1] for t in ltype])
    sig = np.sqrt(KB * temp / (m * MVV2E))
    vel = rng.normal(size=(len(ltype), 3)) * sig[:, None]
    vel -= np.sum(vel * m[:, None], axis=0) / np.sum(m)
    _w(d, "box.f64", box, np.float64)
    _w(d, "boxlo.f64", boxlo, np.float64)
    _w(d, "coord.f64", coord, np.float64)
    _w(d, "vel.f64", vel, np.float64)
    _w(d, "type.i32", ltype, np.int32)
    _w(d, "mass.f64", masses, np.float64)


def _run(d, model_path, precision, steps, dt, skin, dump_every, *opts):
    args = [MD_DRIVER, str(d), model_path, precision, str(steps), repr(dt), repr(skin), str(dump_every)] + list(opts)
    print("case md", precision, steps, dt, skin, opts, flush=True)
    assert os.path.exists(MD_DRIVER), "build the MD driver with __graft_entry__.build()"
    r = subprocess.run(args, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    return _r(d, "thermo.f64").reshape(-1, 12)


def _check_dump(d, step, m, box, boxlo, mtype, thermo, precision, atom=False, keep=None):
    """The pair style's folded outputs at a dump step against the oracle on
    the same local coordinates (in the pair style's frame: x - boxlo)."""
    x = _r(d, f"dump{step}_x.f64").reshape(-1, 3)
    f = _r(d, f"dump{step}_f.f64").reshape(-1, 3)
    keep = np.arange(len(mtype)) if keep is None else keep
    ref = evaluate(m, x[keep] - boxlo, mtype[keep], box, oracle_lib(), keep=True)
    row = thermo[step]
    n = len(keep)
    vref = ref.virial.reshape(9)[[0, 4, 8, 3, 6, 7]]
    if precision == "fp64":
        assert abs(row[1] - ref.energy) <= 1e-10 * abs(ref.energy)
        assert np.max(np.abs(f[keep] - ref.force)) <= 1e-10 * np.max(np.abs(ref.force))
        assert np.max(np.abs(row[6:12] - vref)) <= 1e-10 * np.max(np.abs(vref))
    else:
        assert abs(row[1] - ref.energy) / n <= 1e-4
        assert np.sqrt(np.mean((f[keep] - ref.force) ** 2)) <= 1e-3
        assert np.max(np.abs(row[6:12] - vref)) / n <= 1e-3
    absent = np.setdiff1d(np.arange(len(mtype)), keep)
    assert np.all(f[absent] == 0)
    if atom:
        ea = _r(d, f"dump{step}_eatom.f64")
        cv = _r(d, f"dump{step}_cvatom.f64").reshape(-1, 9)
        tol = 1e-10 if precision == "fp64" else 1e-3
        np.testing.assert_allclose(ea[keep], ref.atom_energy, atol=tol * max(1.0, np.max(np.abs(ref.atom_energy))),
                                   rtol=0)
        # oracle per-atom virials live on extended atoms; fold them onto owners
        av = np.zeros((n, 9))
        np.add.at(av, ref.extra["mapping"], ref.atom_virial_ext.reshape(-1, 9))
        scale = max(1.0, np.max(np.abs(av)))
        np.testing.assert_allclose(cv[keep], av[:, ORDER], atol=tol * scale, rtol=0)
        assert np.allclose(cv.sum(0)[:6], row[6:12], atol=1e-9 * scale * n)


def _nve_error(thermo):
    etot = thermo[:, 3]
    return np.max(np.abs(etot - etot[0]))


@pytest.fixture(scope="module")
def water64(tmp_path_factory):
    coord, atype, box = systems.water(64, seed=21)
    return coord, atype, box


def test_md_fp64_dumps_and_energy_conservation(tmp_path, water_md_model, water64):
    m, path = water_md_model
    coord, atype, box = water64
    boxlo = np.zeros(3)
    _system(tmp_path, coord, atype + 1, box, boxlo, 300.0, 1, [MASS[0], MASS[1]])
    th = _run(tmp_path, path, "fp64", 200, 0.0005, 1.0, 50, "atom")
    builds = int(th[-1, 5])
    assert builds >= 3, f"only {builds} neighbour-list builds"
    # dumps on reused lists (ago > 0) as well as fresh ones
    assert np.any(th[[50, 100, 150, 200], 4] > 0)
    for step in (0, 50, 100, 150, 200):
        _check_dump(tmp_path, step, m, box, boxlo, atype, th, "fp64", atom=True)
    # NVE on the same potential: over the same physical time, halving the step
    # cuts velocity Verlet's energy error by four when the forces are the
    # exact gradient of the energy.  The random-weight model has stiff
    # regions, so the asymptotic regime needs 0.1 / 0.05 fs steps; a 0.3 A
    # skin keeps the lists being rebuilt and reused.
    errs = []
    for k, (nstep, dt) in enumerate(((200, 1e-4), (400, 5e-5))):
        dk = tmp_path / f"nve{k}"
        dk.mkdir()
        _system(dk, coord, atype + 1, box, boxlo, 300.0, 1, [MASS[0], MASS[1]])
        thk = _run(dk, path, "fp64", nstep, dt, 0.3, 0)
        assert thk[-1, 5] >= 2
        errs.append((_nve_error(thk), np.ptp(thk[:, 2]), thk))
    (e1, ke1, t1), (e2, _, t2) = errs
    print(f"NVE max |dE| {e1:.3e} (0.1 fs) {e2:.3e} (0.05 fs), KE range {ke1:.3e} eV", flush=True)
    assert 3.0 < e1 / e2 < 5.5
    assert e1 < 0.02 * ke1
    # the two runs follow the same trajectory to O(dt^2)
    assert abs(t1[-1, 1] - t2[-1, 1]) < 0.01 * np.ptp(t1[:, 1])


def test_md_shuffled_rows_and_special_bits(tmp_path, water_md_model, water64):
    """Rows in random order with LAMMPS special-bond bits on every third
    entry give the same trajectory as plain rows."""
    m, path = water_md_model
    coord, atype, box = water64
    runs = []
    for opts in ((), ("shuffle", "special")):
        d = tmp_path / ("_".join(opts) or "plain")
        d.mkdir()
        _system(d, coord, atype + 1, box, np.zeros(3), 300.0, 2, [MASS[0], MASS[1]])
        runs.append((d, _run(d, path, "fp64", 60, 0.0005, 1.0, 30, *opts)))
    (d0, t0), (d1, t1) = runs
    np.testing.assert_allclose(t1[:, 1], t0[:, 1], rtol=1e-11, atol=0)
    f0, f1 = _r(d0, "dump60_f.f64"), _r(d1, "dump60_f.f64")
    assert np.max(np.abs(f1 - f0)) <= 1e-9 * np.max(np.abs(f0))
    _check_dump(d1, 60, m, box, np.zeros(3), atype, t1, "fp64")


def test_md_triclinic_boxlo_and_null_types(tmp_path, water_md_model):
    """A sheared cell with boxlo != 0, and a third LAMMPS type mapped to NULL
    (absent atoms: no neighbours, no energy, zero force)."""
    m, path = water_md_model
    coord, atype, box = systems.water(96, seed=33)
    L = box[0]
    boxm = np.array([[L, 0, 0], [2.1, L, 0], [-1.3, 1.7, L]])
    boxlo = np.array([-3.0, 1.5, 0.25])
    rng = np.random.default_rng(4)
    frac = np.mod(coord / L, 1.0)
    x = frac @ boxm + boxlo
    nnull = 17
    xn = rng.uniform(0, 1, size=(nnull, 3)) @ boxm + boxlo
    allx = np.concatenate([x, xn])
    ltype = np.concatenate([atype + 1, np.full(nnull, 3, np.int32)])
    _system(tmp_path, allx, ltype, boxm.reshape(9), boxlo, 300.0, 5, [MASS[0], MASS[1], 4.0])
    th = _run(tmp_path, path, "fp64", 40, 0.0005, 0.6, 20, "names", "O", "H", "NULL", "atom")
    keep = np.arange(len(atype))
    mtype = np.concatenate([atype, np.full(nnull, -1, np.int32)])
    for step in (0, 20, 40):
        _check_dump(tmp_path, step, m, boxm.reshape(9), boxlo, mtype, th, "fp64", atom=True, keep=keep)


def test_md_mixed_within_budget(tmp_path, water_md_model, water64):
    m, path = water_md_model
    coord, atype, box = water64
    _system(tmp_path, coord, atype + 1, box, np.zeros(3), 300.0, 1, [MASS[0], MASS[1]])
    th = _run(tmp_path, path, "mixed", 200, 0.0005, 1.0, 100, "atom")
    for step in (0, 100, 200):
        _check_dump(tmp_path, step, m, box, np.zeros(3), atype, th, "mixed", atom=True)
    # NVE at 0.1 fs: the mixed path conserves energy about as well as FP64
    # (its forces are the gradient of its energy to FP32-level rounding)
    errs = {}
    for prec in ("fp64", "mixed"):
        dk = tmp_path / f"nve_{prec}"
        dk.mkdir()
        _system(dk, coord, atype + 1, box, np.zeros(3), 300.0, 1, [MASS[0], MASS[1]])
        errs[prec] = _nve_error(_run(dk, path, prec, 200, 1e-4, 0.3, 0))
    print(f"NVE max |dE| at 0.1 fs: fp64 {errs['fp64']:.3e} mixed {errs['mixed']:.3e} eV", flush=True)
    assert errs["mixed"] < 1.5 * errs["fp64"]


def test_pair_style_error_contract(tmp_path, water_md_model):
    _, path = water_md_model
    for name, a, dt in (("box.f64", np.eye(3).reshape(9) * 20, np.float64), ("boxlo.f64", np.zeros(3), np.float64),
                        ("coord.f64", np.zeros((1, 3)), np.float64), ("vel.f64", np.zeros((1, 3)), np.float64),
                        ("type.i32", [1], np.int32), ("mass.f64", [1.0, 1.0], np.float64)):
        _w(tmp_path, name, a, dt)
    print("case md errors", flush=True)
    r = subprocess.run([MD_DRIVER, str(tmp_path), path, "fp64", "0", "0.001", "1.0", "0", "errors"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    # pair_style without a model, pair_coeff before pair_style, pair_coeff 1 2,
    # an unknown element, a keyword without a value, a 6-component per-atom
    # virial, a LAMMPS type beyond pair_coeff -- then a valid call
    assert list(_r(tmp_path, "errors.i32", np.int32)) == [1, 1, 1, 1, 1, 1, 1, 0]
