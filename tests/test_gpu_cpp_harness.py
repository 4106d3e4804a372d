"""The C++ drop-in surface, called from C++ (tests/cpp/harness.cpp, built by
__graft_entry__.build()) the way the reference's callers call it:

* lib: every deepmd:: template of include/deepmd_lib.h on a water frame --
  normalize_coord_gpu, compute_cell_info, copy_coord_gpu (device Region and
  cell_info), build_nlist_gpu (device InputNlist), convert_nlist_gpu_device +
  prod_env_mat_a_gpu / format_nbor_list_gpu, env_mat_nbor_update (the
  mesh-encoded host list), use_nlist_map, prod_force/virial(_grad)_a_gpu,
  tabulate_fusion_se_a{,_grad,_grad_grad}_gpu; the *_cpu host-array entry
  points on the same inputs; the float instantiations of the geometry chain;
  each against the oracle, the reference's op goldens, or the device result;
* deeppot: dpmi::DeepPot loading a model file of the reference's se_e2_a test
  model (tables built natively) against source/tests/infer/deeppot-testcase.yaml
  (periodic and non-periodic blocks, per-atom energies and virials), the
  float interface, and a LAMMPS-style call (caller ghosts, unsorted rcut +
  skin rows, ago = 0 then a reused list on moved atoms).
"""
import os
import subprocess

import numpy as np
import pytest

from _fixtures import deeppot_sea_cases, deeppot_sea_model, pt_tabulate
from dpmi import model as dmodel
from dpmi import systems
from dpmi._build import HARNESS
from oracle import oracle_lib
from oracle.sea import evaluate

pytestmark = pytest.mark.gpu


def _w(d, name, a, dt):
    np.ascontiguousarray(a, dtype=dt).tofile(os.path.join(d, name))


def _r(d, name, dt):
    return np.fromfile(os.path.join(d, name), dtype=dt)


def _run(mode, d):
    print(f"case harness {mode}", flush=True)
    assert os.path.exists(HARNESS), "build the harness with __graft_entry__.build()"
    r = subprocess.run([HARNESS, mode, str(d)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr


def rel(a, b):
    a, b = np.asarray(a, np.float64).reshape(-1), np.asarray(b, np.float64).reshape(-1)
    assert a.shape == b.shape
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def lib_run(tmp_path_factory):
    d = tmp_path_factory.mktemp("lib")
    m = dmodel.water_model(seed=1)
    coord, atype, box = systems.water(64, seed=11)
    nloc, nnei = len(atype), m.nnei
    rng = np.random.default_rng(4)
    avg = rng.normal(0, 0.05, size=(2, nnei * 4))
    std = rng.uniform(0.5, 1.5, size=(2, nnei * 4))
    _w(d, "params.i32", [nloc, 2] + list(m.sel), np.int32)
    _w(d, "params.f64", [m.rcut, m.rcut_smth], np.float64)
    _w(d, "coord.f64", coord, np.float64)
    _w(d, "atype.i32", atype, np.int32)
    _w(d, "box.f64", box, np.float64)
    _w(d, "avg.f64", avg, np.float64)
    _w(d, "std.f64", std, np.float64)
    nd = rng.normal(size=(nloc, nnei * 4))
    gf, gv = rng.normal(size=nloc * 3), rng.normal(size=9)
    _w(d, "nd.f64", nd, np.float64)
    _w(d, "gforce.f64", gf, np.float64)
    _w(d, "gvir.f64", gv, np.float64)
    p = pt_tabulate()
    dy = rng.normal(size=p["expected_descriptor_tensor"].shape)
    _w(d, "tab_dims.i32", [p["em_tensor"].shape[0], p["em_tensor"].shape[1], p["M"]], np.int32)
    _w(d, "tab_table.f64", p["table_tensor"], np.float64)
    _w(d, "tab_info.f64", p["table_info_tensor"], np.float64)
    _w(d, "tab_emx.f64", p["em_x_tensor"], np.float64)
    _w(d, "tab_em.f64", p["em_tensor"], np.float64)
    _w(d, "tab_dy.f64", dy, np.float64)
    _run("lib", d)
    return d, dict(m=m, coord=coord, atype=atype, box=box, avg=avg, std=std, nd=nd, gf=gf, gv=gv, p=p, dy=dy)


def test_lib_geometry_and_environment(lib_run):
    d, x = lib_run
    O = oracle_lib()
    m, nloc = x["m"], len(x["atype"])
    cn = O.normalize_coord(x["coord"], x["box"])
    assert np.array_equal(_r(d, "norm.f64", np.float64), cn.reshape(-1))
    assert np.array_equal(_r(d, "cellinfo.i32", np.int32), O.compute_cell_info(m.rcut, x["box"]))
    cext, text, mapping = O.copy_coord(cn, x["atype"], x["box"], m.rcut)
    assert np.array_equal(_r(d, "ext_coord.f64", np.float64), cext.reshape(-1))
    assert np.array_equal(_r(d, "ext_type.i32", np.int32), text)
    assert np.array_equal(_r(d, "mapping.i32", np.int32), mapping)
    il, nn, jl = O.build_nlist(cext, nloc, m.rcut, text)
    assert np.array_equal(_r(d, "bn_numneigh.i32", np.int32), nn)
    assert np.array_equal(_r(d, "bn_rows.i32", np.int32), np.concatenate([jl[i, : nn[i]] for i in range(nloc)]))
    em, ed, rij, nlist = O.prod_env_mat_a(cext, text, il, nn, jl, x["avg"], x["std"], nloc, m.rcut, m.rcut_smth,
                                          m.sel)
    assert np.array_equal(_r(d, "nlist.i32", np.int32), nlist.reshape(-1))
    assert np.array_equal(_r(d, "fmt_nlist.i32", np.int32), nlist.reshape(-1))
    assert rel(_r(d, "em.f64", np.float64), em) < 1e-12
    assert rel(_r(d, "em_deriv.f64", np.float64), ed) < 1e-12
    assert rel(_r(d, "rij.f64", np.float64), rij) < 1e-12
    # the mesh-encoded host list gives the same environment matrix
    assert np.array_equal(_r(d, "mesh_em.f64", np.float64), _r(d, "em.f64", np.float64))
    assert int(_r(d, "mesh_max_nbor.i32", np.int32)[0]) == 256
    nl = nlist.reshape(-1)
    assert np.array_equal(_r(d, "mapped_nlist.i32", np.int32), np.where(nl >= 0, mapping[np.maximum(nl, 0)], -1))


def test_lib_force_virial_and_tabulate(lib_run):
    d, x = lib_run
    O = oracle_lib()
    nloc = len(x["atype"])
    ed, rij = _r(d, "em_deriv.f64", np.float64), _r(d, "rij.f64", np.float64)
    nlist = _r(d, "nlist.i32", np.int32)
    nall = len(_r(d, "mapping.i32", np.int32))
    nd = x["nd"].reshape(nloc, -1)
    f_ref = O.prod_force_a(nd, ed.reshape(nloc, -1), nlist.reshape(nloc, -1), nloc, nall)[0]
    assert rel(_r(d, "force.f64", np.float64), f_ref) < 1e-12
    v_ref, av_ref = O.prod_virial_a(nd, ed.reshape(nloc, -1), rij.reshape(nloc, -1), nlist.reshape(nloc, -1), nloc,
                                    nall)
    assert rel(_r(d, "virial.f64", np.float64), v_ref) < 1e-12
    assert rel(_r(d, "atom_virial.f64", np.float64), av_ref) < 1e-12
    g_ref = O.prod_force_grad_a(x["gf"], ed.reshape(nloc, -1), nlist.reshape(nloc, -1), nloc)
    assert rel(_r(d, "gnet_f.f64", np.float64), g_ref) < 1e-12
    g_ref = O.prod_virial_grad_a(x["gv"], ed.reshape(nloc, -1), rij.reshape(nloc, -1), nlist.reshape(nloc, -1), nloc)
    assert rel(_r(d, "gnet_v.f64", np.float64), g_ref) < 1e-12
    p = x["p"]
    np.testing.assert_allclose(_r(d, "tab_out.f64", np.float64), p["expected_descriptor_tensor"].reshape(-1),
                               atol=1e-10, rtol=1e-10)
    dx, dem = _r(d, "tab_dx.f64", np.float64), _r(d, "tab_dem.f64", np.float64)
    ex, emt = p["em_x_tensor"], p["em_tensor"]
    rx, rem = O.tabulate_fusion_se_a_grad(p["table_tensor"].reshape(-1), p["table_info_tensor"], ex, emt, x["dy"])
    assert rel(dx, rx) < 1e-12 and rel(dem, rem) < 1e-12
    rz = O.tabulate_fusion_se_a_grad_grad(p["table_tensor"].reshape(-1), p["table_info_tensor"], ex, emt,
                                          dx.reshape(ex.shape), dem.reshape(emt.shape), p["M"])
    assert rel(_r(d, "tab_dz.f64", np.float64), rz) < 1e-12


def test_lib_host_entry_points(lib_run):
    """The *_cpu functions (host arrays) return the device results."""
    d, _ = lib_run
    for name, dt in (("norm.f64", np.float64), ("fmt_nlist.i32", np.int32), ("nlist.i32", np.int32),
                     ("em.f64", np.float64), ("tab_out.f64", np.float64), ("tab_dx.f64", np.float64),
                     ("tab_dem.f64", np.float64), ("tab_dz.f64", np.float64), ("gnet_f.f64", np.float64),
                     ("gnet_v.f64", np.float64), ("virial.f64", np.float64), ("bn_numneigh.i32", np.int32)):
        assert np.array_equal(_r(d, "cpu_" + name, dt), _r(d, name, dt)), name
    assert np.array_equal(_r(d, "cpu_mapping.i32", np.int32), _r(d, "mapping.i32", np.int32))
    assert np.array_equal(_r(d, "cpu_ext_coord.f64", np.float64), _r(d, "ext_coord.f64", np.float64))
    assert rel(_r(d, "cpu_force.f64", np.float64), _r(d, "force.f64", np.float64)) < 1e-13


def test_lib_float_instantiations(lib_run):
    """normalize / copy_coord / build_nlist / prod_env_mat_a / format in FP32:
    coordinates within FP32 rounding of the FP64 chain, identical ghost
    mapping and rows (pairs within FP32 rounding of the cutoff aside), and the
    environment matrix within FP32 accuracy."""
    d, _ = lib_run
    assert rel(_r(d, "f32_norm.f32", np.float32), _r(d, "norm.f64", np.float64)) < 1e-6
    assert np.array_equal(_r(d, "f32_mapping.i32", np.int32), _r(d, "mapping.i32", np.int32))
    assert rel(_r(d, "f32_ext_coord.f32", np.float32), _r(d, "ext_coord.f64", np.float64)) < 1e-6
    nn32, nn64 = _r(d, "f32_bn_numneigh.i32", np.int32), _r(d, "bn_numneigh.i32", np.int32)
    assert np.mean(nn32 == nn64) > 0.98
    nl32, nl64 = _r(d, "f32_nlist.i32", np.int32), _r(d, "nlist.i32", np.int32)
    nnei = 138
    same = np.all(nl32.reshape(-1, nnei) == nl64.reshape(-1, nnei), axis=1)
    assert np.mean(same) > 0.98
    assert np.array_equal(_r(d, "f32_fmt_nlist.i32", np.int32), nl32)
    e32 = _r(d, "f32_em.f32", np.float32).reshape(-1, nnei * 4)[same]
    e64 = _r(d, "em.f64", np.float64).reshape(-1, nnei * 4)[same]
    assert np.max(np.abs(e32 - e64)) < 1e-5 * np.max(np.abs(e64))


@pytest.mark.parametrize("precision", [0, 1])
def test_deeppot_reference_goldens(tmp_path, precision):
    d = tmp_path
    m = deeppot_sea_model()
    dmodel.save_model_file(m, str(d / "model.dpmi"), 0.5, type_map=["O", "H"])
    cases = deeppot_sea_cases()
    pbc = next(c for c in cases if c["box"] is not None)
    nopbc = next(c for c in cases if c["box"] is None)
    coord = np.array(pbc["coord"]).reshape(-1, 3)
    at = np.array(pbc["atype"], np.int32)
    box = np.array(pbc["box"])
    _w(d, "dp_precision.i32", [precision], np.int32)
    _w(d, "pbc_coord.f64", coord, np.float64)
    _w(d, "pbc_atype.i32", at, np.int32)
    _w(d, "pbc_box.f64", box, np.float64)
    _w(d, "nopbc_coord.f64", nopbc["coord"], np.float64)
    _w(d, "nopbc_atype.i32", nopbc["atype"], np.int32)
    # LAMMPS-style input: ghosts within rcut + skin, shuffled rcut + skin rows
    O = oracle_lib()
    cn = O.normalize_coord(coord, box)
    cext, text, mapping = O.copy_coord(cn, at, box, m.rcut + 1.0)
    nloc = len(at)
    il, nn, jl = O.build_nlist(cext, nloc, m.rcut + 1.0, text)
    rng = np.random.default_rng(3)
    rows = np.concatenate([rng.permutation(jl[i, : nn[i]]) for i in range(nloc)])
    delta = rng.normal(0, 0.03, size=(nloc, 3))
    _w(d, "lmp_coord.f64", cext, np.float64)
    _w(d, "lmp_coord2.f64", cext + delta[mapping], np.float64)
    _w(d, "lmp_atype.i32", text, np.int32)
    _w(d, "lmp_nghost.i32", [len(text) - nloc], np.int32)
    _w(d, "lmp_ilist.i32", il, np.int32)
    _w(d, "lmp_numneigh.i32", nn, np.int32)
    _w(d, "lmp_rows.i32", rows, np.int32)
    _run("deeppot", d)
    assert abs(_r(d, "dp_cutoff.f64", np.float64)[0] - 6.0) < 1e-12
    assert int(_r(d, "dp_ntypes.i32", np.int32)[0]) == 2
    assert bytes(_r(d, "dp_type_map.i32", np.int32).astype(np.uint8)).decode() == "O H"
    # FP64: the goldens (1e-9, as the model-level golden test).  Mixed: the
    # accuracy budget per frame (energy <= 0.1 meV/atom, force RMSE <= 1
    # meV/A, virial <= 1 meV/atom); on this small model (fitting [10, 10, 10],
    # energies dominated by the -93 / -186 eV bias) per-atom energies move by
    # up to 0.2 meV, so they are held to 1 meV.
    tol = 1e-9 if precision == 0 else 1e-3
    for pre, c in (("pbc", pbc), ("nopbc", nopbc)):
        n = len(c["atype"])
        np.testing.assert_allclose(_r(d, f"{pre}_ae.f64", np.float64), c["atomic_energy"], atol=tol, rtol=0)
        f = _r(d, f"{pre}_f.f64", np.float64)
        if precision == 0:
            np.testing.assert_allclose(f, c["force"], atol=tol, rtol=0)
            np.testing.assert_allclose(_r(d, f"{pre}_av.f64", np.float64), c["atomic_virial"], atol=tol, rtol=0)
        else:
            assert np.sqrt(np.mean((f - np.array(c["force"])) ** 2)) <= 1e-3
        np.testing.assert_allclose(_r(d, f"{pre}_v.f64", np.float64),
                                   np.array(c["atomic_virial"]).reshape(-1, 9).sum(0), atol=tol * n, rtol=0)
        assert abs(_r(d, f"{pre}_e.f64", np.float64)[0] - np.sum(c["atomic_energy"])) <= (tol if precision == 0
                                                                                            else 1e-4 * n)
    np.testing.assert_allclose(_r(d, "pbcf_f.f32", np.float32), pbc["force"], atol=max(tol, 1e-6), rtol=0)
    # LAMMPS style: extended outputs, folded by the caller
    nall = len(text)

    def fold(a, k):
        out = np.zeros((nloc, k))
        np.add.at(out, mapping, a.reshape(nall, k))
        return out

    np.testing.assert_allclose(fold(_r(d, "lmp_f.f64", np.float64), 3).reshape(-1), pbc["force"], atol=tol, rtol=0)
    if precision == 0:
        np.testing.assert_allclose(fold(_r(d, "lmp_av.f64", np.float64), 9).reshape(-1), pbc["atomic_virial"],
                                   atol=tol, rtol=0)
    ae = _r(d, "lmp_ae.f64", np.float64)
    np.testing.assert_allclose(ae[:nloc], pbc["atomic_energy"], atol=tol, rtol=0)
    assert np.all(ae[nloc:] == 0)
    # the reused list on moved atoms: the oracle on the moved frame
    m.build_tables(0.5)
    ref = evaluate(m, coord + delta, at, box, oracle_lib())
    e2 = _r(d, "lmp2_e.f64", np.float64)[0]
    f2 = fold(_r(d, "lmp2_f.f64", np.float64), 3)
    if precision == 0:
        assert abs(e2 - ref.energy) <= 1e-10 * abs(ref.energy)
        assert rel(f2, ref.force) < 1e-10
    else:
        assert abs(e2 - ref.energy) <= 1e-4 * nloc and np.sqrt(np.mean((f2 - ref.force) ** 2)) <= 1e-3
