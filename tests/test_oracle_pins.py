"""Pins the CPU oracle before anything is compared against it.

* golden vectors of the reference's own tests (tests/golden/, written by
  tests/golden/make_fixtures.py from the reference checkout);
* the reference CPU library compiled from its sources (oracle/_ref), where
  that build exists: bit-exact integers, floats to 1e-13;
* invariants the mathematics fixes: brute-force minimum-image neighbour sets,
  central finite differences of the environment matrix, table == direct
  embedding evaluation at the knots.
"""
import numpy as np
import pytest

from _fixtures import deeppot_sea_cases, deeppot_sea_model, lib_goldens, pt_tabulate
from dpmi import model as dmodel
from dpmi import systems
from oracle import oracle_lib, ref_available, ref_lib
from oracle.sea import evaluate

BOX13 = np.array([13.0, 0, 0, 0, 13.0, 0, 0, 0, 13.0])


@pytest.fixture(scope="module")
def O():
    return oracle_lib()


def _six_atom_env(O, g, sec, rc=6.0, rs=0.8):
    """The 6-atom fixture of source/lib/tests/test_env_mat_a.cc:24-137."""
    posi = np.array(g["posi"]).reshape(-1, 3)
    atype = np.array(g["atype"], np.int32)
    cext, text, mapping = O.copy_coord(posi, atype, BOX13, rc)
    il, nn, jl = O.build_nlist(cext, len(atype), rc, text)
    sel = np.diff(sec)
    ntypes = len(sel)
    nnei = int(sum(sel))
    avg = np.zeros((ntypes, nnei * 4))
    std = np.ones((ntypes, nnei * 4))
    em, ed, rij, nl = O.prod_env_mat_a(cext, text, il, nn, jl, avg, std, len(atype), rc, rs, sel)
    return em, ed, rij, nl, len(text)


def test_env_mat_golden(O):
    g = lib_goldens()["test_env_mat_a"]
    em, *_ = _six_atom_env(O, g, g["sec_a"])
    np.testing.assert_allclose(em.reshape(-1), g["expected_env"], atol=1e-5)


def test_env_mat_short_sel_golden(O):
    # test_env_mat_a.cc:138-: sel truncates to the nearest two per type
    g = lib_goldens()["test_env_mat_a"]
    em, *_ = _six_atom_env(O, g, g["sec_a@2"])
    np.testing.assert_allclose(em.reshape(-1), g["expected_env@2"], atol=1e-5)


def test_env_mat_deriv_finite_difference(O):
    # test_env_mat_a.cc:281-316: em_deriv is d(em)/d(centre) -> -d/d(rij)
    g = lib_goldens()["test_env_mat_a"]
    em, ed, rij, nl, nall = _six_atom_env(O, g, g["sec_a"])
    posi = np.array(g["posi"]).reshape(-1, 3)
    atype = np.array(g["atype"], np.int32)
    h = 1e-5
    nnei = nl.shape[1]
    for i in (0, 3):
        for d in range(3):
            emp, emm = [], []
            for s in (+1, -1):
                p = posi.copy()
                p[i, d] += s * h
                # keep the same neighbour list: recompute with fixed nl via the env row
                cext, text, _ = O.copy_coord(p, atype, BOX13, 6.0)
                rr = cext[np.maximum(nl[i], 0)] - cext[i]
                out = []
                for k in range(nnei):
                    if nl[i, k] < 0:
                        out.append(np.zeros(4))
                        continue
                    r = np.linalg.norm(rr[k])
                    u = np.clip((r - 0.8) / (6.0 - 0.8), 0, 1)
                    sw = 1.0 if r < 0.8 else (u ** 3 * (-6 * u * u + 15 * u - 10) + 1 if r < 6.0 else 0.0)
                    out.append(np.array([1 / r, rr[k][0] / r ** 2, rr[k][1] / r ** 2, rr[k][2] / r ** 2]) * sw)
                (emp if s > 0 else emm).append(np.concatenate(out))
            fd = (emp[0] - emm[0]) / (2 * h)
            an = ed[i].reshape(nnei * 4, 3)[:, d]
            np.testing.assert_allclose(an, fd, atol=1e-6)


def _force_fixture(O, g):
    sec = g["sec_a"]
    em, ed, rij, nl, nall = _six_atom_env(O, g, sec)
    nloc = nl.shape[0]
    return ed, rij, nl, nall, nloc


def test_prod_force_golden(O):
    g = lib_goldens()["test_prod_force_a"]
    ed, rij, nl, nall, nloc = _force_fixture(O, g)
    nd = 10 - np.arange(nloc * nl.shape[1] * 4) * 0.01
    # two identical frames (test_prod_force_a.cc:112-116)
    f = O.prod_force_a(np.tile(nd, 2), np.tile(ed.reshape(-1), 2), np.tile(nl, (2, 1)), nloc, nall, 2)
    np.testing.assert_allclose(f.reshape(-1), g["expected_force"] + g["expected_force"], atol=1e-5)


def test_prod_virial_golden(O):
    g = lib_goldens()["test_prod_virial_a"]
    ed, rij, nl, nall, nloc = _force_fixture(O, g)
    nd = 10 - np.arange(nloc * nl.shape[1] * 4) * 0.01
    v, av = O.prod_virial_a(nd, ed, rij, nl, nloc, nall)
    np.testing.assert_allclose(v, g["expected_virial"], atol=1e-5)
    np.testing.assert_allclose(av.reshape(-1), g["expected_atom_virial"], atol=1e-5)


def test_prod_force_grad_golden(O):
    g = lib_goldens()["test_prod_force_grad_a"]
    ed, rij, nl, nall, nloc = _force_fixture(O, g)
    grad = 10 - np.arange(nloc * 3) * 0.1
    # two identical frames; each reproduces the single-frame golden
    out = O.prod_force_grad_a(np.tile(grad, 2), np.tile(ed.reshape(-1), 2), np.tile(nl, (2, 1)), nloc, 2)
    np.testing.assert_allclose(out[:nloc].reshape(-1), g["expected_grad_net"], atol=1e-5)
    np.testing.assert_array_equal(out[:nloc], out[nloc:])


def test_prod_virial_grad_golden(O):
    g = lib_goldens()["test_prod_virial_grad_a"]
    ed, rij, nl, nall, nloc = _force_fixture(O, g)
    grad = 10 - np.arange(9) * 1.0
    out = O.prod_virial_grad_a(grad, ed, rij, nl, nloc)
    np.testing.assert_allclose(out.reshape(-1), g["expected_grad_net"], atol=1e-5)


@pytest.mark.parametrize("which,box,rc", [("", [13.0, 0, 0, 0, 13.0, 0, 0, 0, 13.0], 6.0),
                                          ("@2", [20.0, 0, 0, 0, 20.0, 0, 0, 0, 20.0], 4.0)])
def test_copy_coord_golden(O, which, box, rc):
    # test_coord.cc:166-257 / 395-470: ghost sets compared after sorting
    g = lib_goldens()["test_coord"]
    posi = np.array(g["posi@2" if which == "" else "posi@3"]).reshape(-1, 3)
    atype = np.array(g["atype" if which == "" else "atype@2"], np.int32)
    cext, text, mapping = O.copy_coord(posi, atype, np.array(box), rc)
    exp_c = np.array(g["_expected_posi_cpy" + which]).reshape(-1, 3)
    exp_t = np.array(g["_expected_atype_cpy" + which])
    exp_m = np.array(g["_expected_mapping" + which])
    nloc = len(atype)
    assert len(text) == len(exp_t)

    def srt(c, t, m):
        keys = [tuple(np.round(x, 9)) + (int(tt), int(mm)) for x, tt, mm in zip(c[nloc:], t[nloc:], m[nloc:])]
        return sorted(keys)

    a = srt(cext, text, mapping)
    b = srt(exp_c, exp_t, exp_m)
    for ka, kb in zip(a, b):
        np.testing.assert_allclose(ka[:3], kb[:3], atol=1e-9)
        assert ka[3:] == kb[3:]


@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-10), (np.float32, 1e-5)])
def test_tabulate_pt_golden(O, dtype, tol):
    # source/tests/pt/test_tabulate_fusion_se_a.py:24-1490
    p = pt_tabulate()
    t, info, ex, em = (p[k].astype(dtype) for k in ("table_tensor", "table_info_tensor", "em_x_tensor", "em_tensor"))
    out = O.tabulate_fusion_se_a(t, info, ex, em, p["M"])
    np.testing.assert_allclose(out, p["expected_descriptor_tensor"], atol=tol, rtol=tol)
    dy = np.ones_like(out)  # loss = descriptor.sum()
    dx, dem = O.tabulate_fusion_se_a_grad(t, info, ex, em, dy)
    np.testing.assert_allclose(dx, p["expected_dy_dem_x"], atol=tol, rtol=tol)
    np.testing.assert_allclose(dem, p["expected_dy_dem"], atol=tol, rtol=tol)


def test_tabulate_grad_grad_finite_difference(O):
    # tabulate_test_utils.py:38-92: dz/dy of the projected first gradient
    p = pt_tabulate()
    t, info, ex, em = (p[k] for k in ("table_tensor", "table_info_tensor", "em_x_tensor", "em_tensor"))
    M = p["M"]
    rng = np.random.default_rng(3)
    wx = rng.normal(size=ex.shape)
    wem = rng.normal(size=em.shape)
    dz = O.tabulate_fusion_se_a_grad_grad(t, info, ex, em, wx, wem, M)
    dy0 = rng.normal(size=(4, 4, M))
    # linear in dy: projection(dy) = sum(grad_x(dy) * wx) + sum(grad_em(dy) * wem)
    def proj(dy):
        dx, dem = O.tabulate_fusion_se_a_grad(t, info, ex, em, dy)
        return np.sum(dx * wx) + np.sum(dem * wem)
    eps = 1e-6
    fd = np.zeros_like(dy0)
    for idx in np.ndindex(dy0.shape):
        a = dy0.copy(); a[idx] += eps
        b = dy0.copy(); b[idx] -= eps
        fd[idx] = (proj(a) - proj(b)) / (2 * eps)
    np.testing.assert_allclose(dz, fd, atol=1e-6, rtol=1e-6)


def test_model_level_golden():
    """The glue (descriptor contraction, fitting, net_deriv, force, virial)
    with the embedding net evaluated directly reproduces the reference's
    model-level goldens (source/tests/infer/deeppot-testcase.yaml)."""
    m = deeppot_sea_model()
    O = oracle_lib()
    for c in deeppot_sea_cases():
        if c["box"] is None:
            continue
        coord = np.array(c["coord"]).reshape(-1, 3)
        at = np.array(c["atype"], np.int32)
        r = evaluate(m, coord, at, np.array(c["box"]), O, embedding="direct")
        np.testing.assert_allclose(r.atom_energy, c["atomic_energy"], atol=1e-10, rtol=0)
        np.testing.assert_allclose(r.force.reshape(-1), c["force"], atol=1e-10, rtol=0)
        np.testing.assert_allclose(r.virial, np.array(c["atomic_virial"]).reshape(-1, 9).sum(0), atol=1e-10)
        if "descriptor" in c:
            np.testing.assert_allclose(r.descriptor.reshape(-1), c["descriptor"], atol=1e-10)


def test_compressed_equals_direct():
    """Compressed (table) E/F/V equal the direct embedding evaluation to 1e-10
    (test_model_compression_se_a.py:167-186 asserts places=10)."""
    m = deeppot_sea_model()
    m.build_tables(0.5)
    O = oracle_lib()
    c = deeppot_sea_cases()[0]
    coord = np.array(c["coord"]).reshape(-1, 3)
    at = np.array(c["atype"], np.int32)
    a = evaluate(m, coord, at, np.array(c["box"]), O, embedding="direct")
    b = evaluate(m, coord, at, np.array(c["box"]), O, embedding="table")
    assert abs(a.energy - b.energy) < 1e-10
    np.testing.assert_allclose(a.force, b.force, atol=1e-10)
    np.testing.assert_allclose(a.virial, b.virial, atol=1e-10)


def test_table_hermite_conditions():
    """At every knot the table reproduces the embedding value, slope and
    curvature (the defining conditions of the quintic Hermite table)."""
    m = dmodel.water_model(seed=2)
    net = 1
    tab = m.tables[net].reshape(m.tables[net].shape[0], m.M, 6)
    info = m.table_info[net]
    xs = np.concatenate([np.arange(info[0], info[1], info[3]), np.arange(info[1], info[2], info[4])])[: tab.shape[0]]
    v, d1, d2 = m.embed_eval(net, xs)
    np.testing.assert_allclose(tab[:, :, 0], v, atol=1e-14)
    np.testing.assert_allclose(tab[:, :, 1], d1, atol=1e-12)
    np.testing.assert_allclose(2 * tab[:, :, 2], d2, atol=1e-10)
    # right end of interval k = left end of k+1
    t = np.where(np.arange(tab.shape[0]) < int(round((info[1] - info[0]) / info[3])), info[3], info[4])[:, None]
    right = sum(tab[:, :, p] * t ** p for p in range(6))
    np.testing.assert_allclose(right[:-1], v[1:], atol=1e-10)


def test_nlist_bruteforce_minimum_image(O):
    """Cell-list ghosts + formatted rows contain exactly the atoms within rc
    found by an O(N^2) minimum-image search (box > 2 rc)."""
    coord, atype, box = systems.water(64, seed=5)
    L = box[0]
    cn = O.normalize_coord(coord, box)
    cext, text, mapping = O.copy_coord(cn, atype, box, 6.0)
    il, nn, jl = O.build_nlist(cext, len(atype), 6.0, text)
    for i in range(0, len(atype), 7):
        d = cn - cn[i]
        d -= L * np.round(d / L)
        r2 = np.sum(d * d, axis=1)
        expect = set(np.nonzero((r2 < 36.0) & (np.arange(len(atype)) != i))[0].tolist())
        got = mapping[jl[i, : nn[i]]].tolist()
        assert len(got) == len(set(got))
        assert set(got) == expect


needs_ref = pytest.mark.skipif(not ref_available(), reason="reference library not built here")


@needs_ref
@pytest.mark.parametrize("nmol,seed", [(64, 1), (64, 7), (512, 3)])
def test_oracle_matches_reference_lib(nmol, seed):
    O, R = oracle_lib(), ref_lib()
    coord, atype, box = systems.water(nmol, seed=seed)
    a, b = O.normalize_coord(coord, box), R.normalize_coord(coord, box)
    assert np.array_equal(a, b)
    ea, eb = O.copy_coord(a, atype, box, 6.0), R.copy_coord(b, atype, box, 6.0)
    for x, y in zip(ea, eb):
        assert np.array_equal(x, y)
    nloc = len(atype)
    na, nb = O.build_nlist(ea[0], nloc, 6.0, ea[1]), R.build_nlist(eb[0], nloc, 6.0, eb[1])
    for x, y in zip(na, nb):
        assert np.array_equal(x, y)
    m = dmodel.water_model(seed=seed)
    avg = np.random.default_rng(seed).normal(0, 0.05, size=(2, m.nnei * 4))
    std = np.random.default_rng(seed + 1).uniform(0.5, 1.5, size=(2, m.nnei * 4))
    pa = O.prod_env_mat_a(ea[0], ea[1], *na, avg, std, nloc, 6.0, 0.5, m.sel)
    pb = R.prod_env_mat_a(eb[0], eb[1], *nb, avg, std, nloc, 6.0, 0.5, m.sel)
    assert np.array_equal(pa[3], pb[3])
    for x, y in zip(pa[:3], pb[:3]):
        np.testing.assert_allclose(x, y, rtol=1e-13, atol=1e-15)
    # tabulate fwd / grad / grad_grad on the water blocks
    em = pa[0].reshape(nloc, m.nnei, 4)
    rng = np.random.default_rng(seed)
    for tj in range(2):
        rr = np.ascontiguousarray(em[:, m.sec[tj]:m.sec[tj + 1]])
        ss = np.ascontiguousarray(rr[:, :, 0])
        t, info = m.tables[m.net_index(0, tj)], m.table_info[m.net_index(0, tj)]
        np.testing.assert_allclose(O.tabulate_fusion_se_a(t, info, ss, rr, m.M),
                                   R.tabulate_fusion_se_a(t, info, ss, rr, m.M), rtol=1e-13, atol=1e-13)
        dy = rng.normal(size=(nloc, 4, m.M))
        for x, y in zip(O.tabulate_fusion_se_a_grad(t, info, ss, rr, dy), R.tabulate_fusion_se_a_grad(t, info, ss, rr, dy)):
            np.testing.assert_allclose(x, y, rtol=1e-12, atol=1e-12)
        zx, zem = rng.normal(size=ss.shape), rng.normal(size=rr.shape)
        np.testing.assert_allclose(O.tabulate_fusion_se_a_grad_grad(t, info, ss, rr, zx, zem, m.M),
                                   R.tabulate_fusion_se_a_grad_grad(t, info, ss, rr, zx, zem, m.M),
                                   rtol=1e-12, atol=1e-12)
    nd = rng.normal(size=(nloc, m.nnei * 4))
    nall = len(ea[1])
    np.testing.assert_allclose(O.prod_force_a(nd, pa[1], pa[3], nloc, nall), R.prod_force_a(nd, pb[1], pb[3], nloc, nall),
                               rtol=1e-12, atol=1e-12)
    for x, y in zip(O.prod_virial_a(nd, pa[1], pa[2], pa[3], nloc, nall), R.prod_virial_a(nd, pb[1], pb[2], pb[3], nloc, nall)):
        np.testing.assert_allclose(x, y, rtol=1e-11, atol=1e-11)
    g = rng.normal(size=(nloc * 3))
    np.testing.assert_allclose(O.prod_force_grad_a(g, pa[1], pa[3], nloc), R.prod_force_grad_a(g, pb[1], pb[3], nloc),
                               rtol=1e-12, atol=1e-12)
    g9 = rng.normal(size=9)
    np.testing.assert_allclose(O.prod_virial_grad_a(g9, pa[1], pa[2], pa[3], nloc),
                               R.prod_virial_grad_a(g9, pb[1], pb[2], pb[3], nloc), rtol=1e-12, atol=1e-12)


@needs_ref
@pytest.mark.parametrize("box", [np.diag([12.42, 12.42, 12.42]).reshape(9),
                                 np.array([12.0, 0.0, 0.0, 2.5, 11.0, 0.0, 1.5, -2.0, 13.0]),
                                 np.diag([49.68, 30.0, 9.0]).reshape(9)])
@pytest.mark.parametrize("rc", [6.0, 8.0, 2.5])
def test_cell_info_matches_reference_lib(box, rc):
    """compute_cell_info (coord.cc:69-108): the 23 ints of the cell grid."""
    O, R = oracle_lib(), ref_lib()
    assert np.array_equal(O.compute_cell_info(rc, box), R.compute_cell_info(rc, box))


@needs_ref
@pytest.mark.parametrize("sel", [[46, 92], [10, 20], [3, 5]])
def test_format_nlist_matches_reference_lib(sel):
    """format_nlist_cpu (fmt_nlist.cc:99-177) on shuffled rc + skin candidate
    rows (as an MD code hands them over), with per-type truncation."""
    O, R = oracle_lib(), ref_lib()
    coord, atype, box = systems.water(64, seed=4)
    cn = O.normalize_coord(coord, box)
    cext, text, _ = O.copy_coord(cn, atype, box, 6.0)
    nloc = len(atype)
    il, nn, jl = O.build_nlist(cext, nloc, 6.5, text)  # rcut + skin
    rng = np.random.default_rng(1)
    jl = jl.copy()
    for i in range(nloc):
        jl[i, : nn[i]] = rng.permutation(jl[i, : nn[i]])
    a = O.format_nlist(il, nn, jl, cext, text, nloc, 6.0, sel)
    b = R.format_nlist(il, nn, jl, cext, text, nloc, 6.0, sel)
    assert np.array_equal(a, b)
    if sel == [3, 5]:
        assert np.all(a >= 0)  # every row truncated at sel
