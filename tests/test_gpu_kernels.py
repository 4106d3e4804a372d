"""Kernel-level parity: each dpmi_* entry point against the oracle on the same
seeded inputs.  Integer outputs (ghost order, neighbour lists) bit-exact;
FP64 within 1e-10 relative (the tolerance the reference's FP64 op tests use,
source/tests/pt/test_tabulate_fusion_se_a.py:24-29); FP32 within 1e-5."""
import numpy as np
import pytest
import torch

from _fixtures import pt_tabulate
from dpmi import model as dmodel
from dpmi import ops, systems
from oracle import oracle_lib

pytestmark = pytest.mark.gpu
DEV = "cuda"


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = max(np.max(np.abs(b)), 1e-300)
    return float(np.max(np.abs(a - b)) / den) if a.size else 0.0


def T(a, dtype=None):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=DEV)


TRICLINIC = np.array([12.6, 0.0, 0.0, 1.3, 12.1, 0.0, -0.7, 0.9, 13.2])


@pytest.fixture(scope="module")
def O():
    return oracle_lib()


@pytest.mark.parametrize("nmol,box", [(64, None), (4096, None), (64, TRICLINIC), (32, None)])
def test_normalize_and_copy_coord_bitexact(O, nmol, box):
    coord, atype, b = systems.water(nmol, seed=11)
    if box is not None:
        b = box
    print(f"case nmol={nmol} triclinic={box is not None}", flush=True)
    c = T(coord)
    ops.normalize_coord(c, b)
    ref = O.normalize_coord(coord, b)
    assert np.array_equal(c.cpu().numpy(), ref)
    ce, te, me = ops.copy_coord(c, T(atype, torch.int32), b, 6.0)
    rce, rte, rme = O.copy_coord(ref, atype, b, 6.0)
    assert ce.shape[0] == rce.shape[0]
    assert np.array_equal(ce.cpu().numpy(), rce)
    assert np.array_equal(te.cpu().numpy(), rte)
    assert np.array_equal(me.cpu().numpy(), rme)


def test_copy_coord_capacity_error():
    coord, atype, b = systems.water(64, seed=1)
    c = T(O_norm := oracle_lib().normalize_coord(coord, b))
    with pytest.raises(RuntimeError, match="MEMORY"):
        ops.copy_coord(c, T(atype, torch.int32), b, 6.0, mem_nall=500)


def test_build_nlist_and_format_bitexact(O):
    coord, atype, b = systems.water(64, seed=3)
    cn = O.normalize_coord(coord, b)
    cext, text, _ = O.copy_coord(cn, atype, b, 6.0)
    nloc = len(atype)
    il, nn, jl, mx = ops.build_nlist(T(cext), nloc, 6.0, T(text, torch.int32), mem_size=128)
    ril, rnn, rjl = O.build_nlist(cext, nloc, 6.0, text, mem_size=128)
    assert np.array_equal(il.cpu().numpy(), ril)
    assert np.array_equal(nn.cpu().numpy(), rnn)
    for r in range(nloc):
        assert np.array_equal(jl[r, : rnn[r]].cpu().numpy(), rjl[r, : rnn[r]])
    m = dmodel.water_model()
    rng = np.random.default_rng(0)
    avg = rng.normal(0, 0.05, size=(2, m.nnei * 4))
    std = rng.uniform(0.5, 1.5, size=(2, m.nnei * 4))
    em, ed, rij, nl = ops.prod_env_mat_a(T(cext), T(text, torch.int32), il, nn, jl, T(avg), T(std), nloc, 6.0, 0.5,
                                         m.sel)
    rem, red, rrij, rnl = O.prod_env_mat_a(cext, text, ril, rnn, rjl, avg, std, nloc, 6.0, 0.5, m.sel)
    assert np.array_equal(nl.cpu().numpy(), rnl)
    assert rel_err(em.cpu(), rem) < 1e-12
    assert rel_err(ed.cpu(), red) < 1e-12
    assert rel_err(rij.cpu(), rrij) < 1e-14


def test_build_nlist_capacity_retry():
    coord, atype, b = systems.water(64, seed=3)
    O = oracle_lib()
    cext, text, _ = O.copy_coord(O.normalize_coord(coord, b), atype, b, 6.0)
    il, nn, jl, mx = ops.build_nlist(T(cext), len(atype), 6.0, T(text, torch.int32), mem_size=16)
    assert jl.shape[1] >= mx > 16


def _tab_inputs(m, seed, nloc=96):
    coord, atype, b = systems.water(nloc // 3, seed=seed)
    O = oracle_lib()
    cext, text, _ = O.copy_coord(O.normalize_coord(coord, b), atype, b, 6.0)
    il, nn, jl = O.build_nlist(cext, len(atype), 6.0, text)
    em, *_ = O.prod_env_mat_a(cext, text, il, nn, jl, m.davg.reshape(2, -1), m.dstd.reshape(2, -1), len(atype), 6.0,
                              0.5, m.sel)
    return em.reshape(len(atype), m.nnei, 4)


@pytest.mark.parametrize("tj", [0, 1])
def test_tabulate_fp64_water_blocks(O, tj):
    m = dmodel.water_model(seed=5)
    em = _tab_inputs(m, seed=5)
    rr = np.ascontiguousarray(em[:, m.sec[tj]:m.sec[tj + 1]])
    ss = np.ascontiguousarray(rr[:, :, 0])
    net = m.net_index(0, tj)
    t, info = m.tables[net], m.table_info[net]
    out = ops.tabulate_fusion_se_a(T(t), info, T(ss), T(rr), m.M)
    ref = O.tabulate_fusion_se_a(t, info, ss, rr, m.M)
    assert rel_err(out.cpu(), ref) < 1e-10
    dy = np.random.default_rng(1).normal(size=ref.shape)
    dx, dem = ops.tabulate_fusion_se_a_grad(T(t), info, T(ss), T(rr), T(dy))
    rdx, rdem = O.tabulate_fusion_se_a_grad(t, info, ss, rr, dy)
    assert rel_err(dx.cpu(), rdx) < 1e-10
    assert rel_err(dem.cpu(), rdem) < 1e-10
    zx = np.random.default_rng(2).normal(size=ss.shape)
    zm = np.random.default_rng(3).normal(size=rr.shape)
    dz = ops.tabulate_fusion_se_a_grad_grad(T(t), info, T(ss), T(rr), T(zx), T(zm), m.M)
    rdz = O.tabulate_fusion_se_a_grad_grad(t, info, ss, rr, zx, zm, m.M)
    assert rel_err(dz.cpu(), rdz) < 1e-10


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_tabulate_pt_golden(dtype, tol):
    # source/tests/pt/test_tabulate_fusion_se_a.py goldens
    p = pt_tabulate()
    t, ex, em = (T(p[k], dtype) for k in ("table_tensor", "em_x_tensor", "em_tensor"))
    info = p["table_info_tensor"]
    out = ops.tabulate_fusion_se_a(t, info, ex, em, p["M"])
    np.testing.assert_allclose(out.cpu().double().numpy(), p["expected_descriptor_tensor"], atol=tol, rtol=tol)
    dx, dem = ops.tabulate_fusion_se_a_grad(t, info, ex, em, torch.ones_like(out))
    np.testing.assert_allclose(dx.cpu().double().numpy(), p["expected_dy_dem_x"], atol=tol, rtol=tol)
    np.testing.assert_allclose(dem.cpu().double().numpy(), p["expected_dy_dem"], atol=tol, rtol=tol)


@pytest.mark.parametrize("nd", [4, 9, 16, 25])
@pytest.mark.parametrize("M,nnei", [(8, 4), (100, 46), (128, 130), (33, 1)])
def test_tabulate_shapes_and_padding(O, nd, M, nnei):
    rng = np.random.default_rng(nd * 1000 + M + nnei)
    nloc = 7
    t = rng.normal(size=(40, 6 * M)) * 0.1
    info = np.array([0.0, 0.2, 1.0, 0.01, 0.1, -1.0])
    # sorted rows with a zero-padded tail of random length (the fold case)
    ex = np.sort(rng.uniform(-0.1, 1.2, size=(nloc, nnei)), axis=1)[:, ::-1].copy()
    em = rng.normal(size=(nloc, nnei, nd))
    em[:, :, 0] = ex
    for i in range(nloc):
        npad = rng.integers(0, nnei + 1)
        if npad:
            ex[i, nnei - npad:] = 0.0
            em[i, nnei - npad:, :] = 0.0
    em[:, :, 0] = ex
    for sorted_ in (True, False):
        out = ops.tabulate_fusion_se_a(T(t), info, T(ex), T(em), M, sorted_)
        ref = O.tabulate_fusion_se_a(t, info, ex, em, M, sorted_)
        assert rel_err(out.cpu(), ref) < 1e-10
        dy = rng.normal(size=(nloc, nd, M))
        dx, dem = ops.tabulate_fusion_se_a_grad(T(t), info, T(ex), T(em), T(dy), sorted_)
        rdx, rdem = O.tabulate_fusion_se_a_grad(t, info, ex, em, dy, sorted_)
        assert rel_err(dx.cpu(), rdx) < 1e-10
        assert rel_err(dem.cpu(), rdem) < 1e-10
        zx, zm = rng.normal(size=ex.shape), rng.normal(size=em.shape)
        dz = ops.tabulate_fusion_se_a_grad_grad(T(t), info, T(ex), T(em), T(zx), T(zm), M, sorted_)
        rdz = O.tabulate_fusion_se_a_grad_grad(t, info, ex, em, zx, zm, M, sorted_)
        assert rel_err(dz.cpu(), rdz) < 1e-10


def test_tabulate_empty_neighbours():
    t = T(np.ones((4, 48)))
    info = np.array([0.0, 0.2, 1.0, 0.01, 0.1, -1.0])
    em = torch.zeros((3, 0, 4), dtype=torch.float64, device=DEV)
    ex = torch.zeros((3, 0), dtype=torch.float64, device=DEV)
    out = ops.tabulate_fusion_se_a(t, info, ex, em, 8)
    assert out.shape == (3, 4, 8) and torch.all(out == 0)


def test_tabulate_error_contracts():
    t = T(np.ones((4, 48)))
    info = np.array([0.0, 0.2, 1.0, 0.01, 0.1, -1.0])
    em = T(np.zeros((3, 2, 5)))
    ex = T(np.zeros((3, 2)))
    with pytest.raises(RuntimeError, match="basis dimension"):
        ops.tabulate_fusion_se_a(t, info, ex, em, 8)
    em = T(np.zeros((3, 2, 4)))
    with pytest.raises(RuntimeError, match="last_layer_size"):
        ops.tabulate_fusion_se_a(t, info, ex, em, 2048)


def test_prod_force_virial_and_grads(O):
    m = dmodel.water_model(seed=7)
    coord, atype, b = systems.water(64, seed=7)
    cext, text, _ = O.copy_coord(O.normalize_coord(coord, b), atype, b, 6.0)
    nloc, nall = len(atype), len(text)
    il, nn, jl = O.build_nlist(cext, nloc, 6.0, text)
    em, ed, rij, nl = O.prod_env_mat_a(cext, text, il, nn, jl, m.davg.reshape(2, -1), m.dstd.reshape(2, -1), nloc,
                                       6.0, 0.5, m.sel)
    rng = np.random.default_rng(4)
    nd = rng.normal(size=em.shape)
    f = ops.prod_force_a(T(nd), T(ed), T(nl), nloc, nall)
    assert rel_err(f.cpu(), O.prod_force_a(nd, ed, nl, nloc, nall)) < 1e-10
    # two frames
    f2 = ops.prod_force_a(T(np.tile(nd, (2, 1))), T(np.tile(ed, (2, 1))), T(np.tile(nl, (2, 1))), nloc, nall, 2)
    assert rel_err(f2.cpu(), O.prod_force_a(np.tile(nd, (2, 1)), np.tile(ed, (2, 1)), np.tile(nl, (2, 1)), nloc,
                                            nall, 2)) < 1e-10
    v, av = ops.prod_virial_a(T(nd), T(ed), T(rij), T(nl), nloc, nall)
    rv, rav = O.prod_virial_a(nd, ed, rij, nl, nloc, nall)
    assert rel_err(v.cpu(), rv) < 1e-10
    assert rel_err(av.cpu(), rav) < 1e-10
    g = rng.normal(size=nloc * 3)
    assert rel_err(ops.prod_force_grad_a(T(g), T(ed), T(nl), nloc).cpu(), O.prod_force_grad_a(g, ed, nl, nloc)) < 1e-10
    g9 = rng.normal(size=9)
    assert rel_err(ops.prod_virial_grad_a(T(g9), T(ed), T(rij), T(nl), nloc).cpu(),
                   O.prod_virial_grad_a(g9, ed, rij, nl, nloc)) < 1e-10
