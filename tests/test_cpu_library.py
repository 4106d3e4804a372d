"""CPU-side checks of the shipped libraries (no GPU compute is launched):
every entry point declared in include/dpmi.h is exported by libdpmi.so, the
torch op library registers the reference schema, and host-side argument
validation behaves like the reference."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "dpmi.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dpmi_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from dpmi import lib

    so = ctypes.CDLL(lib.LIBPATH)
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing
    # the Python binding covers the whole header
    assert set(names) == set(lib.SIGNATURES)


def test_deepmd_cpp_templates_exported():
    """The deepmd:: drop-in templates (include/deepmd_lib.h) are exported
    with C++ linkage like the reference libdeepmd."""
    import subprocess

    from dpmi import lib

    out = subprocess.run(["nm", "-DC", lib.LIBPATH], capture_output=True, text=True).stdout
    for name in ("deepmd::tabulate_fusion_se_a_gpu<double>", "deepmd::tabulate_fusion_se_a_grad_gpu<double>",
                 "deepmd::tabulate_fusion_se_a_grad_grad_gpu<double>", "deepmd::prod_env_mat_a_gpu<double>",
                 "deepmd::prod_force_a_gpu<double>", "deepmd::prod_virial_a_gpu<double>",
                 "deepmd::prod_force_grad_a_gpu<double>", "deepmd::prod_virial_grad_a_gpu<double>",
                 "deepmd::prod_force_a_gpu<float>", "deepmd::prod_virial_a_gpu<float>",
                 "deepmd::prod_force_grad_a_gpu<float>", "deepmd::prod_virial_grad_a_gpu<float>",
                 "deepmd::build_nlist_gpu<double>", "deepmd::copy_coord_gpu<double>",
                 "deepmd::normalize_coord_gpu<double>", "deepmd::tabulate_fusion_se_a_gpu<float>",
                 "deepmd::normalize_coord_gpu<float>", "deepmd::copy_coord_gpu<float>", "deepmd::build_nlist_gpu<float>",
                 "deepmd::prod_env_mat_a_gpu<float>", "deepmd::format_nbor_list_gpu<double>",
                 "deepmd::format_nbor_list_gpu<float>", "deepmd::compute_cell_info<double>",
                 "deepmd::convert_nlist_gpu_device", "deepmd::free_nlist_gpu_device",
                 "deepmd::validate_nlist_gpu_capacity", "deepmd::use_nlist_map", "deepmd::env_mat_nbor_update",
                 "deepmd::normalize_coord_cpu<double>", "deepmd::copy_coord_cpu<double>",
                 "deepmd::build_nlist_cpu<double>", "deepmd::format_nlist_cpu<double>",
                 "deepmd::prod_env_mat_a_cpu<double>", "deepmd::tabulate_fusion_se_a_cpu<double>",
                 "deepmd::tabulate_fusion_se_a_grad_cpu<float>", "deepmd::tabulate_fusion_se_a_grad_grad_cpu<double>",
                 "deepmd::prod_force_a_cpu<double>", "deepmd::prod_virial_a_cpu<float>",
                 "deepmd::prod_force_grad_a_cpu<double>", "deepmd::prod_virial_grad_a_cpu<double>",
                 "dpmi::DeepPot::init", "dpmi::DeepPot::cutoff"):
        assert name in out, name


def test_version_string():
    from dpmi import lib

    assert b"gfx950" in lib.load().dpmi_version()


def test_torch_op_schema_matches_reference():
    from dpmi import torch_ops

    torch_ops.load()
    # the reference registers the function with m.def(name, fn)
    # (tabulate_multi_device.cc:1721-1723), i.e. an inferred schema of four
    # tensors (table, table_info, em_x, em) and an int last_layer_size
    schema = str(torch.ops.deepmd.tabulate_fusion_se_a.default._schema)
    assert schema == "deepmd::tabulate_fusion_se_a(Tensor _0, Tensor _1, Tensor _2, Tensor _3, int _4) -> Tensor[] _0"
    assert hasattr(torch.ops.deepmd, "enable_mpi")


def test_torch_op_host_tensors_need_a_gpu():
    # host tensors are staged to the HIP device (there is no CPU arithmetic):
    # without a visible GPU the call fails loudly
    from dpmi import torch_ops

    torch_ops.load()
    t = torch.zeros((4, 48), dtype=torch.float64)
    with pytest.raises(RuntimeError, match="HIP"):
        torch.ops.deepmd.tabulate_fusion_se_a(t, torch.zeros(6, dtype=torch.float64), torch.zeros((2, 3), dtype=torch.float64),
                                              torch.zeros((2, 3, 4), dtype=torch.float64), 8)


def test_model_table_layout():
    from dpmi import model

    m = model.water_model(seed=3, env_stat=False)
    assert m.nnets == 4 and m.tables[0].shape == (280, 600)
    np.testing.assert_allclose(m.table_info[0], [0.0, 2.0, 10.0, 0.01, 0.1, -1.0])
    assert [m.net_index(ti, tj) for ti in range(2) for tj in range(2)] == [0, 2, 1, 3]


def test_env_stat_of_a_frame():
    """davg/dstd as the reference's EnvMatStatSe sets them: [avg_r, 0, 0, 0]
    and [std_r, std_a x 3] per centre type, over every slot (empty slots as
    zeros); checked on a two-atom-type frame against a direct evaluation."""
    from dpmi import model, systems

    m = model.SeAModel.random(2, [46, 92], 6.0, 0.5, [8, 16], 4, [16], seed=1)
    coord, atype, box = systems.water(64, seed=5)
    m.set_env_stat([(coord, atype, box)])
    L = box[0]
    x = np.mod(coord, L)
    per_type = {0: [], 1: []}
    for i in range(len(atype)):
        d = x - x[i]
        d -= L * np.round(d / L)
        r = np.linalg.norm(d, axis=1)
        env = np.zeros((m.nnei, 4))
        for tj in range(2):
            js = [j for j in np.argsort(r, kind="stable") if j != i and atype[j] == tj and r[j] < m.rcut]
            js = js[: m.sel[tj]]
            for s, j in enumerate(js):
                u = min(max((r[j] - m.rcut_smth) / (m.rcut - m.rcut_smth), 0.0), 1.0)
                sw = 1.0 if r[j] < m.rcut_smth else u ** 3 * (-6 * u * u + 15 * u - 10) + 1
                env[m.sec[tj] + s] = [sw / r[j], *(sw / r[j] ** 2 * d[j])]
        per_type[atype[i]].append(env)
    for t in range(2):
        e = np.concatenate(per_type[t])
        np.testing.assert_allclose(m.davg[t, :, 0], e[:, 0].mean(), rtol=1e-12)
        assert np.all(m.davg[t, :, 1:] == 0.0)
        np.testing.assert_allclose(m.dstd[t, :, 0], e[:, 0].std(), rtol=1e-10)
        np.testing.assert_allclose(m.dstd[t, :, 1:], e[:, 1:].std(), rtol=1e-10)
