"""torch.ops.deepmd.tabulate_fusion_se_a on HIP tensors: the reference's own
op tests (source/tests/pt/test_tabulate_fusion_se_a.py:1424-1506) against
this op library, including the grad-grad chain."""
import numpy as np
import pytest
import torch

from _fixtures import pt_tabulate
from dpmi import torch_ops

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _load():
    torch_ops.load()


def _inputs(dtype):
    p = pt_tabulate()
    dev = "cuda"
    t = torch.tensor(p["table_tensor"], dtype=dtype, device=dev, requires_grad=True)
    info = torch.tensor(p["table_info_tensor"], dtype=dtype, device="cpu", requires_grad=True)
    ex = torch.tensor(p["em_x_tensor"], dtype=dtype, device=dev, requires_grad=True)
    em = torch.tensor(p["em_tensor"], dtype=dtype, device=dev, requires_grad=True)
    return p, t, info, ex, em


@pytest.mark.parametrize("dtype,prec", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_forward_backward(dtype, prec):
    p, t, info, ex, em = _inputs(dtype)
    out = torch.ops.deepmd.tabulate_fusion_se_a(t, info, ex, em, p["M"])[0]
    exp = torch.tensor(p["expected_descriptor_tensor"], dtype=dtype, device="cuda")
    torch.testing.assert_close(out, exp, atol=prec, rtol=prec)
    out.sum().backward()
    torch.testing.assert_close(ex.grad, torch.tensor(p["expected_dy_dem_x"], dtype=dtype, device="cuda"),
                               atol=prec, rtol=prec)
    torch.testing.assert_close(em.grad, torch.tensor(p["expected_dy_dem"], dtype=dtype, device="cuda"),
                               atol=prec, rtol=prec)
    assert t.grad is None and info.grad is None


def test_second_order_backward_matches_finite_difference():
    # tabulate_test_utils.py:38-92
    p, t, info, ex, em = _inputs(torch.float64)
    desc = torch.ops.deepmd.tabulate_fusion_se_a(t, info, ex, em, p["M"])[0]

    def nonuniform(x):
        return torch.linspace(-0.5, 0.75, x.numel(), dtype=x.dtype, device=x.device).reshape_as(x)

    dy = nonuniform(desc).requires_grad_(True)
    g = torch.autograd.grad(desc, (ex, em), grad_outputs=dy, create_graph=True, retain_graph=True)
    w = tuple(nonuniform(x) for x in g)
    proj = sum((gi * wi).sum() for gi, wi in zip(g, w))
    dz = torch.autograd.grad(proj, dy, retain_graph=True)[0]
    eps = 1e-6
    ref = torch.empty_like(dy)
    base = dy.detach()

    def project(d):
        gg = torch.autograd.grad(desc, (ex, em), grad_outputs=d, retain_graph=True)
        return sum((gi * wi).sum() for gi, wi in zip(gg, w))

    for k in range(dy.numel()):
        a = base.clone()
        b = base.clone()
        a.view(-1)[k] += eps
        b.view(-1)[k] -= eps
        ref.view(-1)[k] = (project(a) - project(b)) / (2 * eps)
    torch.testing.assert_close(dz, ref, atol=1e-6, rtol=1e-6)


def test_device_contract_errors():
    p, t, info, ex, em = _inputs(torch.float64)
    with pytest.raises(RuntimeError, match="table_info must be on the CPU"):
        torch.ops.deepmd.tabulate_fusion_se_a(t, info.cuda(), ex, em, p["M"])
    with pytest.raises(RuntimeError, match="same device as table"):
        torch.ops.deepmd.tabulate_fusion_se_a(t, info, ex.cpu(), em, p["M"])
    with pytest.raises(RuntimeError, match="last_layer_size must be between 1 and 1024"):
        torch.ops.deepmd.tabulate_fusion_se_a(t, info, ex, em, 0)


@pytest.mark.parametrize("dtype,prec", [(torch.float64, 1e-10), (torch.float32, 1e-5)])
def test_host_tensors_staged_to_the_gpu(dtype, prec):
    """Host tensors (the reference's CPU kernels) give the golden values and
    gradients on the host; the arithmetic runs on the HIP device."""
    p = pt_tabulate()
    t = torch.tensor(p["table_tensor"], dtype=dtype)
    info = torch.tensor(p["table_info_tensor"], dtype=dtype)
    ex = torch.tensor(p["em_x_tensor"], dtype=dtype, requires_grad=True)
    em = torch.tensor(p["em_tensor"], dtype=dtype, requires_grad=True)
    out = torch.ops.deepmd.tabulate_fusion_se_a(t, info, ex, em, p["M"])[0]
    assert out.device.type == "cpu"
    torch.testing.assert_close(out, torch.tensor(p["expected_descriptor_tensor"], dtype=dtype), atol=prec, rtol=prec)
    out.sum().backward()
    assert ex.grad.device.type == "cpu"
    torch.testing.assert_close(ex.grad, torch.tensor(p["expected_dy_dem_x"], dtype=dtype), atol=prec, rtol=prec)
    torch.testing.assert_close(em.grad, torch.tensor(p["expected_dy_dem"], dtype=dtype), atol=prec, rtol=prec)


def test_non_contiguous_input_raises():
    p, t, info, ex, em = _inputs(torch.float64)
    wide = torch.cat([em, em], dim=2)[:, :, :4]  # a strided view of the same shape
    assert not wide.is_contiguous()
    with pytest.raises(RuntimeError, match="contiguous"):
        torch.ops.deepmd.tabulate_fusion_se_a(t, info, ex, wide, p["M"])
