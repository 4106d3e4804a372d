"""dpmi_prod_virial_a_f64_noalloc (fixed device scratch, no allocation)
against dpmi_prod_virial_a_f64 through the C ABI: identical kernels, so the
virial (per-row partials summed in a fixed order) agrees bit for bit and the
per-atom virials (FP64 atomics, whose order varies between any two calls)
within 1e-12 of their largest magnitude; two calls in a row, the
second with fewer rows (the scratch keeps the first call's partials), and
the row limit returns DPMI_ERR_ARGUMENT before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from dpmi.lib import load

pytestmark = pytest.mark.gpu


def _inputs(nloc, nall, nnei, seed):
    rng = np.random.default_rng(seed)
    nl = rng.integers(0, nall, size=(nloc, nnei)).astype(np.int32)
    nl[rng.random((nloc, nnei)) < 0.2] = -1
    dev = "cuda"
    nd = torch.as_tensor(rng.normal(size=nloc * nnei * 4), device=dev)
    ed = torch.as_tensor(rng.normal(size=nloc * nnei * 12), device=dev)
    rij = torch.as_tensor(rng.normal(size=nloc * nnei * 3), device=dev)
    return nd, ed, rij, torch.as_tensor(nl.reshape(-1), device=dev)


def _call(fn, nloc, nall, nnei, nd, ed, rij, nl):
    v = torch.full((9,), np.nan, dtype=torch.float64, device="cuda")
    av = torch.full((nall * 9,), np.nan, dtype=torch.float64, device="cuda")
    p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    rc = fn(p(v), p(av), p(nd), p(ed), p(rij), p(nl), nloc, nall, nnei, None)
    torch.cuda.synchronize()
    assert rc == 0, load().dpmi_last_error()
    return v.cpu().numpy(), av.cpu().numpy()


def test_noalloc_virial_matches_allocating_entry():
    lib = load()
    nnei = 138
    for k, (nloc, nall) in enumerate([(3000, 6100), (257, 700)]):
        print(f"case noalloc virial call {k}: nloc={nloc} nall={nall}", flush=True)
        args = _inputs(nloc, nall, nnei, seed=k)
        v0, a0 = _call(lib.dpmi_prod_virial_a_f64, nloc, nall, nnei, *args)
        v1, a1 = _call(lib.dpmi_prod_virial_a_f64_noalloc, nloc, nall, nnei, *args)
        assert np.all(np.isfinite(v1)) and np.all(np.isfinite(a1))
        assert np.array_equal(v0, v1)
        assert np.max(np.abs(a0 - a1)) <= 1e-12 * np.max(np.abs(a0))


def test_noalloc_virial_row_limit():
    lib = load()
    d = torch.zeros(16, dtype=torch.float64, device="cuda")
    n = torch.zeros(16, dtype=torch.int32, device="cuda")
    p = C.c_void_p(d.data_ptr())
    rc = lib.dpmi_prod_virial_a_f64_noalloc(p, p, p, p, p, C.c_void_p(n.data_ptr()), (1 << 17) + 1, 4, 1, None)
    assert rc == 1 and b"fixed scratch" in lib.dpmi_last_error()
