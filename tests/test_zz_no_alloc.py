"""The allocation-free rounding bound (sf_csr_rounding_bound_noalloc) against the allocating entry
point (sf_csr_rounding_bound, CsrMatrix::multiply_rounding_bound of
crates/solidfem-solve/src/csr.rs:142-154), bit for bit, through the C ABI with device pointers.

Two calls in a row share the per-GPU scratch array, the second on a smaller matrix, so a term
left over from the first call would show. Device memory for x and the bound comes from torch.
"""
import ctypes as C

import numpy as np
import pytest

import refcpu
import solidfem_amd as sf
from solidfem_amd import _capi
from workloads import FINGERPRINT_CANTILEVER, loaded, plate_config

pytestmark = pytest.mark.gpu


def bits(v):
    return np.float64(v).view(np.uint64)


def noalloc_bound(ctx, csr, x_host, dot_order, stream=None):
    import torch

    x = torch.from_numpy(np.ascontiguousarray(x_host)).to("cuda:0")
    bound = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    lib = _capi.load()
    status = lib.sf_csr_rounding_bound_noalloc(csr._ptr, C.c_void_p(x.data_ptr()), int(dot_order),
                                              C.c_void_p(bound.data_ptr()), C.c_void_p(stream) if stream else None)
    sf._raise(status)
    if stream:
        torch.cuda.current_stream().synchronize()
    else:
        ctx.synchronize()
    return float(bound.cpu().numpy()[0])


@pytest.mark.parametrize("dot_order", [sf.DotOrder.TREE, sf.DotOrder.REFERENCE], ids=["tree", "reference"])
def test_noalloc_bound_equals_the_allocating_bound_on_two_calls(ctx, dot_order):
    import torch

    larger = refcpu.Problem(plate_config(divisions=(12, 12, 2)))
    smaller = refcpu.Problem(loaded(*FINGERPRINT_CANTILEVER[:2]))
    assert smaller.free_dofs < larger.free_dofs
    rng = np.random.default_rng(11)
    for k, ref in enumerate((larger, smaller)):
        rs, cols, vals = ref.csr()
        csr = sf.CsrMatrix.new(ctx, rs, cols, vals)
        x = rng.standard_normal(ref.free_dofs) * np.exp(rng.uniform(-20, 20, ref.free_dofs))
        expected = csr.multiply_rounding_bound(x, dot_order)
        # the first call on the context's stream, the second on torch's stream
        stream = torch.cuda.current_stream().cuda_stream if k == 1 else None
        got = noalloc_bound(ctx, csr, x, dot_order, stream)
        assert expected > 0
        assert bits(got) == bits(expected), (k, got, expected)


def test_noalloc_bound_refuses_a_matrix_beyond_its_scratch(ctx):
    import torch

    limit = _capi.load().sf_noalloc_max_rows()
    n = int(limit) + 1
    rs = np.arange(n + 1, dtype=np.int64)
    cols = np.arange(n, dtype=np.uint32)
    vals = np.ones(n)
    csr = sf.CsrMatrix.new(ctx, rs, cols, vals)
    x = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    bound = torch.zeros(1, dtype=torch.float64, device="cuda:0")
    status = _capi.load().sf_csr_rounding_bound_noalloc(csr._ptr, C.c_void_p(x.data_ptr()), 0, C.c_void_p(bound.data_ptr()), None)
    assert status == 1  # SF_ERR_INVALID_ARGUMENT
    with pytest.raises(sf.SolidfemError) as e:
        sf._raise(status)
    assert e.value.a == n and e.value.b == limit
