"""A caller's preconditioner through the Preconditioner seam (crates/solidfem-solve/src/pcg.rs sees
only the trait of preconditioner.rs:9-20), on the GPU.

Jacobi written by the caller -- z = r * (1 / diag(A)), the operations of Jacobi::apply
(preconditioner.rs:53-80) -- must give the bits and iteration counts of the library's own Jacobi
solve, in both dot orders, and in the reference's order the oracle's Jacobi solve; in host form
(numpy arrays) and in device form (device addresses and the solve's stream, the staging done here
with the HIP runtime). A failing callback fails the step by name.
"""
import ctypes as C

import numpy as np
import pytest

import refcpu
import solidfem_amd as sf
from workloads import block_config

pytestmark = pytest.mark.gpu

PRECONDITIONER_FAILED = 29


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.fixture(scope="module")
def system():
    ref = refcpu.Problem(block_config(0, divisions=(6, 4, 4)))
    rs, cols, vals = ref.csr()
    diag = np.array([vals[rs[i]:rs[i + 1]][cols[rs[i]:rs[i + 1]] == i][0] for i in range(len(rs) - 1)])
    return rs, cols, vals, ref.rhs(), 1.0 / diag


def host_jacobi(inverse):
    def apply(r, z):
        z[:] = r * inverse
    return apply


def device_jacobi(inverse):
    """The same product, with r brought to the host and z sent back on the solve's stream."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    to_host, to_device = 2, 1

    def apply(r, z, n, stream):
        host_r, host_z = np.empty(n), np.empty(n)
        assert hip.hipMemcpyAsync(host_r.ctypes.data, r, 8 * n, to_host, stream) == 0
        assert hip.hipStreamSynchronize(stream) == 0
        host_z[:] = host_r * inverse
        assert hip.hipMemcpyAsync(z, host_z.ctypes.data, 8 * n, to_device, stream) == 0
        assert hip.hipStreamSynchronize(stream) == 0
    return apply


@pytest.mark.parametrize("dot", [sf.DotOrder.REFERENCE, sf.DotOrder.TREE])
def test_a_callers_jacobi_is_the_librarys_jacobi(ctx, system, dot):
    rs, cols, vals, b, inverse = system
    csr = sf.CsrMatrix.new(ctx, rs, cols, vals)
    options = sf.PcgOptions(tolerance=1e-8, dot_order=dot)
    own = sf.Pcg.new(ctx, csr, None, b, options)
    p_own = own.step(100000)
    assert p_own.outcome == sf.StepOutcome.CONVERGED
    for name, pcg in (("host", sf.Pcg.with_preconditioner(ctx, csr, host_jacobi(inverse), b, options)),
                      ("device", sf.Pcg.with_preconditioner(ctx, csr, device_jacobi(inverse), b, options, device=True))):
        p = pcg.step(100000)
        print(f"{dot.name} {name}: {p.iteration} iterations (library Jacobi {p_own.iteration})", flush=True)
        assert p.outcome == sf.StepOutcome.CONVERGED and p.iteration == p_own.iteration
        assert np.array_equal(bits(pcg.estimate()), bits(own.estimate())), name
    if dot == sf.DotOrder.REFERENCE:
        xr, _ = refcpu.pcg_jacobi(rs, cols, vals, b, tolerance=1e-8)
        assert np.array_equal(bits(own.estimate()), bits(xr))


def test_a_sliced_solve_with_a_callers_preconditioner_has_the_same_bits(ctx, system):
    rs, cols, vals, b, inverse = system
    csr = sf.CsrMatrix.new(ctx, rs, cols, vals)
    options = sf.PcgOptions(tolerance=1e-8)
    whole = sf.Pcg.with_preconditioner(ctx, csr, host_jacobi(inverse), b, options)
    whole.step(100000)
    sliced = sf.Pcg.with_preconditioner(ctx, csr, host_jacobi(inverse), b, options)
    while sliced.step(7).outcome == sf.StepOutcome.RUNNING:
        pass
    assert np.array_equal(bits(sliced.estimate()), bits(whole.estimate()))


def test_a_failing_preconditioner_fails_the_step(ctx, system):
    rs, cols, vals, b, inverse = system
    csr = sf.CsrMatrix.new(ctx, rs, cols, vals)
    calls = [0]

    def flaky(r, z):
        calls[0] += 1
        if calls[0] > 3:
            raise FloatingPointError("the caller's preconditioner failed")
        z[:] = r * inverse

    pcg = sf.Pcg.with_preconditioner(ctx, csr, flaky, b)
    with pytest.raises(FloatingPointError):
        pcg.step(100)
    # the C ABI reports it by name
    lib = sf.library()
    fn = sf._capi.APPLY_FN(lambda r, z, n, s, u: 7)
    pre = sf._capi.Preconditioner(fn, None, 0, 0)
    h = C.c_void_p()
    status = lib.sf_pcg_create_with(ctx._ptr, csr._ptr, C.byref(pre), b.ctypes.data_as(C.POINTER(C.c_double)),
                                    None, C.byref(h))
    assert status == PRECONDITIONER_FAILED
    a, bb, v = C.c_int64(), C.c_int64(), C.c_double()
    lib.sf_last_error_payload(C.byref(a), C.byref(bb), C.byref(v))
    assert a.value == 7
