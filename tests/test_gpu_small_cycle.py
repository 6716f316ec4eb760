"""The cycle below a level in one workgroup against the launch-per-kernel cycle, bit for bit.

Levels of 6 x 6 blocks with at most 8192 unknowns, down to a coarsest level with its explicit
inverse, can run their whole V-cycle (cycle.rs:461-539) as one launch of small_cycle_kernel. Its
phases repeat the per-value arithmetic and order of the kernels they replace (start, fused
Chebyshev step, residual, stored-transpose restriction, prolongation, the f64 inverse product), so
the preconditioner must give the same bits with it and without it, in every precision, and the
batched V-cycle (which never uses it) must equal it column by column.
"""
import os

import numpy as np
import pytest

import refcpu
import solidfem_amd as sf
from workloads import FINGERPRINT_CANTILEVER, block_config, loaded, plate_config

pytestmark = pytest.mark.gpu

# unknowns per level: cantilever 3528, 192, 18; block 20400, 414, 12; jittered plate
# 35280, 24624, 6222, 210, 24 (the fused part starts at level 2)
WORKLOADS = {
    "cantilever": (loaded(*FINGERPRINT_CANTILEVER[:2]), 100, 1),
    "block": (block_config(0, divisions=(12, 8, 8)), 100, 1),
    "jittered-plate": (plate_config(divisions=(24, 24, 2)), 150, 2),
}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def hierarchy(ctx, ref, coarse, precision, small):
    os.environ["SOLIDFEM_SMALL_CYCLE"] = "1" if small else "0"
    try:
        rs, cols, vals = ref.csr()
        csr = sf.CsrMatrix.new(None, rs, cols, vals)
        setup = sf.MultigridSetup.new(ctx, csr, ref.rigid_modes(), ref.free_positions(),
                                      sf.MultigridOptions(coarse_dofs=coarse, precision=precision, coarse_solve=1))
        while not setup.step(2 ** 62):
            pass
        return setup.finish()
    finally:
        del os.environ["SOLIDFEM_SMALL_CYCLE"]


@pytest.mark.parametrize("precision", list(sf.Precision), ids=lambda p: p.name.lower())
@pytest.mark.parametrize("name", list(WORKLOADS))
def test_the_one_workgroup_sub_cycle_gives_the_same_bits(ctx, name, precision):
    w, coarse, first = WORKLOADS[name]
    ref = refcpu.Problem(w)
    fused = hierarchy(ctx, ref, coarse, precision, True)
    apart = hierarchy(ctx, ref, coarse, precision, False)
    assert fused.small_cycle_level() == first and apart.small_cycle_level() == -1
    n = len(ref.csr()[0]) - 1
    rng = np.random.default_rng(17)
    r = rng.standard_normal((3, n))
    for k in range(3):
        assert np.array_equal(bits(fused.apply(r[k])), bits(apart.apply(r[k]))), k
    batch = fused.apply_batch(r)
    for k in range(3):
        assert np.array_equal(bits(batch[k]), bits(fused.apply(r[k]))), k
