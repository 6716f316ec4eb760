"""Stress recovery on the GPU against the oracle, bit for bit.

The device recovery (integration-point stresses, the worst incident node values, the vertex patch
fits and the nodal field of crates/solidfem-elastic/src/results.rs:307-576, element.rs:135-192
and patch.rs) repeats the reference's operations in the reference's order, so every value must
be the oracle's bits: on solved displacements, on random fields, with one and two materials, and
with the 4- and 14-point rules.
"""
import dataclasses

import numpy as np
import pytest

import refcpu
import solidfem_amd as sf
from workloads import FINGERPRINT_CANTILEVER, bar_config, loaded, plate_config

pytestmark = pytest.mark.gpu

FIELDS = {
    "jittered-plate": plate_config(divisions=(8, 8, 2)),
    "shell-core-bar": bar_config(divisions=(8, 4, 4)),
    "shell-core-bar-14-point": dataclasses.replace(bar_config(divisions=(4, 2, 2)), quadrature_points=14),
}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("name", list(FIELDS))
def test_device_recovery_of_a_random_field_is_bit_identical(ctx, name):
    w = FIELDS[name]
    ref = refcpu.Problem(w)
    u = 1e-3 * np.random.default_rng(11).standard_normal((ref.node_count, 3))
    rp, rn = ref.recover(u)
    dp, dn = sf.Model.from_workload(w).recover(ctx, u)
    assert dp.shape == rp.shape and dn.shape == rn.shape
    assert np.array_equal(bits(dp), bits(rp))
    assert np.array_equal(bits(dn), bits(rn))


def test_recovery_of_a_solved_study_is_bit_identical(ctx):
    size, div, _, _ = FINGERPRINT_CANTILEVER
    w = loaded(size, div)
    b = sf.StudyBuilder.new(ctx, sf.Model.from_workload(w), sf.PcgOptions(dot_order=sf.DotOrder.REFERENCE),
                            sf.MultigridOptions(precision=sf.Precision.DOUBLE, coarse_dofs=100))
    b.step(2 ** 62)
    study = b.finish()
    assert study.solve().outcome == sf.StepOutcome.CONVERGED
    u = study.displacements()
    points, nodal = study.recover()
    rp, rn = refcpu.Problem(w).recover(u)
    assert np.array_equal(bits(points), bits(rp)) and np.array_equal(bits(nodal), bits(rn))
    # a cantilever under tip shear: the bending moment peaks at the clamp, so the largest
    # recovered von Mises lies in the first layer of elements (12 layers over the length)
    x = refcpu.Problem(w).nodes()
    assert x[np.argmax(nodal[:, 6]), 0] <= size[0] / div[0] + 1e-12
