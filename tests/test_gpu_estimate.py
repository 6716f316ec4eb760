"""The error estimate (crates/solidfem-elastic/src/error_estimate.rs) on the GPU against the oracle.

The device estimate repeats element_integrals' operations in the reference's order and keeps the
running sums in element order, so every element error and every sum must be the oracle's bits: on
solved displacements and on random fields, over jittered and tube meshes and both solve rules.
The reference's own tests of the estimate's surface run through the C ABI: slicing and cancelling
give one run's bits and an unfinished estimate refuses (convergence.rs:119-164), a preview of an
unconverged solve refuses (exact.rs:79-97), a patch of two materials refuses by vertex
(exact.rs:61-77), a recovery of another mesh refuses (error_estimate.rs:128-134), and an exactly
represented bending field held on the boundary reads as converged through the device study
(exact.rs:27-59).
"""
import dataclasses

import numpy as np
import pytest

import refcpu
import solidfem_amd as sf
from workloads import FINGERPRINT_CANTILEVER, bar_config, loaded, plate_config, tube_config

pytestmark = pytest.mark.gpu

NOT_FINISHED, CANCELLED, MIXED_MATERIAL_PATCH, MESH_MISMATCH = 48, 47, 53, 57

FIELDS = {
    "jittered-plate": plate_config(divisions=(8, 8, 2)),
    "jittered-plate-14-point": dataclasses.replace(plate_config(divisions=(6, 6, 2)), quadrature_points=14),
    "tube": tube_config(divisions=(20, 7, 8), wall_cells=2, poisson=0.3),
}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def device_estimate(ctx, model, u, slice_elements=2 ** 62):
    recovery = model.begin_recovery(ctx, u, converged=True)
    assert recovery.advance(2 ** 62)
    estimation = model.begin_error_estimate(recovery)
    while not estimation.advance(slice_elements):
        pass
    return estimation.estimate()


def assert_same(estimate, oracle):
    assert oracle["status"] == 0
    assert np.array_equal(bits(estimate.element_errors), bits(oracle["element_errors"]))
    for name in ("error_norm_squared", "solution_norm_squared", "relative_error"):
        assert bits(getattr(estimate, name)) == bits(oracle[name]), name


@pytest.mark.parametrize("name", list(FIELDS))
def test_device_estimate_of_a_random_field_is_bit_identical(ctx, name):
    w = FIELDS[name]
    ref = refcpu.Problem(w)
    u = 1e-3 * np.random.default_rng(17).standard_normal((ref.node_count, 3))
    assert_same(device_estimate(ctx, sf.Model.from_workload(w), u), ref.error_estimate(u))


def test_device_estimate_of_a_solved_study_is_bit_identical(ctx):
    w = loaded(*FINGERPRINT_CANTILEVER[:2])
    study = sf.StaticStudy.new(ctx, sf.Model.from_workload(w), sf.PcgOptions(tolerance=1e-10),
                               sf.MultigridOptions(coarse_dofs=100))
    assert study.solve().outcome == sf.StepOutcome.CONVERGED
    (points, nodal), estimate = study.solution_with_error_estimate()
    ref = refcpu.Problem(w)
    u = study.displacements()
    assert_same(estimate, ref.error_estimate(u))
    rp, rn = ref.recover(u)
    assert np.array_equal(bits(points), bits(rp)) and np.array_equal(bits(nodal), bits(rn))
    assert 0.0 < estimate.relative_error < 1.0
    assert study.report().recovery_seconds > 0.0


def test_an_estimate_in_slices_gives_the_bits_of_one_run_and_resumes_after_a_cancel(ctx):
    w = FIELDS["jittered-plate"]
    model = sf.Model.from_workload(w)
    u = 1e-3 * np.random.default_rng(3).standard_normal((refcpu.Problem(w).node_count, 3))
    whole = device_estimate(ctx, model, u)
    sliced = device_estimate(ctx, model, u, slice_elements=5)
    assert np.array_equal(bits(sliced.element_errors), bits(whole.element_errors))
    assert bits(sliced.relative_error) == bits(whole.relative_error)
    recovery = model.begin_recovery(ctx, u, converged=True)
    while not recovery.advance(7):
        pass
    partial = model.begin_error_estimate(recovery)
    with pytest.raises(sf.ModelError) as e:
        partial.estimate()
    assert e.value.status == NOT_FINISHED
    calls = [0]

    def cancel_on_fourth():
        calls[0] += 1
        return calls[0] == 4

    with pytest.raises(sf.ModelError) as e:
        partial.advance(2 ** 62, cancel_on_fourth)
    assert e.value.status == CANCELLED and partial.elements_done() == 3
    assert partial.advance(2 ** 62)
    resumed = partial.estimate()
    assert np.array_equal(bits(resumed.element_errors), bits(whole.element_errors))
    assert bits(resumed.error_norm_squared) == bits(whole.error_norm_squared)


def test_a_preview_of_an_unconverged_solve_refuses_an_estimate(ctx):
    study = sf.StaticStudy.new(ctx, sf.Model.from_workload(loaded([4.0, 1.0, 1.0], [8, 2, 2])),
                               sf.PcgOptions(tolerance=1e-12), sf.MultigridOptions(coarse_dofs=100))
    study.step(2)
    recovery = study.begin_recovery()
    assert study.continue_recovery(recovery, 2 ** 62)
    estimation = study.begin_error_estimate(recovery)
    for call in (lambda: study.continue_error_estimate(estimation, 2 ** 62), lambda: study.finish_error_estimate(estimation)):
        with pytest.raises(sf.ModelError) as e:
            call()
        assert e.value.status == NOT_FINISHED


def test_a_patch_holding_two_materials_refuses_with_the_oracles_vertex(ctx):
    w = bar_config(divisions=(8, 2, 2))
    w.shell_depth = 5.0
    ref = refcpu.Problem(w)
    u = 1e-3 * np.random.default_rng(2).standard_normal((ref.node_count, 3))
    oracle = ref.error_estimate(u)
    assert oracle["status"] == MIXED_MATERIAL_PATCH
    model = sf.Model.from_workload(w)
    recovery = model.begin_recovery(ctx, u, converged=True)
    recovery.advance(2 ** 62)
    estimation = model.begin_error_estimate(recovery)
    with pytest.raises(sf.ModelError) as e:
        estimation.advance(2 ** 62)
    assert e.value.status == MIXED_MATERIAL_PATCH and e.value.a == oracle["a"]
    assert estimation.elements_done() == len(oracle["element_errors"])


def test_a_recovery_of_another_mesh_is_refused(ctx):
    a, b = loaded([2.0, 1.0, 1.0], [4, 2, 2]), loaded([2.0, 1.0, 1.0], [4, 2, 3])
    u = np.zeros((refcpu.Problem(a).node_count, 3))
    recovery = sf.Model.from_workload(a).begin_recovery(ctx, u, converged=True)
    recovery.advance(2 ** 62)
    estimation = sf.Model.from_workload(b).begin_error_estimate(recovery)
    with pytest.raises(sf.ModelError) as e:
        estimation.advance(2 ** 62)
    assert e.value.status == MESH_MISMATCH


def test_an_exactly_represented_bending_field_reads_as_converged(ctx):
    """exact.rs:17-59 through the device study: a 3 x 2 x 2 box of 3 x 2 x 2 cells, E 2100, nu 0.3,
    every boundary node held at the bending field, solved to 1e-14."""
    kappa, nu, size = 1e-3, 0.3, (3.0, 2.0, 2.0)

    def bending(p):
        x, y, z = p
        return [kappa * x * y, -0.5 * kappa * (x * x + nu * (y * y - z * z)), -nu * kappa * y * z]

    w = sf.Workload(size=size, divisions=(3, 2, 2), clamp_axis=-1, load_axis=-1, youngs=2100.0, poisson=nu)
    model = sf.Model.from_workload(w)
    for n, p in enumerate(model.mesh().nodes()):
        if any(p[a] == 0.0 or p[a] == size[a] for a in range(3)):
            for a, v in enumerate(bending(p)):
                model.prescribe_displacement(n, a, v)
    study = sf.StaticStudy.new(ctx, model, sf.PcgOptions(tolerance=1e-14), sf.MultigridOptions(coarse_dofs=100))
    progress = study.solve()
    assert progress.outcome == sf.StepOutcome.CONVERGED
    _, estimate = study.solution_with_error_estimate()
    tolerance = 10.0 * (1e2 * progress.relative_residual + 1e-13)
    assert estimate.relative_error <= tolerance
    scale = np.sqrt(estimate.solution_norm_squared)
    assert np.all(estimate.element_errors <= tolerance * scale)
    assert estimate.solution_norm_squared == pytest.approx(2100.0 * kappa ** 2 * 16.0, rel=1e-12)
