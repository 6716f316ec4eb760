"""Stress recovery on the GPU against the oracle, bit for bit.

The device recovery (integration-point stresses, the worst incident node values, the vertex patch
fits and the nodal field of crates/solidfem-elastic/src/results.rs:307-576, element.rs:135-192
and patch.rs) repeats the reference's operations in the reference's order, so every value must
be the oracle's bits: on solved displacements, on random fields, with one and two materials, and
with the 4- and 14-point rules. The resumable recovery must give the same bits whatever its slices
and cancellations, with the reference's budget order (elements first, then one patch fit per
vertex); the nodal rotations (results.rs:262-305) must be the oracle's bits too, and a rigid
rotation must read its angle.
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


# ---- the resumable recovery (Recovery, results.rs:103-124 and :440-536) ----

MODEL_NOT_FINISHED, MODEL_CANCELLED = 48, 47


@pytest.mark.parametrize("name", ["shell-core-bar", "shell-core-bar-14-point"])
def test_recovery_in_slices_has_the_bits_of_one_run(ctx, name):
    w = FIELDS[name]
    ref = refcpu.Problem(w)
    u = 1e-3 * np.random.default_rng(12).standard_normal((ref.node_count, 3))
    rp, rn = ref.recover(u)
    model = sf.Model.from_workload(w)
    elements, vertices = rp.shape[0], ref.vertex_count
    print(f"recovery slices: {name}, {elements} elements, {vertices} vertices", flush=True)
    for size in (1, 7, 333, elements, elements + 1, elements + vertices, 2 ** 62):
        r = model.begin_recovery(ctx, u)
        st = r.state()
        assert (st.elements, st.vertices, st.elements_done, st.fits_done, st.finished) == (elements, vertices, 0, 0, 0)
        with pytest.raises(sf.ModelError) as e:
            r.field()
        assert e.value.status == MODEL_NOT_FINISHED
        calls = 0
        while not r.advance(size):
            calls += 1
            st = r.state()
            # the budget goes to elements first; fits start once every element is done
            assert st.elements_done == min(elements, calls * size)
            assert st.fits_done == max(0, calls * size - elements)
        assert calls == (elements + vertices + size - 1) // size - 1
        dp, dn = r.field()
        assert np.array_equal(bits(dp), bits(rp)) and np.array_equal(bits(dn), bits(rn)), size
        assert np.array_equal(r.displacements(), u) and not r.is_converged()


def test_a_cancelled_recovery_keeps_its_units_and_resumes_to_the_same_bits(ctx):
    w = FIELDS["shell-core-bar"]
    ref = refcpu.Problem(w)
    u = 1e-3 * np.random.default_rng(13).standard_normal((ref.node_count, 3))
    rp, rn = ref.recover(u)
    elements = rp.shape[0]
    model = sf.Model.from_workload(w)
    for allowed in (0, 5, elements - 1, elements, elements + 9):
        r = model.begin_recovery(ctx, u)
        calls = [0]

        def cancel():
            calls[0] += 1
            return calls[0] > allowed

        with pytest.raises(sf.ModelError) as e:
            r.advance(2 ** 62, cancel)
        assert e.value.status == MODEL_CANCELLED
        st = r.state()
        assert st.elements_done + st.fits_done == allowed and calls[0] == allowed + 1
        assert r.advance(2 ** 62)
        dp, dn = r.field()
        assert np.array_equal(bits(dp), bits(rp)) and np.array_equal(bits(dn), bits(rn)), allowed


def test_a_study_recovers_its_preview_and_its_answer_in_slices(ctx):
    size, div, _, _ = FINGERPRINT_CANTILEVER
    w = loaded(size, div)
    study = sf.StaticStudy.new(ctx, sf.Model.from_workload(w), sf.PcgOptions(dot_order=sf.DotOrder.REFERENCE),
                               sf.MultigridOptions(precision=sf.Precision.DOUBLE, coarse_dofs=100))
    ref = refcpu.Problem(w)
    # a preview of the solve in progress
    study.step(5)
    preview = study.begin_recovery()
    assert not preview.is_converged()
    while not study.continue_recovery(preview, 100):
        pass
    pp, pn = study.finish_recovery(preview)
    rp, rn = ref.recover(study.displacements())
    assert np.array_equal(bits(pp), bits(rp)) and np.array_equal(bits(pn), bits(rn))
    assert study.solve().outcome == sf.StepOutcome.CONVERGED
    answer = study.begin_recovery()
    assert answer.is_converged()
    before = study.report().recovery_seconds
    assert before > 0.0
    while not study.continue_recovery(answer, 37):
        pass
    assert study.report().recovery_seconds > before
    ap, an = study.finish_recovery(answer)
    rp, rn = ref.recover(study.displacements())
    assert np.array_equal(bits(ap), bits(rp)) and np.array_equal(bits(an), bits(rn))
    # the one-call recovery is the same field
    op, on = study.recover()
    assert np.array_equal(bits(op), bits(ap)) and np.array_equal(bits(on), bits(an))
    # the recovery outlives the study it came from
    del study
    assert np.array_equal(bits(answer.field()[1]), bits(an))


def test_the_peak_of_a_recovered_field_is_the_oracle_peak(ctx):
    w = FIELDS["jittered-plate"]
    ref = refcpu.Problem(w)
    u = 1e-3 * np.random.default_rng(14).standard_normal((ref.node_count, 3))
    dp, dn = sf.Model.from_workload(w).recover(ctx, u)
    x = ref.nodes()
    points = dp.reshape(-1, 10)
    for inside in (lambda p: True, lambda p: p[0] >= 50.0, lambda p: p[0] > 5.0 and p[1] < 40.0):
        node_in = np.array([inside(p) for p in x])
        point_in = np.array([inside(p[:3]) for p in points])
        value, position, at_point, index = refcpu.peak(x, dn[:, 6], points, node_in, point_in)
        peak = sf.peak_stress(x, dn[:, 6], points, inside)
        assert bits(peak.von_mises) == bits(value) and np.array_equal(peak.position, position)
        assert (peak.at_point, peak.index) == (at_point, index)


# ---- nodal rotations (results.rs:262-305) ----


@pytest.mark.parametrize("name", list(FIELDS))
def test_nodal_rotations_are_bit_identical(ctx, name):
    w = FIELDS[name]
    ref = refcpu.Problem(w)
    u = 1e-3 * np.random.default_rng(15).standard_normal((ref.node_count, 3))
    r = sf.Model.from_workload(w).nodal_rotations(ctx, u)
    assert np.array_equal(bits(r), bits(ref.nodal_rotations(u)))


def test_nodal_rotations_of_a_rigid_rotation_are_its_angle(ctx):
    w = FIELDS["jittered-plate"]
    model = sf.Model.from_workload(w)
    x = refcpu.Problem(w).nodes()
    omega = np.array([1e-3, -2e-3, 5e-4])
    r = model.nodal_rotations(ctx, np.cross(omega, x))
    assert np.abs(r - omega).max() <= 1e-12 * np.abs(omega).max()
    with pytest.raises(ValueError):
        model.nodal_rotations(ctx, np.zeros((len(x) - 1, 3)))
