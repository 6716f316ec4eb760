"""A caller's MaterialField (crates/solidfem-elastic/src/material.rs:35-50) through
sf_model_set_material_field, on the GPU.

The field is evaluated once per integration point at the position point_geometry forms, so:

* a callback restating the shell-core field gives the built-in shell-core model's matrix,
  recovery and solve bit for bit, and a constant callback the uniform model's matrix and error
  estimate -- those built-in fields are the oracle's to the bit (test_gpu_parity.py);
* a field split at x = 1 on cells aligned with it gives every element the stiffness of the uniform
  model of its own material, bit for bit, and refuses an error estimate by a mixed-material vertex
  (crates/solidfem-elastic/tests/error_estimate/exact.rs:61-77);
* an invalid material is refused by name at the first element and point that meets it, and a
  callback that raises is reported.
"""
import dataclasses

import numpy as np
import pytest

import solidfem_amd as sf
from workloads import bar_config

pytestmark = pytest.mark.gpu

INVALID_MATERIAL, MIXED_MATERIAL_PATCH = 46, 53


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def shell_core(w):
    """The built-in shell-core field of a workload, restated: the depth below the two long-face
    pairs across the longest axis, the core where it is not below the shell depth."""
    lo, hi = np.array(w.origin), np.array(w.origin) + np.array(w.size)
    longest = int(np.argmax(w.size))
    axes = ((longest + 1) % 3, (longest + 2) % 3)

    def at(p):
        depth = 1e300
        for ax in axes:
            depth = min(depth, min(p[ax] - lo[ax], hi[ax] - p[ax]))
        return (w.youngs, w.poisson) if depth < w.shell_depth else (w.core_youngs, w.core_poisson)
    return at


def study_of(ctx, model):
    study = sf.StaticStudy.new(ctx, model, sf.PcgOptions(tolerance=1e-10, dot_order=sf.DotOrder.REFERENCE),
                               sf.MultigridOptions(precision=sf.Precision.DOUBLE, coarse_dofs=100))
    return study


@pytest.mark.parametrize("points", [4, 14])
def test_a_restated_shell_core_field_is_the_built_in_one(ctx, points):
    w = dataclasses.replace(bar_config(divisions=(12, 4, 4)), quadrature_points=points)
    built_in = sf.Model.from_workload(w)
    caller = sf.Model.from_workload(w)
    caller.set_material_field(shell_core(w))
    a, b = study_of(ctx, built_in), study_of(ctx, caller)
    for x, y in zip(a.matrix(), b.matrix()):
        assert np.array_equal(bits(x), bits(y))
    pa, pb = a.solve(), b.solve()
    assert pa.outcome == pb.outcome == sf.StepOutcome.CONVERGED and pa.iteration == pb.iteration
    assert np.array_equal(bits(a.displacements()), bits(b.displacements()))
    for x, y in zip(a.recover(), b.recover()):
        assert np.array_equal(bits(x), bits(y))


def test_a_constant_field_is_the_uniform_model(ctx):
    w = bar_config(divisions=(12, 4, 4))
    w.material_kind = 0
    uniform = sf.Model.from_workload(w)
    caller = sf.Model.from_workload(w)
    caller.set_material_field(lambda p: (w.youngs, w.poisson))
    a, b = study_of(ctx, uniform), study_of(ctx, caller)
    for x, y in zip(a.matrix(), b.matrix()):
        assert np.array_equal(bits(x), bits(y))
    a.solve(), b.solve()
    (_, ea), (_, eb) = a.solution_with_error_estimate(), b.solution_with_error_estimate()
    assert np.array_equal(bits(ea.element_errors), bits(eb.element_errors))
    assert bits(ea.relative_error) == bits(eb.relative_error)


def test_a_field_split_at_a_cell_boundary_gives_each_element_its_materials_stiffness(ctx):
    steel, soft = (2100.0, 0.3), (700.0, 0.3)
    w = sf.Workload(size=(3.0, 2.0, 2.0), divisions=(3, 2, 2), clamp_axis=-1, load_axis=-1, youngs=steel[0],
                    poisson=steel[1])
    split = sf.Model.from_workload(w)
    split.set_material_field(lambda p: steel if p[0] < 1.0 else soft)
    uniform = {m: sf.Model.from_workload(dataclasses.replace(w, youngs=m[0], poisson=m[1])) for m in (steel, soft)}
    elements = len(split.mesh().tets())
    ke = split.element_stiffness(ctx, 0, elements)
    ke_of = {m: uniform[m].element_stiffness(ctx, 0, elements) for m in (steel, soft)}
    x, tets = split.mesh().nodes(), split.mesh().tets()
    for e in range(elements):
        m = steel if x[tets[e][:4], 0].max() <= 1.0 else soft
        assert np.array_equal(bits(ke[e]), bits(ke_of[m][e])), e
    # held at the bending field of exact.rs, the two-material patches refuse an estimate
    kappa, nu = 1e-3, 0.3
    for n, p in enumerate(x):
        if any(p[a] == 0.0 or p[a] == w.size[a] for a in range(3)):
            value = [kappa * p[0] * p[1], -0.5 * kappa * (p[0] ** 2 + nu * (p[1] ** 2 - p[2] ** 2)), -nu * kappa * p[1] * p[2]]
            for a in range(3):
                split.prescribe_displacement(n, a, value[a])
    study = sf.StaticStudy.new(ctx, split, sf.PcgOptions(tolerance=1e-14))
    assert study.solve().outcome == sf.StepOutcome.CONVERGED
    with pytest.raises(sf.ModelError) as e:
        study.solution_with_error_estimate()
    assert e.value.status == MIXED_MATERIAL_PATCH


def test_an_invalid_material_is_refused_at_its_element_and_point(ctx):
    w = sf.Workload(size=(3.0, 2.0, 2.0), divisions=(3, 2, 2), clamp_axis=0, load_axis=0, load_value=3.0,
                    traction=(0.0, 0.0, -1.0), youngs=2100.0, poisson=0.3)
    model = sf.Model.from_workload(w)
    model.set_material_field(lambda p: (2100.0, 0.3) if p[0] < 2.0 else (-1.0, 0.3))
    x, tets = model.mesh().nodes(), model.mesh().tets()
    rule = sf.quadrature_points(4)
    first = next((e, q) for e in range(len(tets)) for q in range(4)
                 if sf.point_geometry(x[tets[e]], rule[q, :3]).position[0] >= 2.0)
    with pytest.raises(sf.ModelError) as e:
        sf.StaticStudy.new(ctx, model)
    assert e.value.status == INVALID_MATERIAL and (e.value.a, e.value.b) == first


def test_a_field_that_raises_is_reported():
    model = sf.Model.from_workload(sf.Workload(size=(1.0, 1.0, 1.0), divisions=(1, 1, 1)))

    def broken(p):
        raise KeyError("material")

    with pytest.raises(KeyError):
        model.set_material_field(broken)


def test_a_recovery_refuses_an_invalid_material_at_its_element_and_point(ctx):
    w = sf.Workload(size=(3.0, 2.0, 2.0), divisions=(3, 2, 2), clamp_axis=0, youngs=2100.0, poisson=0.3)
    model = sf.Model.from_workload(w)
    model.set_material_field(lambda p: (2100.0, 0.3) if p[0] < 2.0 else (2100.0, 0.6))
    x, tets = model.mesh().nodes(), model.mesh().tets()
    rule = sf.quadrature_points(4)
    first = next((e, q) for e in range(len(tets)) for q in range(4)
                 if sf.point_geometry(x[tets[e]], rule[q, :3]).position[0] >= 2.0)
    u = np.zeros((len(x), 3))
    with pytest.raises(sf.ModelError) as e:
        model.recover(ctx, u)
    assert e.value.status == INVALID_MATERIAL and (e.value.a, e.value.b) == first
    recovery = model.begin_recovery(ctx, u, converged=True)
    with pytest.raises(sf.ModelError) as e:
        recovery.advance(2 ** 62)
    assert e.value.status == INVALID_MATERIAL and recovery.elements_done() == first[0]
