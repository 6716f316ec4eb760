"""Prescribed displacement components through the device study, against the oracle.

The stretched unit cube of crates/solidfem/tests/validation/prescribed.rs:40-114: the eliminated
system (StudyBuilder::apply_prescribed, crates/solidfem-elastic/src/study.rs:567-593) and the solve
in the reference's order are the oracle's to the bit, and the answer meets the closed forms --
prescribed values exact, the uniform stress E * stretch along x, the transverse contraction
-nu * stretch.
"""
import numpy as np
import pytest

import refcpu
import solidfem_amd as sf
from workloads import STRETCH, stretch_components, stretched_cube

pytestmark = pytest.mark.gpu
E, NU = 200000.0, 0.3


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("divisions", [(1, 1, 1), (3, 3, 3)])
def test_a_stretched_cube_is_the_oracles_and_the_closed_forms(ctx, divisions):
    w = stretched_cube(divisions)
    model = sf.Model.from_workload(w)
    nodes = model.mesh().nodes()
    comps = stretch_components(nodes)
    for n, a, v in comps:
        model.prescribe_displacement(n, a, v)
    options = sf.MultigridOptions(precision=sf.Precision.DOUBLE, coarse_dofs=100)
    builder = sf.StudyBuilder.new(ctx, model, sf.PcgOptions(tolerance=1e-12, dot_order=sf.DotOrder.REFERENCE), options)
    builder.step(2 ** 62)
    study = builder.finish()
    ref = refcpu.Problem(w, prescribed=comps)
    # the eliminated system
    rs, cols, vals = study.matrix()
    r_rs, r_cols, r_vals = ref.csr()
    assert np.array_equal(rs, r_rs) and np.array_equal(cols, r_cols)
    assert np.array_equal(bits(vals), bits(r_vals))
    assert np.array_equal(bits(study.rhs()), bits(ref.rhs()))
    # the solve, in the reference's order
    progress = study.solve()
    disp = study.displacements()
    r_disp, r_res = refcpu.Multigrid(ref, refcpu.options(precision=0, coarse_dofs=100)).solve(1e-12)
    assert progress.outcome == sf.StepOutcome.CONVERGED and r_res.status == 0
    assert progress.iteration == r_res.iterations
    assert np.array_equal(bits(disp), bits(r_disp))
    # closed forms
    for n, p in enumerate(nodes):
        if p[0] == 1.0:
            assert disp[n, 0] == STRETCH
        if p[0] == 0.0:
            assert disp[n, 0] == 0.0
    corner = next(n for n, p in enumerate(nodes) if tuple(p) == (1.0, 1.0, 1.0))
    assert abs(disp[corner, 1] + NU * STRETCH) < 1e-6 * STRETCH
    points, _ = study.recover()
    stress = points[:, :, 3:9].reshape(-1, 6)
    expected = E * STRETCH
    assert np.all(np.abs(stress[:, 0] - expected) < 1e-5 * expected)
    assert np.all(np.abs(stress[:, 1:]) < 1e-5 * expected)
