"""Mesh comparisons of device solves and recoveries (crates/solidfem/tests/validation/comparison.rs):
the reference's cantilevers solved and recovered on the GPU and compared on the host read what the
reference's own test reads -- the peaks' holders, exact element and volume ratios, the verdicts --
and so do its graded cantilevers, meshes of mapped box vertices that only the product builds here
(the oracle takes box workloads): a fine mesh refined only outside the region is refused, one
halved in x reads 32/7 at the peak and 8 over the whole part, and one refined away from its peak
reads too alike.
"""
import math

import numpy as np
import pytest

import solidfem_amd as sf

pytestmark = pytest.mark.gpu

CANTILEVER = (4.0, 1.0, 1.0)
STEEL = (1000.0, 0.3)


def away(p):
    return p[0] >= 2.0 - 1e-9


def solve(ctx, model):
    study = sf.StaticStudy.new(ctx, model, sf.PcgOptions(tolerance=1e-9), sf.MultigridOptions(coarse_dofs=100))
    assert study.solve().outcome == sf.StepOutcome.CONVERGED
    points, nodal = study.recover()
    return (model.mesh(), nodal, points)


def clamped(mesh, length):
    """validation/comparison.rs:24-38 with support.rs:19-39: every node of x = 0 fixed, a unit
    downward traction (load / area, area 1) on the faces of x = length."""
    model = sf.Model.new(mesh, *STEEL)
    x = mesh.nodes()
    for n in np.nonzero(np.abs(x[:, 0]) < 1e-9)[0]:
        model.fix_node(int(n))
    for face in mesh.boundary_faces():
        if all(abs(x[face[k], 0] - length) < 1e-9 for k in (2, 3, 4)):
            model.add_traction(face, (0.0, 0.0, -1.0))
    return model


def box_cantilever(ctx, cells):
    mesh = sf.QuadraticTetMesh.from_linear(sf.box_tet_mesh((0.0, 0.0, 0.0), CANTILEVER, (4 * cells, cells, cells)))
    return solve(ctx, clamped(mesh, CANTILEVER[0]))


def graded_cantilever(ctx, widths):
    """validation/comparison.rs:316-340: unit box cells moved to the planes of the given widths."""
    planes = np.concatenate([[0.0], np.cumsum(widths)])
    box = sf.box_tet_mesh((0.0, 0.0, 0.0), (float(len(widths)), 1.0, 1.0), (len(widths), 1, 1))
    p = box.positions().copy()
    p[:, 0] = planes[np.rint(p[:, 0]).astype(int)]
    mesh = sf.QuadraticTetMesh.from_linear(sf.TetMesh.new(p, box.tets()))
    return solve(ctx, clamped(mesh, float(planes[-1])))


def test_device_cantilevers_compare_as_the_reference_reads_them(ctx):
    meshes = [box_cantilever(ctx, n) for n in (1, 2, 4, 8)]
    changes = []
    for a, b in zip(meshes, meshes[1:]):
        c = sf.MeshComparison.between(a, b, away)
        assert c.peak_change() < 0.0
        assert np.array_equal(c.coarse().peak.position, [2.0, 1.0, 1.0])
        assert np.array_equal(c.fine().peak.position, [2.0, 1.0, 1.0]) and c.peak_distance() == 0.0
        assert c.element_ratio() == 8.0 and c.peak_element_ratio() == 8.0
        assert c.verdict(math.nan) == sf.Verdict.STILL_MOVING
        changes.append(c.peak_change())
        whole = sf.MeshComparison.between(a, b)
        assert whole.peak_change() > 0.0 and whole.verdict(0.05) == sf.Verdict.STILL_MOVING
    assert all(abs(b) < 0.5 * abs(a) for a, b in zip(changes, changes[1:]))
    first = sf.MeshComparison.between(meshes[0], meshes[1])
    assert np.array_equal(first.coarse().peak.position, [0.5, 1.0, 0.0])
    assert np.array_equal(first.fine().peak.position, [0.25, 1.0, 0.0]) and first.peak_distance() == 0.25


def test_a_mesh_refined_only_outside_the_region_read_is_not_finer_there(ctx):
    coarse = graded_cantilever(ctx, [1.0] * 4)
    same_inside = graded_cantilever(ctx, [0.125] * 16 + [1.0, 1.0])
    twice_inside = graded_cantilever(ctx, [0.125] * 16 + [0.5] * 4)
    assert [len(m[0].tets()) for m in (coarse, same_inside, twice_inside)] == [24, 108, 120]
    with pytest.raises(sf.ModelError) as e:
        sf.MeshComparison.between(coarse, same_inside, away)
    assert e.value.status == 72 and (e.value.a, e.value.b) == (12, 12)
    in_region = sf.MeshComparison.between(coarse, twice_inside, away)
    assert in_region.element_ratio() == 2.0
    assert np.array_equal(in_region.fine().peak.position, [2.0, 1.0, 1.0])
    assert abs(in_region.peak_element_ratio() - 32.0 / 7.0) / (32.0 / 7.0) < 1e-12
    assert in_region.verdict(math.nan) == sf.Verdict.STILL_MOVING
    whole = sf.MeshComparison.between(coarse, twice_inside)
    assert whole.element_ratio() == 5.0
    assert np.array_equal(whole.fine().peak.position, [0.3125, 1.0, 0.0])
    assert whole.peak_element_ratio() == 8.0 and whole.verdict(math.nan) == sf.Verdict.STILL_MOVING


def test_a_region_refined_away_from_its_peak_reads_too_alike(ctx):
    coarse = graded_cantilever(ctx, [1.0] * 4)
    fine = graded_cantilever(ctx, [1.0] * 3 + [0.125] * 8)
    c = sf.MeshComparison.between(coarse, fine, away)
    assert (c.coarse().elements, c.fine().elements) == (12, 54)
    assert c.element_ratio() >= sf.MIN_ELEMENT_RATIO
    assert np.array_equal(c.coarse().peak.position, [2.0, 1.0, 1.0])
    assert np.array_equal(c.fine().peak.position, [2.0, 1.0, 1.0])
    assert c.peak_element_ratio() == 1.0
    assert abs(c.peak_change()) <= 0.01 and c.verdict(0.01) == sf.Verdict.MESHES_TOO_ALIKE
