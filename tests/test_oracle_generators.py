"""The oracle's workload generators pinned by what the mathematics fixes (CPU only).

The product builds the same inputs bit for bit (tests/test_host_parity.py), so these checks pin
both. Each is chosen so that a plausible mistake fails it:

* Jitter (interior vertices moved by U(-5%, +5%) of the cell per axis, before the Tet10 upgrade):
  boundary vertices unmoved; every offset within 5% of its cell size; offsets with mean 0 and
  variance (0.05 h)^2 / 3 of a uniform law, uncorrelated across axes; midside nodes still the
  midpoints of their edges. Jittering the boundary, a normal instead of a uniform draw, jitter
  after the upgrade or one draw for all axes fail here.
* Random tractions (each component N(0, 1 MPa) per face and load case): mean 0, variance 1 and
  kurtosis 3 of a normal law (a uniform law has 1.8); uncorrelated across components and across
  load cases.
* Torsion traction (0.5 MPa along e_theta about the bar axis at each face's corner centroid):
  magnitude 0.5 and perpendicular to both the radius and the axis; zero resultant force; the
  consistent nodal loads' resultant and torque equal the faces' sum of area * t and
  area * (r x t), as crates/solidfem/tests/validation/torsion.rs:35-61 forms the torque; and the
  twist rate of a uniform square bar away from its ends matches Saint-Venant's
  (torsion.rs:21-77) within 1%.
"""
import numpy as np
import pytest

import refcpu
from solidfem_amd import Workload
from workloads import bar_config, block_config, plate_config


def test_jitter_moves_interior_vertices_by_a_uniform_law_before_the_upgrade():
    div = (30, 30, 4)
    jittered = refcpu.Problem(plate_config(divisions=div))
    plain = refcpu.Problem(Workload(**{**plate_config(divisions=div).__dict__, "jitter_fraction": 0.0}))
    nv = jittered.vertex_count
    p0, p1 = plain.nodes(), jittered.nodes()
    off = (p1 - p0)[:nv]
    h = np.array([100.0 / div[0], 100.0 / div[1], 0.8 / div[2]])
    grid = np.rint((p0[:nv] - 0.0) / h).astype(int)
    interior = np.all((grid > 0) & (grid < np.array(div)), axis=1)
    assert np.count_nonzero(interior) == np.prod(np.array(div) - 1)
    assert not np.any(off[~interior]), "a boundary vertex moved"
    u = off[interior] / (0.05 * h)  # U(-1, 1) per axis
    n = len(u)
    assert np.all(np.abs(u) <= 1.0)
    assert np.all(np.abs(u.mean(axis=0)) < 4 * np.sqrt(1 / 3 / n))
    assert np.all(np.abs(u.var(axis=0) - 1 / 3) < 4 * np.sqrt(4 / 45 / n))
    corr = np.corrcoef(u.T)
    assert np.all(np.abs(corr[np.triu_indices(3, 1)]) < 4 / np.sqrt(n))
    # a uniform law's fourth moment is 1/5 (a normal of the same variance gives 1/3)
    assert np.all(np.abs((u ** 4).mean(axis=0) - 0.2) < 4 * np.sqrt((1 / 9 - 1 / 25) / n))
    # midside nodes are midpoints of their edges in the jittered mesh
    tets = jittered.tets()
    edges = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
    for k, (a, b) in enumerate(edges):
        mid = p1[tets[:, 4 + k]]
        assert np.allclose(mid, 0.5 * (p1[tets[:, a]] + p1[tets[:, b]]), rtol=0, atol=1e-12)


def test_random_tractions_are_independent_standard_normals():
    # the 30 x 20 top faces of the block (its depth does not change them)
    cases = [refcpu.Problem(block_config(c, divisions=(30, 20, 1))).loaded_faces()[1] for c in range(8)]
    t = np.stack(cases)  # (case, face, component)
    x = t.reshape(-1, 3)
    n = len(x)
    assert np.all(np.abs(x.mean(axis=0)) < 4 / np.sqrt(n))
    assert np.all(np.abs(x.var(axis=0) - 1.0) < 4 * np.sqrt(2 / n))
    kurt = (x ** 4).mean(axis=0) / x.var(axis=0) ** 2
    assert np.all(np.abs(kurt - 3.0) < 4 * np.sqrt(24 / n))
    corr = np.corrcoef(x.T)
    assert np.all(np.abs(corr[np.triu_indices(3, 1)]) < 4 / np.sqrt(n))
    per_case = t.reshape(8, -1)
    cc = np.corrcoef(per_case)
    assert np.all(np.abs(cc[np.triu_indices(8, 1)]) < 4 / np.sqrt(per_case.shape[1]))


def face_geometry(problem, faces):
    nodes, tets = problem.nodes(), problem.tets()
    corner = {0: (1, 2, 3), 1: (0, 3, 2), 2: (0, 1, 3), 3: (0, 2, 1)}
    out = []
    for tet, face in faces:
        a, b, c = (nodes[tets[tet, k]] for k in corner[int(face)])
        area = 0.5 * np.linalg.norm(np.cross(b - a, c - a))
        out.append((area, (a + b + c) / 3.0))
    return out


def torsion_loads(w):
    p = refcpu.Problem(w)
    faces, t = p.loaded_faces()
    geo = face_geometry(p, faces)
    return p, t, geo


def test_torsion_traction_is_tangential_and_its_loads_carry_the_faces_torque():
    w = bar_config(divisions=(16, 4, 4))
    p, t, geo = torsion_loads(w)
    axis = np.array([10.0, 10.0])
    area = np.array([g[0] for g in geo])
    r = np.array([g[1][1:] - axis for g in geo])
    assert np.allclose(np.linalg.norm(t, axis=1), 0.5, rtol=1e-12)
    assert np.allclose(t[:, 0], 0.0)
    assert np.allclose(np.einsum("ij,ij->i", r, t[:, 1:]), 0.0, atol=1e-12)
    assert np.all(r[:, 0] * t[:, 2] - r[:, 1] * t[:, 1] > 0)  # one sense of rotation
    force = (area[:, None] * t).sum(axis=0)
    torque = float((area * (r[:, 0] * t[:, 2] - r[:, 1] * t[:, 1])).sum())
    assert np.allclose(force, 0.0, atol=1e-9 * torque)
    # the consistent nodal loads carry the same resultant and torque (a flat face's area centroid
    # is its corner centroid, where the per-face traction is evaluated)
    rhs = p.rhs().reshape(-1, 3)
    fm = p.free_map()
    pos = p.nodes()[fm != 0xFFFFFFFF]
    assert np.allclose(rhs.sum(axis=0), force, atol=1e-9 * torque)
    arm = pos[:, 1:] - axis
    nodal_torque = float((arm[:, 0] * rhs[:, 2] - arm[:, 1] * rhs[:, 1]).sum())
    assert abs(nodal_torque - torque) < 1e-12 * torque


def saint_venant_beta():
    k = 2 * np.arange(200) + 1.0
    return (1.0 - 192.0 / np.pi ** 5 * np.sum(np.tanh(k * np.pi / 2.0) / k ** 5)) / 3.0


def twist_rate(w, disp, nodes, side):
    centre = np.array([w.axis_u, w.axis_v])

    def twist(x):
        at = np.abs(nodes[:, 0] - x) < 1e-9
        y, z = (nodes[at, 1:] - centre).T
        u = disp[at]
        return float((y * u[:, 2] - z * u[:, 1]).sum() / (y * y + z * z).sum())

    length = w.size[0]
    x1, x2 = 0.3 * length, 0.7 * length
    return (twist(x2) - twist(x1)) / (x2 - x1)


def test_a_uniform_square_bar_twists_at_the_saint_venant_rate():
    # crates/solidfem/tests/validation/torsion.rs:21-77 at 20x its size: end effects three sides
    # from each end decay below exp(-3 pi); four quadratic cells across resolve the warping
    w = Workload(**{**bar_config(divisions=(20, 4, 4)).__dict__, "material_kind": 0, "youngs": 2100.0, "poisson": 0.3})
    p, t, geo = torsion_loads(w)
    axis = np.array([w.axis_u, w.axis_v])
    torque = sum(g[0] * ((g[1][1] - axis[0]) * tt[2] - (g[1][2] - axis[1]) * tt[1]) for g, tt in zip(geo, t))
    mg = refcpu.Multigrid(p, refcpu.options(precision=0, coarse_dofs=100))
    disp, res = mg.solve(1e-10)
    assert res.status == 0
    side = w.size[1]
    g = w.youngs / (2 * (1 + w.poisson))
    expected = torque / (g * saint_venant_beta() * side ** 4)
    rate = twist_rate(w, disp, p.nodes(), side)
    assert abs(rate - expected) < 0.01 * expected, (rate, expected)
