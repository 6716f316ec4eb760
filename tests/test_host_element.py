"""The element functions on the host (sf_shape_values .. sf_von_mises; crates/solidfem-elastic/src/
{shape,quadrature,element}.rs), pinned by:

* the reference's patch test (crates/solidfem-elastic/tests/element/patch.rs:17-76): a linear
  displacement field gives its constant strain at every node of every element of a distorted mesh,
  to 1e-11 of the largest strain component;
* the oracle's recovery: the stress at every integration point of a jittered plate, formed here
  from the element's nodes and displacements, is the oracle's bit for bit (the same operations);
* what the mathematics fixes: the shape functions are a partition of unity and interpolate (N_a at
  node b is exactly 1 or 0), their derivatives sum to zero, the rules' weights sum to one and
  integrate monomials of degree two (4 points) and five (14 points) exactly, and von Mises reads
  |sigma| in uniaxial stress and sqrt(3) |tau| in pure shear.
"""
import math

import numpy as np
import pytest

import refcpu
import solidfem_amd as sf
from workloads import plate_config

NODE_NATURAL = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.5, 0, 0], [0.5, 0.5, 0], [0, 0.5, 0], [0, 0, 0.5],
                         [0.5, 0, 0.5], [0, 0.5, 0.5]], dtype=float)
GRADIENT = np.array([[0.01, 0.002, -0.003], [0.004, -0.005, 0.006], [0.007, 0.001, 0.008]])
OFFSET = np.array([0.3, -0.2, 0.1])


def distorted_mesh():
    """patch.rs:17-38: a 3^3 box of side 3, interior vertices moved by 0.15 h sin(...)."""
    box = sf.box_tet_mesh((0.0, 0.0, 0.0), (3.0, 3.0, 3.0), (3, 3, 3))
    p = box.positions().copy()
    inner = ~np.any((p == 0.0) | (p == 3.0), axis=1)
    q = p[inner]
    p[inner] = q + 0.15 * np.sin(np.stack([13 * q[:, 0] + 7 * q[:, 1] + 3 * q[:, 2], 5 * q[:, 0] + 11 * q[:, 1] + 2 * q[:, 2],
                                           3 * q[:, 0] + 4 * q[:, 1] + 17 * q[:, 2]], axis=1))
    return sf.QuadraticTetMesh.from_linear(sf.TetMesh.new(p, box.tets()))


def test_a_linear_field_gives_constant_strain_on_a_distorted_mesh():
    mesh = distorted_mesh()
    x, tets = mesh.nodes(), mesh.tets()
    exact = np.array([GRADIENT[0, 0], GRADIENT[1, 1], GRADIENT[2, 2], GRADIENT[0, 1] + GRADIENT[1, 0],
                      GRADIENT[1, 2] + GRADIENT[2, 1], GRADIENT[2, 0] + GRADIENT[0, 2]])
    for tet in tets:
        nodes = x[tet]
        u = OFFSET + nodes @ GRADIENT.T
        for natural in NODE_NATURAL:
            strain = sf.strain_at(sf.point_geometry(nodes, natural), u)
            assert np.abs(strain - exact).max() < 1e-11 * 0.01


def test_point_stresses_are_the_oracles_bit_for_bit():
    w = plate_config(divisions=(6, 6, 2))
    ref = refcpu.Problem(w)
    x, tets = ref.nodes(), ref.tets()
    u = 1e-3 * np.random.default_rng(4).standard_normal((ref.node_count, 3))
    points, _ = ref.recover(u)
    rule = sf.quadrature_points(4)
    for e in range(0, len(tets), 7):
        for q in range(4):
            g = sf.point_geometry(x[tets[e]], rule[q, :3])
            stress = sf.stress_at(w.youngs, w.poisson, sf.strain_at(g, u[tets[e]]))
            assert np.array_equal(g.position, points[e, q, :3])
            assert np.array_equal(stress.view(np.uint64), points[e, q, 3:9].view(np.uint64))
            assert np.float64(sf.von_mises(stress)).view(np.uint64) == np.float64(points[e, q, 9]).view(np.uint64)


def test_shape_functions_interpolate_and_partition_unity():
    for b, natural in enumerate(NODE_NATURAL):
        assert np.array_equal(sf.shape_values(natural), np.eye(10)[b])
    rng = np.random.default_rng(6)
    for _ in range(20):
        n = rng.dirichlet(np.ones(4))[1:]
        assert abs(sf.shape_values(n).sum() - 1.0) < 1e-15
        assert np.abs(sf.shape_derivatives(n).sum(axis=0)).max() < 1e-14


@pytest.mark.parametrize("rule,degree", [(4, 2), (14, 5)])
def test_rules_integrate_their_degree_exactly(rule, degree):
    pts = sf.quadrature_points(rule)
    assert len(pts) == rule and abs(pts[:, 3].sum() - 1.0) < 1e-15
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            for c in range(degree + 1 - a - b):
                exact = 6.0 * math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)
                got = np.sum(pts[:, 3] * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c)
                assert abs(got - exact) < 1e-14, (a, b, c)
    with pytest.raises(ValueError):
        sf.quadrature_points(5)


def test_an_inverted_element_is_refused_and_von_mises_reads_its_closed_forms():
    nodes = np.array([x for x in NODE_NATURAL])
    nodes[[1, 2]] = nodes[[2, 1]]  # two corners swapped: negative volume
    nodes[4:] = 0.5 * (nodes[[0, 1, 2, 0, 1, 2]] + nodes[[1, 2, 0, 3, 3, 3]])
    with pytest.raises(sf.ModelError) as e:
        sf.point_geometry(nodes, [0.25, 0.25, 0.25])
    assert e.value.status == 45
    assert sf.von_mises([-7.0, 0, 0, 0, 0, 0]) == 7.0
    assert abs(sf.von_mises([0, 0, 0, 2.0, 0, 0]) - 2.0 * math.sqrt(3.0)) < 1e-15
