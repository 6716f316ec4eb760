"""The oracle's stress recovery (results.rs:307-576, element.rs:135-192, patch.rs) pinned on the CPU.

Closed forms: an affine displacement field has constant strain, so every integration point and
every recovered node must read the constitutive stress of that strain; a quadratic field has a
linear stress, which Tet10 represents exactly and the linear patch fit must reproduce at every
node. An independent numpy evaluation of the Tet10 gradients checks the point stresses and the
worst-value fallback where a patch holds two materials. The nodal rotations (results.rs:262-305)
are pinned by a rigid rotation and a linear field (their skew part, everywhere), a translation
(zero), and an independent numpy evaluation of the determinant-weighted mean on a cubic field, where
the elements disagree at shared nodes and an unweighted mean differs.
"""
import dataclasses

import numpy as np
import pytest

import refcpu
from workloads import bar_config, block_config, plate_config

# node natural coordinates (shape.rs:9-20) and the 4-point rule (quadrature.rs:44-51)
NATURAL = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0], [0.5, 0.0, 0.0],
                    [0.5, 0.5, 0.0], [0.0, 0.5, 0.0], [0.0, 0.0, 0.5], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
A4, B4 = 0.1381966011250105, 0.5854101966249685
RULE4 = np.array([[A4, A4, A4], [B4, A4, A4], [A4, B4, A4], [A4, A4, B4]])
EDGES = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]


def lame(e, nu):
    return e * nu / ((1 + nu) * (1 - 2 * nu)), e / (2 * (1 + nu))


def stress_of(grad_u, e, nu):
    """sigma = lambda tr(eps) I + 2 mu eps, as [xx, yy, zz, xy, yz, zx]; grad_u[i][j] = du_i/dx_j."""
    lam, mu = lame(e, nu)
    eps = 0.5 * (grad_u + grad_u.T)
    s = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    return np.array([s[0, 0], s[1, 1], s[2, 2], s[0, 1], s[1, 2], s[2, 0]])


def von_mises(s):
    a, b, c = s[0] - s[1], s[1] - s[2], s[2] - s[0]
    return np.sqrt(0.5 * (a * a + b * b + c * c) + 3 * (s[3] ** 2 + s[4] ** 2 + s[5] ** 2))


def tet10_gradients(x, nat):
    """Physical gradients (10 x 3) of the P2 shape functions at a natural point: an independent
    numpy form (barycentric derivatives, then J^-1)."""
    r, s, t = nat
    l = np.array([1 - r - s - t, r, s, t])
    dl = np.array([[-1.0, -1.0, -1.0], [1.0, 0, 0], [0, 1.0, 0], [0, 0, 1.0]])
    d = np.zeros((10, 3))
    for i in range(4):
        d[i] = (4 * l[i] - 1) * dl[i]
    for k, (a, b) in enumerate(EDGES):
        d[4 + k] = 4 * (dl[a] * l[b] + l[a] * dl[b])
    jac = d.T @ x  # J[i][k] = sum_a dN_a/dxi_i x_a,k
    return np.linalg.solve(jac, d.T).T


@pytest.mark.parametrize("w", [plate_config(divisions=(4, 4, 2)), bar_config(divisions=(6, 2, 2))], ids=["jittered", "bar"])
def test_affine_field_recovers_the_constant_stress_everywhere(w):
    w = dataclasses.replace(w, material_kind=0, youngs=2000.0, poisson=0.3)
    p = refcpu.Problem(w)
    g = np.array([[1e-3, 2e-4, -3e-4], [5e-4, -2e-3, 1e-4], [-4e-4, 3e-4, 7e-4]])
    u = p.nodes() @ g.T + np.array([0.1, -0.2, 0.3])
    points, nodal = p.recover(u)
    s = stress_of(g, 2000.0, 0.3)
    scale = np.abs(s).max()
    assert np.abs(points[..., 3:9] - s).max() <= 1e-10 * scale
    assert np.abs(points[..., 9] - von_mises(s)).max() <= 1e-10 * scale
    assert np.abs(nodal[:, :6] - s).max() <= 1e-9 * scale
    assert np.abs(nodal[:, 6] - von_mises(s)).max() <= 1e-9 * scale


def test_quadratic_field_recovers_its_linear_stress_at_every_node():
    # u = (a x y, b x^2 + c z^2, d y z): grad u is linear, so is the stress; the linear patch fit
    # of a linear field is exact at vertices and midsides alike
    w = dataclasses.replace(block_config(0, divisions=(4, 3, 3)), youngs=1000.0, poisson=0.25)
    p = refcpu.Problem(w)
    x = p.nodes()
    a, b, c, d = 2e-4, -1e-4, 3e-4, 1.5e-4
    u = np.stack([a * x[:, 0] * x[:, 1], b * x[:, 0] ** 2 + c * x[:, 2] ** 2, d * x[:, 1] * x[:, 2]], axis=1)
    points, nodal = p.recover(u)

    def grad_at(pt):
        px, py, pz = pt
        return np.array([[a * py, a * px, 0.0], [2 * b * px, 0.0, 2 * c * pz], [0.0, d * pz, d * py]])

    scale = np.abs(nodal[:, :6]).max()
    for q in points.reshape(-1, 10)[::7]:
        assert np.abs(q[3:9] - stress_of(grad_at(q[:3]), 1000.0, 0.25)).max() <= 1e-9 * scale
    for n in range(len(x)):
        assert np.abs(nodal[n, :6] - stress_of(grad_at(x[n]), 1000.0, 0.25)).max() <= 1e-8 * scale, n


def test_points_and_the_two_material_fallback_match_an_independent_evaluation():
    # a shell/core bar: a patch that straddles the interface is not fitted, and its node reads the
    # largest von Mises among its elements' values there, each with the element's own material
    # (that of the element's integration point nearest the node)
    w = bar_config(divisions=(4, 4, 4))
    p = refcpu.Problem(w)
    x, tets = p.nodes(), p.tets()
    u = 1e-3 * np.random.default_rng(4).standard_normal(x.shape)
    points, nodal = p.recover(u)
    assert points.shape[1] == 4

    def material(pos):  # the shell is 2 mm under the four long faces (y, z in [0, 20])
        shell = min(pos[1], 20 - pos[1], pos[2], 20 - pos[2]) < 2.0
        return (3500.0, 0.36) if shell else (700.0, 0.3)

    nearest = [int(np.argmin(((RULE4 - NATURAL[n]) ** 2).sum(axis=1))) for n in range(10)]
    worst = {}
    for e, t in enumerate(tets):
        xe, ue = x[t], u[t]
        mats = []
        for q in range(4):
            gr = tet10_gradients(xe, RULE4[q])
            mats.append(material(points[e, q, :3]))
            s = stress_of(ue.T @ gr, *mats[-1])
            assert np.abs(points[e, q, 3:9] - s).max() <= 1e-9 * max(1.0, np.abs(s).max()), (e, q)
        for n in range(10):
            s = stress_of(ue.T @ tet10_gradients(xe, NATURAL[n]), *mats[nearest[n]])
            vm = von_mises(s)
            if t[n] not in worst or vm > worst[t[n]][0]:
                worst[t[n]] = (vm, s)
    vertices = p.vertex_count
    incident = [[] for _ in range(vertices)]
    for e, t in enumerate(tets):
        for v in t[:4]:
            incident[v].append(e)
    mixed = 0
    for v in range(vertices):
        if len({material(points[e, q, :3]) for e in incident[v] for q in range(4)}) > 1:
            mixed += 1
            assert np.abs(nodal[v, :6] - worst[v][1]).max() <= 1e-9 * max(1.0, np.abs(worst[v][1]).max()), v
    assert mixed > 0


# ---- nodal rotations (results.rs:262-305) ----


def rotation_of(grad_u):
    """Half the curl of u from grad_u[i][j] = du_i/dx_j."""
    return 0.5 * np.array([grad_u[2, 1] - grad_u[1, 2], grad_u[0, 2] - grad_u[2, 0], grad_u[1, 0] - grad_u[0, 1]])


def test_rigid_and_linear_fields_rotate_by_their_skew_part():
    w = plate_config(divisions=(4, 4, 2))
    p = refcpu.Problem(w)
    x = p.nodes()
    omega = np.array([2e-3, -5e-4, 1e-3])
    rigid = np.cross(omega, x) + np.array([0.3, -0.1, 0.2])
    r = p.nodal_rotations(rigid)
    assert np.abs(r - omega).max() <= 1e-13 * np.abs(omega).max() * 100
    a = np.array([[1e-3, 2e-4, -3e-4], [5e-4, -2e-3, 1e-4], [-4e-4, 3e-4, 7e-4]])
    r = p.nodal_rotations(x @ a.T)
    assert np.abs(r - rotation_of(a)).max() <= 1e-11 * np.abs(a).max()
    # a translation: the gradients sum to zero up to rounding, of order eps |u| / h
    assert np.abs(p.nodal_rotations(np.tile([1.0, 2.0, 3.0], (len(x), 1)))).max() <= 1e-13


def test_rotations_are_the_volume_weighted_mean_of_the_elements_values():
    # a cubic field is outside the Tet10 space: each element's gradient at a shared node differs,
    # so the determinant weighting (six times the volume of a straight-edged element) shows
    w = plate_config(divisions=(3, 3, 2))
    p = refcpu.Problem(w)
    x, tets = p.nodes(), p.tets()
    u = np.stack([x[:, 1] ** 3 * 1e-4, x[:, 0] ** 3 * 2e-5 + x[:, 2] ** 2 * 1e-3, x[:, 0] * x[:, 1] * x[:, 2] * 1e-4], axis=1)
    weighted, weights = np.zeros((len(x), 3)), np.zeros(len(x))
    for t in tets:
        xe, ue = x[t], u[t]
        det = np.linalg.det(np.array([xe[1] - xe[0], xe[2] - xe[0], xe[3] - xe[0]]))
        for n in range(10):
            weighted[t[n]] += det * rotation_of(ue.T @ tet10_gradients(xe, NATURAL[n]))
            weights[t[n]] += det
    expected = weighted / weights[:, None]
    r = p.nodal_rotations(u)
    assert np.abs(r - expected).max() <= 1e-12 * np.abs(expected).max()
    # the unweighted mean is a different answer on this jittered mesh
    plain, counts = np.zeros((len(x), 3)), np.zeros(len(x))
    for t in tets:
        for n in range(10):
            plain[t[n]] += rotation_of(u[t].T @ tet10_gradients(x[t], NATURAL[n]))
            counts[t[n]] += 1
    assert np.abs(plain / counts[:, None] - expected).max() > 1e-6 * np.abs(expected).max()
