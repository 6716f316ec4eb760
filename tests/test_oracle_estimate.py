"""The oracle's error estimate (crates/solidfem-elastic/src/error_estimate.rs) pinned on the CPU by
the reference's own tests and by closed forms:

* exact.rs:27-59: pure bending, which 10-node elements hold exactly, held at every boundary node
  and solved to the rounding floor, reads as converged (eta and every element's error within
  10 (1e2 rho + 1e-13)); and its solution norm is the closed form E kappa^2 int y^2 dV.
* exact.rs:61-77: a patch of two materials refuses the estimate (the shell-core bar here).
* convergence.rs:71-117: on the harmonic field u = grad(e^x cos(y/sqrt2) cos(z/sqrt2)) held on the
  unit cube's boundary, the estimate and the true energy-norm error both fall at rate 2 in [1.7,
  2.3] per halving and their ratio stays in [0.5, 2]. The true error is integrated here with the
  14-point rule from the oracle's point stresses and the exact strain.
* loaded.rs:81-129: on a clamped, end-loaded cantilever the relative error falls at every
  refinement and stays within a factor of two of the true error from the energy identity
  ||u - u_h||^2 = ||u||^2 - ||u_h||^2, ||u||^2 extrapolated (Aitken) from four meshes.
* The solution norm is twice the strain energy: sum f.u over the free unknowns, to 1e-9.
"""
import numpy as np
import pytest

import refcpu
from solidfem_amd import Workload
from workloads import bar_config

NU, KAPPA = 0.3, 1e-3
A = 1.0 / np.sqrt(2.0)
# crates/solidfem-elastic/src/quadrature.rs:60-77: the 14-point rule's weights by orbit
W14 = np.array([0.11268792571801585] * 4 + [0.07349304311636195] * 4 + [0.042546020777081466] * 6)
MIXED_MATERIAL_PATCH = 53


def bending(p):
    """exact.rs:17-25: sigma_xx = E kappa y, every other component zero."""
    x, y, z = p
    return [KAPPA * x * y, -0.5 * KAPPA * (x * x + NU * (y * y - z * z)), -NU * KAPPA * y * z]


def harmonic(p):
    """convergence.rs:27-34"""
    e = np.exp(p[0])
    sy, cy, sz, cz = np.sin(A * p[1]), np.cos(A * p[1]), np.sin(A * p[2]), np.cos(A * p[2])
    return [e * cy * cz, -A * e * sy * cz, -A * e * cy * sz]


def harmonic_strain(p):
    """convergence.rs:37-49: the Hessian of phi, [xx, yy, zz, 2xy, 2yz, 2zx]."""
    e = np.exp(p[:, 0])
    sy, cy, sz, cz = np.sin(A * p[:, 1]), np.cos(A * p[:, 1]), np.sin(A * p[:, 2]), np.cos(A * p[:, 2])
    return np.stack([e * cy * cz, -A * A * e * cy * cz, -A * A * e * cy * cz, 2 * (-A * e * sy * cz),
                     2 * (A * A * e * sy * sz), 2 * (-A * e * cy * sz)], axis=1)


def isotropic_stress(youngs, nu, e):
    lam, mu = youngs * nu / ((1 + nu) * (1 - 2 * nu)), youngs / (2 * (1 + nu))
    trace = e[:, 0] + e[:, 1] + e[:, 2]
    return np.stack([lam * trace + 2 * mu * e[:, 0], lam * trace + 2 * mu * e[:, 1], lam * trace + 2 * mu * e[:, 2],
                     mu * e[:, 3], mu * e[:, 4], mu * e[:, 5]], axis=1)


def energy_density(youngs, nu, s):
    """s : C^-1 : s, from the compliance written as a matrix."""
    c = np.zeros((6, 6))
    c[:3, :3] = -nu / youngs
    np.fill_diagonal(c[:3, :3], 1.0 / youngs)
    c[3:, 3:] = np.eye(3) * 2.0 * (1.0 + nu) / youngs
    return np.einsum("pi,ij,pj->p", s, c, s)


def held(size, cells, youngs, exact, **extra):
    """support.rs:20-61: every boundary node held at `exact`, no load, solved to 1e-14."""
    w = Workload(size=tuple(size), divisions=tuple(cells), clamp_axis=-1, load_axis=-1, youngs=youngs, poisson=NU, **extra)
    components = []
    for n, p in enumerate(refcpu.Problem(w).nodes()):
        if any(p[a] == 0.0 or p[a] == size[a] for a in range(3)):
            components += [(n, a, v) for a, v in enumerate(exact(p))]
    problem = refcpu.Problem(w, prescribed=components)
    u, result = refcpu.Multigrid(problem, refcpu.options(precision=0, coarse_dofs=100)).solve(1e-14)
    assert result.status == 0
    return w, problem, u, result


def strain_energy_twice(problem, u):
    fm = problem.free_map()
    free = fm != 0xFFFFFFFF
    uf = np.zeros(problem.free_dofs)
    uf.reshape(-1, 3)[fm[free]] = u[free]
    return problem.rhs() @ uf


def test_an_exactly_represented_bending_field_reads_as_converged():
    _, problem, u, result = held((3.0, 2.0, 2.0), (3, 2, 2), 2100.0, bending)
    e = problem.error_estimate(u)
    assert e["status"] == 0
    tolerance = 10.0 * (1e2 * result.relative_residual + 1e-13)
    assert e["relative_error"] <= tolerance
    scale = np.sqrt(e["solution_norm_squared"])
    assert len(e["element_errors"]) == 72 and np.all(e["element_errors"] <= tolerance * scale)
    # E kappa^2 int y^2 dV over 3 x 2 x 2 = 2100 * 1e-6 * 3 * (8/3) * 2
    assert e["solution_norm_squared"] == pytest.approx(2100.0 * KAPPA ** 2 * 16.0, rel=1e-12)


def test_a_patch_holding_two_materials_refuses_an_estimate():
    w = bar_config(divisions=(8, 2, 2))
    w.shell_depth = 5.0
    problem = refcpu.Problem(w)
    u = 1e-3 * np.random.default_rng(2).standard_normal((problem.node_count, 3))
    e = problem.error_estimate(u)
    assert e["status"] == MIXED_MATERIAL_PATCH and 0 <= e["a"] < problem.vertex_count


def test_the_estimate_falls_at_the_quadratic_rate_and_tracks_the_true_error():
    rows = []
    for cells in (2, 4, 8):
        w, problem, u, _ = held((1.0, 1.0, 1.0), (cells,) * 3, 1.0, harmonic)
        e = problem.error_estimate(u)
        assert e["status"] == 0
        # the true error: the 14-point rule over each straight element (measure w_q * volume)
        fourteen = refcpu.Problem(Workload(**{**w.__dict__, "quadrature_points": 14}))
        points, _ = fourteen.recover(u)
        x, tets = fourteen.nodes(), fourteen.tets()
        c = x[tets[:, :4]]
        volume = np.abs(np.einsum("ij,ij->i", np.cross(c[:, 1] - c[:, 0], c[:, 2] - c[:, 0]), c[:, 3] - c[:, 0])) / 6.0
        position, element_stress = points[:, :, :3].reshape(-1, 3), points[:, :, 3:9].reshape(-1, 6)
        difference = isotropic_stress(1.0, NU, harmonic_strain(position)) - element_stress
        truth = np.sqrt(np.sum((W14[None, :] * volume[:, None]).ravel() * energy_density(1.0, NU, difference)))
        rows.append((np.sqrt(e["error_norm_squared"]), truth))
    for k in (0, 1):
        rate = np.log2(rows[1][k] / rows[2][k])
        assert 1.7 <= rate <= 2.3, (k, rate)
    for estimated, truth in rows[1:]:
        assert 0.5 <= estimated / truth <= 2.0


def test_a_clamped_cantilever_estimate_falls_and_stays_within_a_factor_of_two():
    rows = []
    for cells in (1, 2, 4, 8):
        w = Workload(size=(4.0, 1.0, 1.0), divisions=(4 * cells, cells, cells), clamp_axis=0, clamp_value=0.0,
                     load_axis=0, load_value=4.0, traction=(0.0, 0.0, -1.0), youngs=1.0, poisson=NU)
        problem = refcpu.Problem(w)
        u, result = refcpu.Multigrid(problem, refcpu.options(precision=0)).solve(1e-12)
        assert result.status == 0
        e = problem.error_estimate(u)
        assert e["status"] == 0
        energy = strain_energy_twice(problem, u)
        assert e["solution_norm_squared"] == pytest.approx(energy, rel=1e-9)
        assert e["error_norm_squared"] == pytest.approx(np.sum(e["element_errors"] ** 2), rel=1e-12)
        assert e["relative_error"] == pytest.approx(
            np.sqrt(e["error_norm_squared"] / (e["solution_norm_squared"] + e["error_norm_squared"])), rel=1e-15)
        rows.append((energy, e))
    e2, e3, e4 = rows[1][0], rows[2][0], rows[3][0]
    ratio = (e3 - e2) / (e4 - e3)
    assert ratio > 1.0
    exact = e4 + (e4 - e3) / (ratio - 1.0)
    previous = np.inf
    for energy, e in rows[:3]:
        assert e["relative_error"] < previous
        previous = e["relative_error"]
        assert 0.5 <= np.sqrt(e["error_norm_squared"]) / np.sqrt(exact - energy) <= 2.0
