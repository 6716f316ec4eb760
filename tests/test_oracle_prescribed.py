"""The oracle's prescribed-displacement elimination (StudyBuilder::apply_prescribed,
crates/solidfem-elastic/src/study.rs:567-593) pinned on the CPU.

* Structure, against an independent numpy statement of symmetric elimination: the free block of
  the stiffness keeps its bits, the couplings between free and prescribed unknowns vanish, a
  prescribed row keeps only its diagonal, the free right-hand side becomes f_F - K_FP g and a
  prescribed one K_pp g.
* Closed forms of crates/solidfem/tests/validation/prescribed.rs:79-114 on the stretched unit cube:
  exact prescribed values, the uniform stress E * stretch along x and nothing else, and the
  transverse contraction -nu * stretch at the free far corner.
"""
import numpy as np
import pytest

import refcpu
from workloads import STRETCH, Workload, stretch_components, stretched_cube

E, NU = 200000.0, 0.3


def dense(rs, cols, vals):
    n = len(rs) - 1
    k = np.zeros((n, n))
    for i in range(n):
        k[i, cols[rs[i]:rs[i + 1]]] = vals[rs[i]:rs[i + 1]]
    return k


def free_dofs_of(problem, components):
    fm = problem.free_map()
    return sorted({3 * int(fm[n]) + a for n, a, _ in components})


def test_elimination_matches_symmetric_elimination_of_the_assembled_system():
    # a loaded cube (traction on x = 1) with prescribed components on x = 0 and a few inside
    w = Workload(size=(1.0, 1.0, 1.0), divisions=(2, 2, 2), clamp_axis=-1, load_axis=0, load_value=1.0,
                 traction=(0.3, -0.2, 0.5), youngs=E, poisson=NU)
    plain = refcpu.Problem(w)
    nodes = plain.nodes()
    comps = [(n, a, 0.001 * (n % 5) - 0.002 * a) for n, p in enumerate(nodes) if p[0] == 0.0 for a in range(3)]
    comps += [(n, 1, 0.0005 * n) for n, p in enumerate(nodes) if p[0] == 0.5 and p[2] == 0.5]
    elim = refcpu.Problem(w, prescribed=comps)
    k, f = dense(*plain.csr()), plain.rhs()
    k2, f2 = dense(*elim.csr()), elim.rhs()
    pres = free_dofs_of(plain, comps)
    g = np.zeros(len(f))
    for n, a, v in comps:
        g[3 * int(plain.free_map()[n]) + a] = v
    mask = np.zeros(len(f), dtype=bool)
    mask[pres] = True
    free = ~mask
    # the free block is untouched, bit for bit
    assert np.array_equal(k2[np.ix_(free, free)].view(np.uint64), k[np.ix_(free, free)].view(np.uint64))
    # couplings vanish and prescribed rows keep only their diagonal
    assert not np.any(k2[np.ix_(free, mask)]) and not np.any(k2[np.ix_(mask, free)])
    kpp = k2[np.ix_(mask, mask)]
    assert np.array_equal(np.diag(kpp), np.diag(k)[mask]) and not np.any(kpp - np.diag(np.diag(kpp)))
    # right-hand sides: f_F - K_FP g (the oracle subtracts term by term in row order) and K_pp g
    expected = f[free] - k[np.ix_(free, mask)] @ g[mask]
    assert np.allclose(f2[free], expected, rtol=1e-12, atol=1e-15 * np.abs(k).max())
    assert np.array_equal(f2[mask], np.diag(k)[mask] * g[mask])
    # and the eliminated system's exact solution is the constrained minimiser
    u = np.linalg.solve(k2, f2)
    assert np.allclose(u[mask], g[mask], rtol=1e-9, atol=1e-15)
    uf = np.linalg.solve(k[np.ix_(free, free)], f[free] - k[np.ix_(free, mask)] @ g[mask])
    assert np.allclose(u[free], uf, rtol=1e-9, atol=1e-12 * np.abs(uf).max())


@pytest.mark.parametrize("divisions", [(1, 1, 1), (3, 3, 3)])
def test_a_stretched_cube_has_the_exact_uniform_stress(divisions):
    w = stretched_cube(divisions)
    probe = refcpu.Problem(w)
    nodes = probe.nodes()
    comps = stretch_components(nodes)
    problem = refcpu.Problem(w, prescribed=comps)
    mg = refcpu.Multigrid(problem, refcpu.options(precision=0, coarse_dofs=100))
    disp, result = mg.solve(1e-12)
    assert result.status == 0
    for n, p in enumerate(nodes):
        if p[0] == 1.0:
            assert disp[n, 0] == STRETCH
        if p[0] == 0.0:
            assert disp[n, 0] == 0.0
    corner = next(n for n, p in enumerate(nodes) if tuple(p) == (1.0, 1.0, 1.0))
    assert abs(disp[corner, 1] - (-NU * STRETCH)) < 1e-6 * STRETCH
    assert abs(disp[corner, 2] - (-NU * STRETCH)) < 1e-6 * STRETCH
    points, _ = problem.recover(disp)
    expected = E * STRETCH
    stress = points[:, :, 3:9].reshape(-1, 6)
    assert np.all(np.abs(stress[:, 0] - expected) < 1e-5 * expected)
    assert np.all(np.abs(stress[:, 1:]) < 1e-5 * expected)


def test_the_last_prescription_of_a_component_wins_and_fixed_nodes_are_refused():
    w = stretched_cube()
    nodes = refcpu.Problem(w).nodes()
    comps = stretch_components(nodes)
    again = comps + [(comps[0][0], comps[0][1], 0.25)]
    a, b = refcpu.Problem(w, prescribed=again), refcpu.Problem(w, prescribed=comps[1:] + [(comps[0][0], comps[0][1], 0.25)])
    assert np.array_equal(a.rhs(), b.rhs())
    clamped = Workload(size=(1.0, 1.0, 1.0), divisions=(1, 1, 1), clamp_axis=0, clamp_value=0.0, load_axis=-1)
    with pytest.raises(RuntimeError):
        refcpu.Problem(clamped, prescribed=[(0, 0, 0.0)])
