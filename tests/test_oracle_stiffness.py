"""The oracle's full stiffness -- assemble_stiffness (crates/solidfem-elastic/src/assembly.rs:17-27),
no node eliminated, three unknowns per node in node order -- pinned on the CPU by:

* the reference's own patch test (crates/solidfem-elastic/tests/element/patch.rs:79-118): a linear
  displacement field is balanced at every interior node to 1e-12 of max|diag| * max|u|, while the
  boundary nodes carry forces (so the check is not empty);
* the six rigid-body null vectors its doc line promises (three translations, three rotations
  omega x p built here in numpy), to the same bound;
* exact symmetry;
* elimination (assembly.rs:29-52): the free stiffness of the same mesh clamped at x = 0 -- the
  matrix the solve fingerprints pin -- is the full stiffness's restriction to the free nodes, bit
  for bit, since each entry sums the same element terms in the same order.
"""
import numpy as np
import scipy.sparse as sp

import refcpu
from solidfem_amd import Workload

GRADIENT = np.array([[0.01, 0.002, -0.003], [0.004, -0.005, 0.006], [0.007, 0.001, 0.008]])
OFFSET = np.array([0.3, -0.2, 0.1])
FIXED = 0xFFFFFFFF


def free_box(**extra):
    """A 3 x 2 x 2 box of 3 x 2 x 2 cells, interior vertices jittered by 10% (seed 7), E 2100, nu 0.3
    (patch.rs:80-83), no clamp and no load."""
    base = dict(size=(3.0, 2.0, 2.0), divisions=(3, 2, 2), jitter_fraction=0.1, jitter_seed=7, clamp_axis=-1,
                load_axis=-1, youngs=2100.0, poisson=0.3)
    return Workload(**{**base, **extra})


def matrix(problem):
    rs, cols, vals = problem.csr()
    n = len(rs) - 1
    return sp.csr_matrix((vals, cols, rs), shape=(n, n)), (rs, cols, vals)


def rigid_modes(x):
    """Translations along x, y, z and rotations about them, omega x p, 3 per node."""
    modes = []
    for k in range(3):
        t = np.zeros_like(x)
        t[:, k] = 1.0
        modes.append(t.ravel())
    for k in range(3):
        omega = np.zeros(3)
        omega[k] = 1.0
        modes.append(np.cross(omega, x).ravel())
    return modes


def test_the_full_stiffness_numbers_every_node_in_node_order():
    p = refcpu.Problem(free_box())
    assert p.free_dofs == 3 * p.node_count
    assert np.array_equal(p.free_map(), np.arange(p.node_count, dtype=np.uint32))


def test_the_full_stiffness_passes_the_reference_patch_test():
    p = refcpu.Problem(free_box())
    k, _ = matrix(p)
    x = p.nodes()
    u = OFFSET + x @ GRADIENT.T
    f = (k @ u.ravel()).reshape(-1, 3)
    scale = np.abs(k.diagonal()).max() * np.abs(u).max()
    interior = np.all((x > 1e-9) & (x < np.array([3.0, 2.0, 2.0]) - 1e-9), axis=1)
    assert interior.sum() > 0
    assert np.abs(f[interior]).max() < 1e-12 * scale
    assert np.abs(f[~interior]).max() > 1e-6 * scale


def test_the_full_stiffness_has_six_rigid_null_vectors_and_is_symmetric():
    p = refcpu.Problem(free_box(quadrature_points=14))
    k, _ = matrix(p)
    scale = np.abs(k.diagonal()).max()
    for r in rigid_modes(p.nodes()):
        assert np.abs(k @ r).max() < 1e-12 * scale * np.abs(r).max()
    assert (k != k.T).nnz == 0


def test_the_free_stiffness_is_the_full_stiffness_restricted_to_free_nodes():
    full, _ = matrix(refcpu.Problem(free_box()))
    clamped = refcpu.Problem(free_box(clamp_axis=0, clamp_value=0.0))
    fm = clamped.free_map()
    free_nodes = np.nonzero(fm != FIXED)[0]
    assert 0 < len(free_nodes) < len(fm)
    dofs = (3 * free_nodes[:, None] + np.arange(3)).ravel()
    restricted = full[dofs][:, dofs].tocsr()
    restricted.sort_indices()
    rs, cols, vals = clamped.csr()
    assert np.array_equal(restricted.indptr, rs) and np.array_equal(restricted.indices, cols)
    assert np.array_equal(restricted.data.view(np.uint64), vals.view(np.uint64))
