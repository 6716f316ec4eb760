"""The oracle pinned by the reference's own known answers (CPU only).

* Bit fingerprints and iteration counts of the Double path, and the pinned (f64, f32)
  iteration pairs: crates/solidfem/tests/validation/multigrid_precision.rs:134-250.
* Closed forms: uniaxial tension (validation/tension.rs:13-57), Timoshenko cantilever
  (validation/cantilever.rs:39-80), the 1-D Laplacian (crates/solidfem-solve/tests/solver/pcg.rs).
* Element invariants: six zero-energy modes, exact symmetry, rule agreement
  (crates/solidfem-elastic/tests/element/rigid_body.rs:90-148).
"""
import numpy as np
import pytest

import refcpu
from workloads import FINGERPRINT_CANTILEVER, FINGERPRINT_PLATE, loaded
from solidfem_amd import Workload


@pytest.mark.parametrize("case", [FINGERPRINT_CANTILEVER, FINGERPRINT_PLATE], ids=["cantilever", "plate"])
def test_double_path_reproduces_the_reference_fingerprint(case):
    size, div, iterations, fingerprint = case
    disp, r = refcpu.solve(loaded(size, div), precision=0, coarse_dofs=100, tolerance=1e-9)
    assert r.status == 0
    assert (r.iterations, r.fingerprint) == (iterations, fingerprint)
    assert refcpu.fingerprint(disp) == fingerprint


@pytest.mark.parametrize("case,precision,expected,floor", [
    (FINGERPRINT_CANTILEVER, 2, (61, 65), False),
    (FINGERPRINT_PLATE, 2, (85, 90), True),
    (FINGERPRINT_CANTILEVER, 1, (61, 61), False),
    (FINGERPRINT_PLATE, 1, (85, 85), True),
], ids=["cantilever-single", "plate-single", "cantilever-storage", "plate-storage"])
def test_pinned_iteration_pairs(case, precision, expected, floor):
    size, div, _, _ = case
    d64, r64 = refcpu.solve(loaded(size, div), precision=0, coarse_dofs=100)
    d32, r32 = refcpu.solve(loaded(size, div), precision=precision, coarse_dofs=100)
    assert (r64.iterations, r32.iterations) == expected
    assert bool(r64.at_rounding_floor) == floor
    largest = np.abs(d64).max()
    assert np.all(np.abs(d64 - d32) <= 2e-3 * largest)
    assert r32.memory_bytes < r64.memory_bytes


def test_a_planted_off_by_one_ulp_changes_the_fingerprint():
    # The fingerprint sees every bit: one ulp in one displacement moves it.
    size, div, _, fingerprint = FINGERPRINT_CANTILEVER
    disp, _ = refcpu.solve(loaded(size, div), precision=0, coarse_dofs=100)
    planted = disp.copy().ravel()
    i = int(np.flatnonzero(planted)[7])
    planted[i] = np.nextafter(planted[i], np.inf)
    assert refcpu.fingerprint(planted) != fingerprint


def test_zero_poisson_tension_is_exact_to_the_solver_tolerance():
    # validation/tension.rs:13-57: u = F L / (A E) at the loaded end, limit 1e-5.
    w = Workload(size=(20.0, 1.0, 1.0), divisions=(20, 2, 2), load_value=20.0, traction=(1000.0, 0.0, 0.0),
                 youngs=2100.0, poisson=0.0)
    disp, r = refcpu.solve(w, precision=0, tolerance=1e-9)
    nodes = refcpu.Problem(w).nodes()
    end = np.abs(nodes[:, 0] - 20.0) < 1e-9
    elongation = disp[end, 0].mean()
    assert abs(elongation - 1000.0 * 20.0 / 2100.0) / (1000.0 * 20.0 / 2100.0) < 1e-5


def test_cantilever_tip_converges_to_timoshenko():
    # validation/cantilever.rs:39-80: E 1000, nu 0, L 10, h 1, unit end load; 1% at 40x4x4.
    e, length, depth = 1000.0, 10.0, 1.0
    inertia, area, g = depth ** 4 / 12.0, depth * depth, e / 2.0
    expected = length ** 3 / (3 * e * inertia) + length / (5.0 / 6.0 * g * area)
    prev_err, prev_tip = np.inf, -np.inf
    for div in [(10, 1, 1), (20, 2, 2), (40, 4, 4)]:
        w = Workload(size=(length, depth, depth), divisions=div, load_value=length, traction=(0.0, 0.0, -1.0),
                     youngs=e, poisson=0.0)
        disp, r = refcpu.solve(w, precision=0, tolerance=1e-9)
        nodes = refcpu.Problem(w).nodes()
        tip = -disp[np.abs(nodes[:, 0] - length) < 1e-9, 2].mean()
        err = abs(tip - expected) / expected
        assert err < prev_err and tip > prev_tip
        prev_err, prev_tip = err, tip
    assert prev_err < 0.01


def laplacian(n):
    rs = np.zeros(n + 1, dtype=np.int64)
    cols, vals = [], []
    for i in range(n):
        for j in (i - 1, i, i + 1):
            if 0 <= j < n:
                cols.append(j)
                vals.append(2.0 if i == j else -1.0)
        rs[i + 1] = len(cols)
    return rs, np.array(cols, dtype=np.uint32), np.array(vals)


def test_laplacian_closed_form_and_slicing():
    # solver/pcg.rs:43-72: x[i] = (i + 1) / (n + 1); sliced runs give identical bits
    n = 50
    rs, cols, vals = laplacian(n)
    b = np.zeros(n)
    b[-1] = 1.0
    x, r = refcpu.pcg_jacobi(rs, cols, vals, b)
    assert r.status == 0
    assert np.all(np.abs(x - (np.arange(n) + 1) / (n + 1)) < 1e-8)
    x3, _ = refcpu.pcg_jacobi(rs, cols, vals, b, slice_iterations=3)
    assert np.array_equal(x.view(np.uint64), x3.view(np.uint64))


def test_a_tolerance_below_rounding_stops_at_the_floor():
    # solver/pcg.rs:103-126
    n = 50
    rs, cols, vals = laplacian(n)
    b = np.zeros(n)
    b[-1] = 1.0
    x, r = refcpu.pcg_jacobi(rs, cols, vals, b, tolerance=1e-18)
    assert r.status == 0 and r.at_rounding_floor and r.relative_residual > 1e-18
    assert np.all(np.abs(x - (np.arange(n) + 1) / (n + 1)) < 1e-10)


def test_an_indefinite_matrix_is_refused():
    # solver/pcg.rs:89-101
    rs = np.array([0, 2, 4], dtype=np.int64)
    cols = np.array([0, 1, 0, 1], dtype=np.uint32)
    vals = np.array([1.0, 3.0, 3.0, 1.0])
    _, r = refcpu.pcg_jacobi(rs, cols, vals, np.array([1.0, -1.0]))
    assert r.status == 2


def test_element_has_six_zero_modes_is_symmetric_and_rules_agree():
    # element/rigid_body.rs:90-148 on a jittered mesh's elements
    w = Workload(size=(2.0, 1.0, 1.0), divisions=(2, 2, 2), jitter_fraction=0.1, jitter_seed=3, load_value=2.0,
                 traction=(1.0, 0.0, 0.0), youngs=1000.0, poisson=0.3)
    p4 = refcpu.Problem(w)
    w14 = Workload(**{**w.__dict__, "quadrature_points": 14})
    p14 = refcpu.Problem(w14)
    for tet in (0, 5, 17, 40):
        ke = p4.element_stiffness(tet)
        assert np.array_equal(ke, ke.T)
        ev = np.sort(np.abs(np.linalg.eigvalsh(ke)))
        assert np.all(ev[:6] < 1e-10 * ev[-1]) and ev[6] > 1e-6 * ev[-1]
        ke14 = p14.element_stiffness(tet)
        assert np.abs(ke - ke14).max() <= 1e-12 * np.abs(ke).max()
