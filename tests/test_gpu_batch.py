"""Several load cases over one matrix and one hierarchy, on the GPU, through the C ABI.

The reference solves one right-hand side per Pcg (crates/solidfem-solve/src/pcg.rs:113-283); the
batch runs that recurrence per column in lockstep. Every column must be bit for bit the
single-column solve of that column -- and so, with the dot products in the reference's order,
the oracle's solve of that load case -- whatever the other columns hold.
"""
import numpy as np
import pytest

import refcpu
import solidfem_amd as sf
from workloads import FINGERPRINT_PLATE, block_config, loaded, plate_config

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def device_multigrid(ctx, ref, precision, coarse, coarse_solve=0):
    rs, cols, vals = ref.csr()
    csr = sf.CsrMatrix.new(ctx, rs, cols, vals)
    setup = sf.MultigridSetup.new(ctx, csr, ref.rigid_modes(), ref.free_positions(),
                                  sf.MultigridOptions(coarse_dofs=coarse, precision=precision, coarse_solve=coarse_solve))
    while not setup.step(2 ** 62):
        pass
    return csr, setup.finish()


def free_to_nodal(ref, x):
    """The full nodal displacement (fixed components zero) from free unknowns."""
    fm = ref.free_map()
    out = np.zeros((len(fm), 3))
    free = fm != 0xFFFFFFFF
    out[free] = x.reshape(-1, 3)[fm[free]]
    return out


@pytest.mark.parametrize("precision", [0, 1, 2], ids=["double", "single-storage", "single"])
@pytest.mark.parametrize("coarse_solve", [0, 1], ids=["reference-coarse", "inverse-coarse"])
@pytest.mark.parametrize("columns", [1, 5, 33])
def test_batched_v_cycle_is_bit_identical_per_column(ctx, precision, coarse_solve, columns):
    ref = refcpu.Problem(loaded(*FINGERPRINT_PLATE[:2]))
    _, mg = device_multigrid(ctx, ref, precision, 100, coarse_solve)
    assert mg.stats().levels >= 3 and mg.threads() == 1
    r = np.random.default_rng(columns).standard_normal((columns, ref.free_dofs))
    z = mg.apply_batch(r)
    for m in range(columns):
        assert np.array_equal(bits(z[m]), bits(mg.apply(r[m]))), f"column {m}"
    assert np.array_equal(bits(mg.apply_batch(r)), bits(z)), "a second application differs"


def test_batched_solve_reproduces_the_oracle_per_load_case(ctx):
    # five random-traction load cases of the clamped block, one hierarchy, the reference's order
    cases = [block_config(c, divisions=(6, 4, 4)) for c in range(5)]
    refs = [refcpu.Problem(w) for w in cases]
    csr, mg = device_multigrid(ctx, refs[0], 0, 100)
    rhs = np.stack([r.rhs() for r in refs])
    batch = sf.PcgBatch.new(ctx, csr, mg, rhs, sf.PcgOptions(dot_order=sf.DotOrder.REFERENCE))
    running, progress = batch.step(10000)
    assert running == 0
    x = batch.estimate()
    for c, ref in enumerate(refs):
        mg_ref = refcpu.Multigrid(ref, refcpu.options(precision=0, coarse_dofs=100))
        disp, r = mg_ref.solve(1e-9)
        assert progress[c].outcome == sf.StepOutcome.CONVERGED
        assert progress[c].iteration == r.iterations, c
        assert np.array_equal(bits(free_to_nodal(ref, x[c])), bits(disp)), f"load case {c}"


@pytest.mark.parametrize("precision", [sf.Precision.DOUBLE, sf.Precision.SINGLE], ids=["double", "single"])
def test_batched_tree_solve_equals_each_single_solve(ctx, precision):
    ref = refcpu.Problem(plate_config(divisions=(12, 12, 2)))
    csr, mg0 = device_multigrid(ctx, ref, 0, 150)
    mg = mg0.with_precision(precision, coarse_solve=1)
    rng = np.random.default_rng(21)
    rhs = np.stack([ref.rhs(), rng.standard_normal(ref.free_dofs), 1e-3 * rng.standard_normal(ref.free_dofs)])
    opts = sf.PcgOptions(tolerance=1e-8, dot_order=sf.DotOrder.TREE)
    batch = sf.PcgBatch.new(ctx, csr, mg, rhs, opts)
    running, progress = batch.step(100000)
    assert running == 0
    x = batch.estimate()
    for m in range(len(rhs)):
        single = sf.Pcg.new(ctx, csr, mg, rhs[m], opts)
        p = single.step(100000)
        assert p.outcome == progress[m].outcome == sf.StepOutcome.CONVERGED
        assert (p.iteration, p.at_rounding_floor) == (progress[m].iteration, progress[m].at_rounding_floor), m
        assert np.array_equal(bits(single.estimate()), bits(x[m])), f"column {m}"
    # the same columns in another order and another company
    other = sf.PcgBatch.new(ctx, csr, mg, np.stack([rhs[2], rhs[0]]), opts)
    other.step(100000)
    y = other.estimate()
    assert np.array_equal(bits(y[0]), bits(x[2])) and np.array_equal(bits(y[1]), bits(x[0]))
    # slicing the lockstep iterations changes nothing
    sliced = sf.PcgBatch.new(ctx, csr, mg, rhs, opts)
    while sliced.step(7)[0]:
        pass
    assert np.array_equal(bits(sliced.estimate()), bits(x))
    # a restart with the right-hand sides swapped
    sliced.restart(rhs[::-1].copy())
    sliced.step(100000)
    assert np.array_equal(bits(sliced.estimate()), bits(x[::-1]))


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


@pytest.mark.parametrize("dot", [sf.DotOrder.REFERENCE, sf.DotOrder.TREE])
def test_batched_jacobi_columns_keep_their_own_outcomes(ctx, dot):
    n = 50
    rs, cols, vals = laplacian(n)
    csr = sf.CsrMatrix.new(ctx, rs, cols, vals)
    e = np.zeros(n)
    e[-1] = 1.0
    ramp = np.linspace(-1.0, 2.0, n)
    rhs = np.stack([e, np.zeros(n), ramp, 3.0 * e])
    opts = sf.PcgOptions(dot_order=dot)
    batch = sf.PcgBatch.new(ctx, csr, None, rhs, opts)
    running, progress = batch.step(10000)
    assert running == 0
    x = batch.estimate()
    assert progress[1].outcome == sf.StepOutcome.CONVERGED and progress[1].iteration == 0
    assert np.all(x[1] == 0.0)
    assert np.all(np.abs(x[0] - (np.arange(n) + 1) / (n + 1)) < 1e-8)
    for m in (0, 2, 3):
        single = sf.Pcg.new(ctx, csr, None, rhs[m], opts)
        p = single.step(10000)
        assert p.iteration == progress[m].iteration
        assert np.array_equal(bits(single.estimate()), bits(x[m])), m
    # an iteration limit fails each column by name at the single solve's iteration
    limited = sf.PcgBatch.new(ctx, csr, None, rhs, sf.PcgOptions(dot_order=dot, max_iterations=5))
    limited.step(10000)
    got = limited.progress()
    assert got[1].outcome == sf.StepOutcome.CONVERGED
    for m in (0, 2, 3):
        assert got[m].outcome == sf.PcgBatch.FAILED and got[m].error == 11 and got[m].error_iteration == 5
        with pytest.raises(sf.PcgError) as err:
            sf.Pcg.new(ctx, csr, None, rhs[m], sf.PcgOptions(dot_order=dot, max_iterations=5)).step(10000)
        assert err.value.value == got[m].error_value
    # cancel stops every running column; a later step keeps the outcome
    c = sf.PcgBatch.new(ctx, csr, None, rhs, opts)
    c.step(3)
    running, got = c.step(1000, cancel=lambda: True)
    assert running == 0
    assert [g.outcome for g in got] == [sf.StepOutcome.CANCELLED, sf.StepOutcome.CONVERGED] + [sf.StepOutcome.CANCELLED] * 2
    assert all(g.iteration == 3 for g in (got[0], got[2], got[3]))


def test_an_indefinite_column_fails_alone(ctx):
    rs = np.array([0, 2, 4], dtype=np.int64)
    cols = np.array([0, 1, 0, 1], dtype=np.uint32)
    csr = sf.CsrMatrix.new(ctx, rs, cols, np.array([1.0, 3.0, 3.0, 1.0]))
    batch = sf.PcgBatch.new(ctx, csr, None, np.array([[1.0, -1.0], [1.0, 1.0]]))
    running, got = batch.step(100)
    assert running == 0
    assert got[0].outcome == sf.PcgBatch.FAILED and got[0].error == 10 and got[0].error_iteration == 0
    assert got[0].error_value < 0
    assert got[1].outcome == sf.StepOutcome.CONVERGED
    assert np.allclose(batch.estimate()[1], [0.25, 0.25])


def test_batch_limits_are_refused(ctx):
    n = 8
    rs, cols, vals = laplacian(n)
    csr = sf.CsrMatrix.new(ctx, rs, cols, vals)
    with pytest.raises(sf.SolidfemError):
        sf.PcgBatch.new(ctx, csr, None, np.ones((65, n)))


def test_solves_outlive_their_context_handle():
    # Handles caught in a reference cycle are finalized together in any order; a solve keeps the
    # stream it launches on alive after its context handle is gone.
    own = sf.Context(0)
    n = 8
    rs, cols, vals = laplacian(n)
    csr = sf.CsrMatrix.new(own, rs, cols, vals)
    pcg = sf.Pcg.new(own, csr, None, np.ones(n))
    batch = sf.PcgBatch.new(own, csr, None, np.ones((2, n)))
    own.close()
    assert pcg.step(100).outcome == sf.StepOutcome.CONVERGED
    running, progress = batch.step(100)
    assert running == 0 and all(p.outcome == sf.StepOutcome.CONVERGED for p in progress)
    assert np.array_equal(bits(batch.estimate()[0]), bits(pcg.estimate()))
    csr.close()
    pcg.close()
    batch.close()


# the batched product forms: lane groups with 1, 2 or 4 blocks in flight, one wavefront per block
# row, and lanes that take two or four consecutive columns each
FORMS = {"group-1": 1, "group-4": 4, "node-2": 16 + 2, "pair-1": 32 + 1, "pair-2": 32 + 2, "quad-1": 48 + 1, "quad-2": 48 + 2,
         "oct-1": 64 + 1, "oct-2": 64 + 2, "group-8": 8, "staged": 80}


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("precision", [0, 2], ids=["double", "single"])
def test_each_batched_form_gives_each_column_its_single_bits(ctx, forced_variant, form, precision):
    # every product and Chebyshev step of the batched V-cycle forced to one form (where it is a
    # candidate for the level's block shape): 16 and 8 columns, multiples of two, four and (16)
    # eight, then 12 and 6, which reuse the choices made for 16 and 8 (the same group sizes), so
    # the eight- and four-column forms fall back to one column per lane
    forced_variant(FORMS[form])
    ref = refcpu.Problem(loaded(*FINGERPRINT_PLATE[:2]))
    _, mg = device_multigrid(ctx, ref, precision, 100)
    for columns in (16, 8, 12, 6):
        r = np.random.default_rng(columns).standard_normal((columns, ref.free_dofs))
        z = mg.apply_batch(r)
        for m in range(columns):
            assert np.array_equal(bits(z[m]), bits(mg.apply(r[m]))), f"{columns} columns: column {m}"
