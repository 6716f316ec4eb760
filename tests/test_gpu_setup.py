"""The multigrid setup's numeric phases on the device against the host phases, bit for bit.

The power iteration (crates/solidfem-solve/src/multigrid/setup.rs:390-449), the smoothed
prolongator (setup.rs:527-573, transfer.rs:108-152), A P (transfer.rs:155-169) and the Galerkin
product (transfer.rs:204-238) run on the device when the setup has a context. Every level's
matrix, block inverse, spectrum, prolongator and aggregates, and the coarse factor, must be the
host setup's bits, which the CPU suite pins to the oracle; here the oracle checks the device
hierarchy directly too. With every level on the device (SetupDevice.DEVICE) the seed pass of the
aggregation (aggregate.rs:119-136) runs there as well, in parallel rounds that build the serial
pass's greedy set, so the aggregates, and with them every level, keep their bits.
"""
import numpy as np
import pytest

import refcpu
import solidfem_amd as sf
from workloads import FINGERPRINT_CANTILEVER, FINGERPRINT_PLATE, bar_config, block_config, loaded, plate_config

pytestmark = pytest.mark.gpu

WORKLOADS = {
    "cantilever": (loaded(*FINGERPRINT_CANTILEVER[:2]), 100),
    "plate": (loaded(*FINGERPRINT_PLATE[:2]), 100),
    "jittered-plate": (plate_config(divisions=(12, 12, 2)), 150),
    "random-block": (block_config(3, divisions=(6, 4, 4)), 60),
    "shell-core-torsion": (bar_config(divisions=(16, 4, 4)), 90),
}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def run_setup(ctx, ref, options, budget=2 ** 62):
    rs, cols, vals = ref.csr()
    csr = sf.CsrMatrix.new(None, rs, cols, vals)
    setup = sf.MultigridSetup.new(ctx, csr, ref.rigid_modes(), ref.free_positions(), options)
    while not setup.step(budget):
        pass
    return setup


def assert_same_setup(a, b):
    assert a.level_count() == b.level_count()
    levels = a.level_count()
    for level in range(levels):
        ma, mb = a.level_matrix(level), b.level_matrix(level)
        assert np.array_equal(ma["starts"], mb["starts"]) and np.array_equal(ma["columns"], mb["columns"]), level
        assert np.array_equal(bits(ma["values"]), bits(mb["values"])), f"level {level} matrix"
        assert np.array_equal(bits(a.level_inverse_diagonal(level)), bits(b.level_inverse_diagonal(level)))
        assert a.smoother_spectrum(level) == b.smoother_spectrum(level), f"level {level} spectrum"
        if level + 1 < levels:
            pa, pb = a.level_matrix(level, prolongator=True), b.level_matrix(level, prolongator=True)
            assert np.array_equal(pa["starts"], pb["starts"]) and np.array_equal(pa["columns"], pb["columns"])
            assert np.array_equal(bits(pa["values"]), bits(pb["values"])), f"level {level} prolongator"
            assert np.array_equal(a.aggregates(level), b.aggregates(level))
    fa, fb = a.coarse_factor(), b.coarse_factor()
    assert (fa is None) == (fb is None)
    if fa is not None:
        assert np.array_equal(bits(np.tril(fa)), bits(np.tril(fb)))


def assert_matches_oracle(setup, mg):
    assert setup.level_count() == mg.levels
    for level in range(mg.levels):
        starts, cols, vals = mg.matrix(level)
        m = setup.level_matrix(level)
        assert np.array_equal(m["columns"], cols) and np.array_equal(bits(m["values"]), bits(vals)), f"level {level}"
        assert setup.smoother_spectrum(level) == mg.spectrum(level)
        if level + 1 < mg.levels:
            ps, pc, pv = mg.prolongator(level)
            p = setup.level_matrix(level, prolongator=True)
            assert np.array_equal(p["columns"], pc) and np.array_equal(bits(p["values"]), bits(pv)), f"level {level} P"


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_device_setup_is_the_host_setup_and_the_oracle(ctx, name):
    w, coarse = WORKLOADS[name]
    ref = refcpu.Problem(w)
    options = sf.MultigridOptions(coarse_dofs=coarse, precision=sf.Precision.DOUBLE, setup_device=sf.SetupDevice.DEVICE)
    device = run_setup(ctx, ref, options)
    host = run_setup(None, ref, options)
    assert_same_setup(device, host)
    assert_matches_oracle(device, refcpu.Multigrid(ref, refcpu.options(precision=0, coarse_dofs=coarse)))


@pytest.mark.parametrize("budget", [1, 977])
def test_a_sliced_device_setup_gives_the_same_bits(ctx, budget):
    w, coarse = WORKLOADS["jittered-plate"]
    ref = refcpu.Problem(w)
    options = sf.MultigridOptions(coarse_dofs=coarse, precision=sf.Precision.DOUBLE, setup_device=sf.SetupDevice.DEVICE)
    assert_same_setup(run_setup(ctx, ref, options, budget=budget), run_setup(None, ref, options))


def test_the_device_seed_pass_on_a_larger_plate_is_the_serial_pass(ctx):
    # about 36,000 block rows on level 0: many rounds of the parallel seed pass
    ref = refcpu.Problem(plate_config(divisions=(40, 40, 2)))
    device = run_setup(ctx, ref, sf.MultigridOptions(precision=sf.Precision.SINGLE, setup_device=sf.SetupDevice.DEVICE))
    host = run_setup(ctx, ref, sf.MultigridOptions(precision=sf.Precision.SINGLE, setup_device=sf.SetupDevice.HOST))
    assert_same_setup(device, host)


def test_a_coarse_factor_of_hundreds_of_unknowns_is_factored_on_the_device_to_the_hosts_bits(ctx):
    # a 10:1:1 bar of 36,300 free DOFs coarsens to a dense level of about 950 unknowns, whose
    # Cholesky factor runs on the device column by column (the host's order): the factor, the
    # whole hierarchy and the oracle's factor agree to the bit, also with the setup sliced
    from solidfem_amd import Workload
    w = Workload(size=(50.0, 5.0, 5.0), divisions=(50, 5, 5), clamp_axis=0, clamp_value=0.0, load_axis=0, load_value=50.0,
                 traction_kind=1, traction_scale=1.0, traction_seed=1, youngs=70000.0, poisson=0.3)
    ref = refcpu.Problem(w)
    options = sf.MultigridOptions(precision=sf.Precision.SINGLE, coarse_dofs=1500)
    auto = run_setup(ctx, ref, options)
    n = auto.coarse_factor().shape[0]
    assert 128 <= n <= 1500, n
    host = run_setup(None, ref, options)
    assert_same_setup(auto, host)
    assert_same_setup(run_setup(ctx, ref, options, budget=977), host)
    factor = refcpu.Multigrid(ref, refcpu.options(precision=2, coarse_dofs=1500)).coarse_factor()
    assert np.array_equal(bits(np.tril(auto.coarse_factor())), bits(np.tril(factor)))


def test_automatic_placement_on_a_larger_plate(ctx):
    # about 36,000 block rows: level 0 runs on the device, the coarse levels on the host
    ref = refcpu.Problem(plate_config(divisions=(40, 40, 2)))
    auto = run_setup(ctx, ref, sf.MultigridOptions(precision=sf.Precision.SINGLE))
    host = run_setup(ctx, ref, sf.MultigridOptions(precision=sf.Precision.SINGLE, setup_device=sf.SetupDevice.HOST))
    assert auto.level_matrix(0)["starts"].size - 1 >= 4096
    assert_same_setup(auto, host)


@pytest.mark.parametrize("case", [FINGERPRINT_CANTILEVER, FINGERPRINT_PLATE], ids=["cantilever", "plate"])
def test_a_device_built_study_solves_to_the_fingerprint(ctx, case):
    # the study builder hands its assembled values on the device to level 0's setup
    size, div, iterations, fingerprint = case
    b = sf.StudyBuilder.new(ctx, sf.Model.from_workload(loaded(size, div)),
                            sf.PcgOptions(tolerance=1e-9, dot_order=sf.DotOrder.REFERENCE),
                            sf.MultigridOptions(precision=sf.Precision.DOUBLE, coarse_dofs=100, setup_device=sf.SetupDevice.DEVICE))
    while True:
        p = b.step(997)
        if p.elements_done == p.elements_total and p.setup_finished:
            break
    s = b.finish()
    p = s.solve()
    assert p.outcome == sf.StepOutcome.CONVERGED
    assert (p.iteration, refcpu.fingerprint(s.displacements())) == (iterations, fingerprint)


@pytest.mark.parametrize("limit", [4, 8, 16])
def test_device_phases_that_overflow_take_the_host_path_with_the_same_bits(ctx, slot_limit, limit):
    # rows with more output blocks than the limit send their phase to the host. On this
    # cantilever's finest level the prolongator rows hold at most 7 blocks, A P rows 12 and
    # Galerkin rows 19: 4 sends the smoothing (and so A P and Galerkin) to the host, 8 sends A P
    # and Galerkin, 16 the Galerkin product alone. The host then fetches what the device kept (the
    # study's matrix, the prolongator, A P), and the study must still solve to the reference's
    # fingerprint with the host setup's hierarchy
    slot_limit(limit)
    size, div, iterations, fingerprint = FINGERPRINT_CANTILEVER
    w = loaded(size, div)
    options = sf.MultigridOptions(precision=sf.Precision.DOUBLE, coarse_dofs=100, setup_device=sf.SetupDevice.DEVICE)
    b = sf.StudyBuilder.new(ctx, sf.Model.from_workload(w), sf.PcgOptions(tolerance=1e-9, dot_order=sf.DotOrder.REFERENCE), options)
    b.step(2 ** 62)
    s = b.finish()
    p = s.solve()
    assert p.outcome == sf.StepOutcome.CONVERGED
    assert (p.iteration, refcpu.fingerprint(s.displacements())) == (iterations, fingerprint)
    mg, host = s.multigrid(), run_setup(None, refcpu.Problem(w), options)
    assert len(mg.level_sizes()) == host.level_count()
    for level in range(host.level_count()):
        for prolongator in ([False, True] if level + 1 < host.level_count() else [False]):
            a, b = mg.level_matrix(level, prolongator), host.level_matrix(level, prolongator)
            assert np.array_equal(a["columns"], b["columns"]) and np.array_equal(bits(a["values"]), bits(b["values"]))
        assert np.array_equal(bits(mg.level_inverse_diagonal(level)), bits(host.level_inverse_diagonal(level)))


@pytest.mark.parametrize("polls", [1, 3, 40])
def test_a_device_setup_cancelled_midway_refuses_further_steps_and_a_new_one_matches(ctx, polls):
    # setup.rs:278-284: the cancel closure is polled before every setup chunk. Cancelled after a
    # few polls, while the level's power iteration may still be running beside the aggregation,
    # the setup refuses further steps by name and can be dropped; a new setup gives the bits.
    w, coarse = WORKLOADS["jittered-plate"]
    ref = refcpu.Problem(w)
    rs, cols, vals = ref.csr()
    csr = sf.CsrMatrix.new(None, rs, cols, vals)
    options = sf.MultigridOptions(coarse_dofs=coarse, precision=sf.Precision.DOUBLE, setup_device=sf.SetupDevice.DEVICE)
    setup = sf.MultigridSetup.new(ctx, csr, ref.rigid_modes(), ref.free_positions(), options)
    count = [0]

    def cancel():
        count[0] += 1
        return count[0] > polls

    with pytest.raises(sf.MultigridError) as e:
        while not setup.step(97, cancel=cancel):
            pass
    assert e.value.status == 23  # SF_ERR_MG_CANCELLED
    with pytest.raises(sf.MultigridError):
        setup.step(2 ** 62)
    setup.close()
    assert_same_setup(run_setup(ctx, ref, options), run_setup(None, ref, options))
