"""The product's host side against the oracle, bit for bit (CPU only, no device call).

Mesh generation (crates/solidfem-mesh), numbering, load and rigid-body modes
(crates/solidfem-elastic/src/{study,loads,rigid_modes}.rs), the free block pattern
(assembly.rs:55-81) and the whole multigrid setup (multigrid/setup.rs), including slicing,
cancellation and the named refusals (crates/solidfem-solve/tests/solver/multigrid.rs).
"""
import numpy as np
import pytest

import refcpu
import solidfem_amd as sf
from workloads import FINGERPRINT_CANTILEVER, FINGERPRINT_PLATE, block_config, bar_config, loaded, plate_config, tube_config

WORKLOADS = {
    "cantilever": loaded(*FINGERPRINT_CANTILEVER[:2]),
    "plate": loaded(*FINGERPRINT_PLATE[:2]),
    "jittered-plate": plate_config(divisions=(12, 12, 2)),
    "random-block": block_config(3, divisions=(6, 4, 4)),
    "shell-core-torsion": bar_config(divisions=(16, 4, 4)),
    "tube": tube_config(divisions=(20, 7, 8), wall_cells=2, poisson=0.3),
}


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("name", list(WORKLOADS))
def test_mesh_load_and_pattern_match_the_oracle(name):
    w = WORKLOADS[name]
    ref = refcpu.Problem(w)
    model = sf.Model.from_workload(w)
    mesh = model.mesh()
    assert np.array_equal(bits(mesh.nodes()), bits(ref.nodes()))
    assert np.array_equal(mesh.tets(), ref.tets())
    assert np.array_equal(mesh.boundary_faces(), ref.boundary_faces())
    free_map, rhs, starts, cols, modes = model.free_problem()
    assert np.array_equal(free_map, ref.free_map())
    assert np.array_equal(bits(rhs), bits(ref.rhs()))
    assert np.array_equal(bits(modes), bits(ref.rigid_modes()))
    # block pattern against the oracle's scalar CSR: row 3n lists 3 columns per block
    rs, rcols, _ = ref.csr()
    nodes = len(starts) - 1
    for node in range(0, nodes, max(1, nodes // 97)):
        row = rcols[rs[3 * node]:rs[3 * node + 1]]
        assert np.array_equal(row[::3] // 3, cols[starts[node]:starts[node + 1]])
    assert 9 * len(cols) == len(rcols)


def host_setup(ref, options, budget=2 ** 62):
    rs, cols, vals = ref.csr()
    csr = sf.CsrMatrix.new(None, rs, cols, vals)
    setup = sf.MultigridSetup.new(None, csr, ref.rigid_modes(), ref.free_positions(), options)
    while not setup.step(budget):
        pass
    return setup


def assert_same_hierarchy(setup, mg):
    assert setup.level_count() == mg.levels
    for level in range(mg.levels):
        m = setup.level_matrix(level)
        starts, cols, vals = mg.matrix(level)
        assert np.array_equal(m["starts"], starts) and np.array_equal(m["columns"], cols)
        assert np.array_equal(bits(m["values"]), bits(vals)), f"level {level} matrix"
        assert np.array_equal(bits(setup.level_inverse_diagonal(level)), bits(mg.inverse_diagonal(level)))
        lo, hi = setup.smoother_spectrum(level)
        rlo, rhi = mg.spectrum(level)
        assert (lo, hi) == (rlo, rhi), f"level {level} spectrum"
        if level + 1 < mg.levels:
            p = setup.level_matrix(level, prolongator=True)
            ps, pc, pv = mg.prolongator(level)
            assert np.array_equal(p["starts"], ps) and np.array_equal(p["columns"], pc)
            assert np.array_equal(bits(p["values"]), bits(pv)), f"level {level} prolongator"
            assert np.array_equal(setup.aggregates(level), mg.aggregates(level))
    f = setup.coarse_factor()
    rf = mg.coarse_factor()
    assert (f is None) == (rf is None)
    if f is not None:
        assert np.array_equal(bits(np.tril(f)), bits(np.tril(rf)))


@pytest.mark.parametrize("name,coarse", [("cantilever", 100), ("plate", 100), ("jittered-plate", 150),
                                         ("random-block", 60), ("shell-core-torsion", 90), ("tube", 120)])
def test_host_multigrid_setup_is_bit_identical_to_the_oracle(name, coarse):
    w = WORKLOADS[name]
    ref = refcpu.Problem(w)
    mg = refcpu.Multigrid(ref, refcpu.options(precision=0, coarse_dofs=coarse))
    setup = host_setup(ref, sf.MultigridOptions(coarse_dofs=coarse, precision=sf.Precision.DOUBLE))
    assert_same_hierarchy(setup, mg)


@pytest.mark.parametrize("budget", [1, 7, 53, 977])
def test_a_sliced_setup_gives_the_same_bits(budget):
    ref = refcpu.Problem(WORKLOADS["plate"])
    mg = refcpu.Multigrid(ref, refcpu.options(precision=0, coarse_dofs=100))
    setup = host_setup(ref, sf.MultigridOptions(coarse_dofs=100, precision=sf.Precision.DOUBLE), budget=budget)
    assert_same_hierarchy(setup, mg)


def block_chain(nodes):
    # solver/multigrid.rs:11-28: SPD 3x3 couplings along a chain
    rs = [0]
    cols, vals = [], []
    for node in range(nodes):
        for axis in range(3):
            row = 3 * node + axis
            for nb in (node - 1, node, node + 1):
                if 0 <= nb < nodes:
                    for q in range(3):
                        cols.append(3 * nb + q)
                        vals.append((4.0 if nb == node else -1.0) if q == axis else 0.0)
            rs.append(len(cols))
    return np.array(rs, dtype=np.int64), np.array(cols, dtype=np.uint32), np.array(vals)


def test_refusals_are_named():
    rs, cols, vals = block_chain(4)
    csr = sf.CsrMatrix.new(None, rs, cols, vals)
    modes = np.arange(12 * 6, dtype=np.float64) % 7 + 1
    pos = np.array([[i, 0, 0] for i in range(4)], dtype=np.float64)
    with pytest.raises(sf.MultigridError) as e:
        sf.MultigridSetup.new(None, csr, modes[:-18], pos)
    assert e.value.status == 21 and (e.value.a, e.value.b) == (72, 54)
    with pytest.raises(sf.MultigridError) as e:
        sf.MultigridSetup.new(None, csr, modes, pos[:3])
    assert e.value.status == 24 and (e.value.a, e.value.b) == (4, 3)
    # block size 1: not 3x3 nodal
    srs = np.array([0, 2, 5, 8, 11, 14, 16], dtype=np.int64)
    scols = np.array([0, 1, 0, 1, 2, 1, 2, 3, 2, 3, 4, 3, 4, 5, 4, 5], dtype=np.uint32)
    scsr = sf.CsrMatrix.new(None, srs, scols, np.ones(16))
    with pytest.raises(sf.MultigridError) as e:
        sf.MultigridSetup.new(None, scsr, np.ones(36), pos[:2])
    assert e.value.status == 20


def test_a_cancelled_setup_stays_cancelled():
    # solver/multigrid.rs:104-127
    rs, cols, vals = block_chain(4000)
    csr = sf.CsrMatrix.new(None, rs, cols, vals)
    modes = np.arange(12000 * 6, dtype=np.float64) % 7 + 1
    pos = np.array([[i, 0, 0] for i in range(4000)], dtype=np.float64)
    setup = sf.MultigridSetup.new(None, csr, modes, pos)
    polls = [0]

    def cancel():
        polls[0] += 1
        return polls[0] > 10

    with pytest.raises(sf.MultigridError) as e:
        setup.step(2 ** 62, cancel)
    assert e.value.status == 23 and polls[0] >= 11
    with pytest.raises(sf.MultigridError) as e:
        setup.step(2 ** 62)
    assert e.value.status == 23
    with pytest.raises(sf.MultigridError) as e:
        setup.finish()
    assert e.value.status == 25


def test_mesh_refusals_are_named():
    with pytest.raises(sf.MeshError) as e:
        sf.box_tet_mesh((0, 0, 0), (1, 1, 1), (0, 1, 1))
    assert e.value.status == 37
    # an inverted tetrahedron
    with pytest.raises(sf.MeshError) as e:
        sf.TetMesh.new([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], [[0, 2, 1, 3]])
    assert e.value.status == 34
    with pytest.raises(sf.MeshError) as e:
        sf.TetMesh.new([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [5, 5, 5]], [[0, 1, 2, 3]])
    assert e.value.status == 33 and e.value.a == 4
