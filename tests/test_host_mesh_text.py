"""The on-wire text form of a mesh request (CPU only): `vertex-count tet-count`, x y z per vertex,
a b c d per tetrahedron -- the surface text of crates/solidfem-bench/src/bin/gate.rs:9-14 with
tetrahedra. A mesh written and read back is the same mesh to the bit, and malformed text or an
invalid mesh is refused by name (crates/solidfem-mesh/src/error.rs:7-27)."""
import numpy as np
import pytest

import solidfem_amd as sf


def jittered_box():
    mesh = sf.box_tet_mesh((0.0, 0.0, 0.0), (12.0, 3.0, 2.0), (6, 3, 2))
    pos, tets = mesh.positions(), mesh.tets()
    rng = np.random.default_rng(3)
    interior = np.all((pos > 1e-9) & (pos < np.array([12.0, 3.0, 2.0]) - 1e-9), axis=1)
    pos[interior] += rng.uniform(-0.1, 0.1, (interior.sum(), 3))
    return pos, tets


def test_a_mesh_written_and_read_back_is_the_same_mesh():
    pos, tets = jittered_box()
    text = sf.TetMesh.text_of(pos, tets)
    back = sf.TetMesh.read_text(text)
    assert np.array_equal(back.positions().view(np.uint64), pos.view(np.uint64))
    assert np.array_equal(back.tets(), tets)
    q = sf.QuadraticTetMesh.from_linear(back)
    r = sf.QuadraticTetMesh.from_linear(sf.TetMesh.new(pos, tets))
    assert np.array_equal(q.nodes().view(np.uint64), r.nodes().view(np.uint64))
    # any white space separates the numbers
    assert np.array_equal(sf.TetMesh.read_text(" \t".join(text.split())).tets(), tets)


@pytest.mark.parametrize("text,line", [
    # a token missing at the end is reported on the line after the last line break
    ("", 1), ("4", 1), ("4 1\n0 0 0\n1 0 0\n0 1 0\n", 5), ("4 1\n0 0 zero\n", 2), ("-4 1\n", 1),
    ("4 1\n0 0 0\n1 0 0\n0 1 0\n0 0 1\n0 1 2 3\n7\n", 7), ("4 1\n0 0 0\n1 0 0\n0 1 0\n0 0 1\n0 1 2 x\n", 6),
])
def test_malformed_text_is_refused_with_its_line(text, line):
    with pytest.raises(sf.SolidfemError) as e:
        sf.TetMesh.read_text(text)
    assert e.value.status == 1 and e.value.a == line, e.value


def test_an_invalid_mesh_is_refused_as_tet_mesh_new_refuses_it():
    good = "4 1\n0 0 0\n1 0 0\n0 1 0\n0 0 1\n"
    with pytest.raises(sf.MeshError) as e:
        sf.TetMesh.read_text(good + "0 1 2 9\n")
    assert e.value.status == 32  # NodeOutOfRange
    with pytest.raises(sf.MeshError) as e:
        sf.TetMesh.read_text("4 1\n0 0 0\n1 0 0\n0 1 0\n1 1 0\n0 1 2 3\n")
    assert e.value.status in (34, 35)  # degenerate or flat
    with pytest.raises(sf.MeshError) as e:
        sf.TetMesh.read_text("4 1\n0 0 0\n1 0 0\n0 1 0\n0 0 nan\n0 1 2 3\n")
    assert e.value.status == 31  # non-finite position
    assert sf.TetMesh.read_text(good + "0 1 2 3\n").tets().shape == (1, 4)
