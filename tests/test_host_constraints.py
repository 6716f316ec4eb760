"""The study's refusal of a model that is not held (crates/solidfem-elastic/src/constraints.rs),
on the host, against the reference's own refusal tests
(crates/solidfem/tests/validation/refusals.rs:29-101, prescribed.rs:117-134) and the closed form
of the rigid motions a set of held components leaves.
"""
import numpy as np
import pytest

import solidfem_amd as sf

NO_FIXED_NODES, BODY_NOT_HELD, DOFS_DO_NOT_HOLD = 40, 51, 52
FREE, ABOUT_POINT, ABOUT_LINE = 0, 1, 2


def bar(size=(4.0, 1.0, 1.0), divisions=(4, 1, 1)):
    return sf.QuadraticTetMesh.from_linear(sf.box_tet_mesh((0.0, 0.0, 0.0), size, divisions))


def loaded_bar():
    # refusals.rs:20-27: traction (1, 0, 0) on the faces at x = 4
    mesh = bar()
    model = sf.Model.new(mesh, 2100.0, 0.3)
    x = mesh.nodes()
    for f in mesh.boundary_faces():
        if all(abs(x[n, 0] - 4.0) < 1e-9 for n in f[2:5]):
            model.add_traction(f, (1.0, 0.0, 0.0))
    return model, x


def refusal(model):
    with pytest.raises(sf.ModelError) as e:
        model.free_problem()
    return e.value


def test_no_fixed_node_is_refused():
    model, _ = loaded_bar()
    assert refusal(model).status == NO_FIXED_NODES


def test_one_fixed_node_leaves_a_rotation_about_it():
    model, _ = loaded_bar()
    model.fix_node(0)
    e = refusal(model)
    assert (e.status, e.value, e.b) == (BODY_NOT_HELD, ABOUT_POINT, 1)


def test_fixed_nodes_along_one_line_leave_a_rotation_about_it():
    model, x = loaded_bar()
    line = [n for n in range(len(x)) if abs(x[n, 1]) < 1e-9 and abs(x[n, 2]) < 1e-9]
    assert len(line) >= 3
    for n in line:
        model.fix_node(n)
    e = refusal(model)
    assert (e.status, e.value, e.b) == (BODY_NOT_HELD, ABOUT_LINE, len(line))


def test_a_clamped_face_holds_the_bar():
    model, x = loaded_bar()
    for n in range(len(x)):
        if abs(x[n, 0]) < 1e-9:
            model.fix_node(n)
    free_map, rhs, *_ = model.free_problem()
    assert np.count_nonzero(free_map == np.iinfo(np.uint32).max) == np.count_nonzero(np.abs(x[:, 0]) < 1e-9)


def two_tets(second):
    positions = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]] + second[0]
    return sf.QuadraticTetMesh.from_linear(sf.TetMesh.new(positions, [[0, 1, 2, 3], second[1]]))


def test_a_second_piece_with_nothing_holding_it_is_refused():
    # refusals.rs:72-101: two separate tetrahedra, only the first held
    mesh = two_tets(([[5.0, 0.0, 0.0], [6.0, 0.0, 0.0], [5.0, 1.0, 0.0], [5.0, 0.0, 1.0]], [4, 5, 6, 7]))
    model = sf.Model.new(mesh, 2100.0, 0.3)
    for n in range(3):
        model.fix_node(n)
    model.add_nodal_force(5, (0.0, 0.0, 1.0))
    e = refusal(model)
    assert (e.status, e.a, e.value, e.b) == (BODY_NOT_HELD, 1, FREE, 0)


def test_a_piece_is_held_through_the_nodes_it_shares_with_a_held_piece():
    # the second element shares only the edge (0, 1) with the first, so it is another body; once
    # the first is held, the shared edge's three nodes hold the second about that line, and one
    # more held node off the line holds it
    second = ([[0.0, -1.0, 0.0], [0.0, 0.0, -1.0]], [0, 1, 4, 5])
    mesh = two_tets(second)
    model = sf.Model.new(mesh, 2100.0, 0.3)
    for n in range(3):
        model.fix_node(n)
    model.add_nodal_force(5, (0.0, 0.0, 1.0))
    e = refusal(model)
    assert (e.status, e.a, e.value, e.b) == (BODY_NOT_HELD, 1, ABOUT_LINE, 3)
    model.fix_node(4)
    free_map, *_ = model.free_problem()
    assert free_map[5] != np.iinfo(np.uint32).max


def cube():
    # prescribed.rs:17-39: the unit cube as six tetrahedra around its diagonal
    points = [[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [1.0, 1.0, 0.0], [0.0, 1.0, 0.0],
              [0.0, 0.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 1.0], [0.0, 1.0, 1.0]]
    tets = [[0, 1, 2, 6], [0, 2, 3, 6], [0, 3, 7, 6], [0, 7, 4, 6], [0, 4, 5, 6], [0, 5, 1, 6]]
    return sf.QuadraticTetMesh.from_linear(sf.TetMesh.new(points, tets))


@pytest.mark.parametrize("axes,rank", [((0,), 3), ((0, 1), 5), ((0, 1, 2), 6)])
def test_prescribed_components_leave_the_rigid_motions_their_rank_says(axes, rank):
    # prescribed.rs:117-134: components prescribed on the face x = 0. Rows (1,0,0,0,z,-y) over a
    # plane x = 0 span translation x and the rotations about y and z (rank 3); adding y
    # components adds translation y and the rotation about x (rank 5); z completes all six.
    mesh = cube()
    model = sf.Model.new(mesh, 200000.0, 0.3)
    x = mesh.nodes()
    for n in range(len(x)):
        if x[n, 0] == 0.0:
            for a in axes:
                model.prescribe_displacement(n, a, 0.01 if a == 0 else 0.0)
    if rank < 6:
        e = refusal(model)
        assert (e.status, e.b) == (DOFS_DO_NOT_HOLD, rank)
    else:
        model.free_problem()
