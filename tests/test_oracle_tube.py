"""The oracle's tube (demo/wasm/src/part.rs:48-185) pinned by its geometry and by beam theory.

The mesh is the box grid with the hollow's cells removed: its element count, its volume, the grid
lines its vertices sit on and the empty hollow follow from the section alone. Clamped at one end
and sheared at the other with nu = 0, the tube is a beam, so its tip deflection is
F L^3 / (3 E I) plus the webs' shear term F L / (G A_webs), with I = (b^4 - (b - 2t)^4) / 12.
"""
import numpy as np

import refcpu
from workloads import tube_config


def test_the_tube_is_the_box_grid_without_its_hollow():
    w = tube_config()
    p = refcpu.Problem(w)
    x, tets = p.nodes(), p.tets()
    # 6 tetrahedra per kept cell: 50 along, 6 x 6 across less the 2 x 2 of the hollow
    assert len(tets) == 6 * 50 * (36 - 4)
    c = x[tets[:, :4]]
    volumes = np.einsum("ij,ij->i", np.cross(c[:, 1] - c[:, 0], c[:, 2] - c[:, 0]), c[:, 3] - c[:, 0]) / 6.0
    assert (volumes > 0).all()
    assert abs(volumes.sum() - 100.0 * (10.0 ** 2 - 6.0 ** 2)) < 1e-9 * 6400.0
    corners = np.unique(tets[:, :4])
    # wall lines 0, 1, 2 and 8, 9, 10; the hollow's span cut once, at 5
    for axis in (1, 2):
        assert sorted(set(x[corners, axis])) == [0.0, 1.0, 2.0, 5.0, 8.0, 9.0, 10.0]
    inside = (x[:, 1] > 2.0) & (x[:, 1] < 8.0) & (x[:, 2] > 2.0) & (x[:, 2] < 8.0)
    assert not inside.any()


def test_a_sheared_tube_bends_as_a_beam():
    w = tube_config()
    disp, r = refcpu.solve(w, precision=0, tolerance=1e-9)
    assert r.status == 0
    x = refcpu.Problem(w).nodes()
    tip = -disp[np.abs(x[:, 0] - 100.0) < 1e-9, 2].mean()
    e, length, force = 200000.0, 100.0, 64.0  # 1 MPa over the 64 mm^2 end annulus
    inertia = (10.0 ** 4 - 6.0 ** 4) / 12.0
    bending = force * length ** 3 / (3.0 * e * inertia)
    shear = force * length / (e / 2.0 * 2.0 * 2.0 * 10.0)  # G = E / 2 at nu = 0; webs 2 x t x h
    # the webs' shear area is uncertain by about t / h (0.3% of the deflection); the 1% band
    # covers that and the discretisation
    assert abs(tip - (bending + shear)) < 0.01 * (bending + shear)
    assert tip > bending
