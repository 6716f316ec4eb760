"""The workloads the tests and the benchmark share (sizes and seeds stated here)."""
import numpy as np

from solidfem_amd import Workload


def loaded(size, divisions):
    """crates/solidfem/tests/validation/multigrid_precision.rs:24-33: steel, clamped at x = 0,
    traction (0, 3, -7) on the far end."""
    return Workload(size=tuple(size), divisions=tuple(divisions), clamp_axis=0, clamp_value=0.0, load_axis=0,
                    load_value=float(size[0]), traction=(0.0, 3.0, -7.0), youngs=200000.0, poisson=0.3)


FINGERPRINT_CANTILEVER = ([10.0, 1.0, 1.0], [12, 3, 3], 61, 4639424340864456019)
FINGERPRINT_PLATE = ([20.0, 20.0, 1.0], [8, 8, 2], 85, 5373319653865270473)


def cantilever_config(divisions=(40, 4, 4)):
    """Config 1: 100x10x10 mm, E 210 GPa, nu 0.3, 1000 N shear on x = 100 (10 MPa over 100 mm^2)."""
    return Workload(size=(100.0, 10.0, 10.0), divisions=tuple(divisions), clamp_axis=0, clamp_value=0.0, load_axis=0,
                    load_value=100.0, traction=(0.0, 0.0, -10.0), youngs=210000.0, poisson=0.3)


def plate_config(divisions=(100, 100, 2), seed=20261202):
    """Config 2: 100x100x0.8 mm plate, E 3.5 GPa, nu 0.36, clamped x = 0, 0.01 MPa pressure on
    z = 0.8, interior vertices jittered by U(-5%, 5%) of the cell size (seed 20261202)."""
    return Workload(size=(100.0, 100.0, 0.8), divisions=tuple(divisions), jitter_fraction=0.05, jitter_seed=seed,
                    clamp_axis=0, clamp_value=0.0, load_axis=2, load_value=0.8, traction=(0.0, 0.0, -0.01),
                    youngs=3500.0, poisson=0.36)


def block_config(case, divisions=(30, 20, 20), seed=7):
    """Config 3: 60x40x40 mm block, E 70 GPa, nu 0.33, clamped z = 0, random tractions N(0, 1 MPa)
    per component on the z = 40 faces, load case `case` (seed 7)."""
    return Workload(size=(60.0, 40.0, 40.0), divisions=tuple(divisions), clamp_axis=2, clamp_value=0.0, load_axis=2,
                    load_value=40.0, traction_kind=1, traction_scale=1.0, traction_seed=seed, load_case=case,
                    youngs=70000.0, poisson=0.33)


def bar_config(divisions=(160, 16, 16)):
    """Config 4: 200x20x20 mm bar, 2 mm shell E 3.5 GPa nu 0.36 around a core E 0.7 GPa nu 0.3,
    clamped x = 0, tangential traction 0.5 MPa about (y, z) = (10, 10) on x = 200."""
    return Workload(size=(200.0, 20.0, 20.0), divisions=tuple(divisions), clamp_axis=0, clamp_value=0.0, load_axis=0,
                    load_value=200.0, traction_kind=2, traction_scale=0.5, axis_u=10.0, axis_v=10.0,
                    material_kind=1, youngs=3500.0, poisson=0.36, core_youngs=700.0, core_poisson=0.3,
                    shell_depth=2.0)


def tube_config(divisions=(50, 6, 6), wall_cells=2, poisson=0.0):
    """A square tube (demo/wasm/src/part.rs:48-185): 100x10x10 mm with 2 mm walls, `wall_cells`
    cells through each wall and the rest of the 6 across over the hollow, steel, clamped x = 0,
    shear traction (0, 0, -1) MPa on the x = 100 end (64 mm^2 of wall: 64 N)."""
    return Workload(size=(100.0, 10.0, 10.0), divisions=tuple(divisions), clamp_axis=0, clamp_value=0.0, load_axis=0,
                    load_value=100.0, traction=(0.0, 0.0, -1.0), youngs=200000.0, poisson=poisson,
                    wall_cells=wall_cells, wall_thickness=2.0)


STRETCH = 0.001


def stretched_cube(divisions=(1, 1, 1)):
    """crates/solidfem/tests/validation/prescribed.rs:16-63: a unit steel cube (E 200 GPa, nu 0.3,
    4-point rule) of box tetrahedra, no clamp and no load; its held components come from
    `stretch_components`."""
    return Workload(size=(1.0, 1.0, 1.0), divisions=tuple(divisions), clamp_axis=-1, load_axis=-1, youngs=200000.0,
                    poisson=0.3)


def stretch_components(nodes):
    """prescribed.rs:40-63: every node of x = 0 held in x, x = 1 stretched by STRETCH in x, the corner
    (0, 0, 0) held in y and z and (0, 1, 0) in z, as (node, axis, value) in the reference's call
    order."""
    out = []
    for n, p in enumerate(nodes):
        if p[0] == 0.0:
            out.append((n, 0, 0.0))
        if p[0] == 1.0:
            out.append((n, 0, STRETCH))
    corner = lambda target: next(n for n, p in enumerate(nodes) if tuple(p) == target)
    out += [(corner((0.0, 0.0, 0.0)), 1, 0.0), (corner((0.0, 0.0, 0.0)), 2, 0.0), (corner((0.0, 1.0, 0.0)), 2, 0.0)]
    return out


def serving_requests(count, rate, seed, min_dofs=5e3, max_dofs=6e5):
    """Synthetic serving traffic: `count` requests with Poisson arrivals at `rate` per second. A
    request is a bar (10:1:1), a square tube (10:1:1, walls a quarter of the side) or a plate
    (1:1:0.02) of Tet10, one kind in three each, with a log-uniform size between min_dofs and max_dofs (5k and 600k DOFs by default),
    E ~ U[1, 210] GPa, nu ~ U[0.25, 0.4], clamped on a random face and loaded on another by random
    tractions N(0, 1 MPa) per component and face; 30% of requests reuse the mesh, material and
    clamp of an earlier request (its cached hierarchy) with new tractions. A tube's clamp and load
    planes are its ends or its outer walls."""
    rng = np.random.default_rng(seed)
    out, systems = [], []
    t = 0.0
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        if systems and rng.random() < 0.3:
            base = systems[rng.integers(len(systems))]
        else:
            dofs = float(np.exp(rng.uniform(np.log(min_dofs), np.log(max_dofs))))
            kind = rng.integers(3)
            tube = {}
            if kind == 0:  # bar: 3 (20k + 1)(2k + 1)^2 DOFs
                k = max(1, int(round((dofs / 240.0) ** (1.0 / 3.0))))
                while k > 1 and 3 * (20 * k + 1) * (2 * k + 1) ** 2 > max_dofs:
                    k -= 1
                size, div = (10.0 * k, 1.0 * k, 1.0 * k), (10 * k, k, k)
            elif kind == 1:  # tube of side k cells, k/4 through each wall: 3 (20k + 1)((2k + 1)^2 - (k - 1)^2) DOFs
                k = max(4, 4 * int(round((dofs / 180.0) ** (1.0 / 3.0) / 4.0)))
                while k > 4 and 3 * (20 * k + 1) * ((2 * k + 1) ** 2 - (k - 1) ** 2) > max_dofs:
                    k -= 4
                size, div = (10.0 * k, 1.0 * k, 1.0 * k), (10 * k, k, k)
                tube = dict(wall_cells=k // 4, wall_thickness=float(k // 4))
            else:  # plate: 3 (2k + 1)^2 5 DOFs
                k = max(2, int(round(np.sqrt(dofs / 60.0))))
                while k > 2 and 15 * (2 * k + 1) ** 2 > max_dofs:
                    k -= 1
                size, div = (1.0 * k, 1.0 * k, 0.4), (k, k, 2)
            clamp = int(rng.integers(6))
            base = dict(size=size, divisions=div, clamp_axis=clamp // 2,
                        clamp_value=0.0 if clamp % 2 == 0 else size[clamp // 2],
                        youngs=float(rng.uniform(1e3, 2.1e5)), poisson=float(rng.uniform(0.25, 0.4)), **tube)
            systems.append(base)
        faces = [f for f in range(6) if f != 2 * base["clamp_axis"] + (0 if base["clamp_value"] == 0.0 else 1)]
        load = faces[rng.integers(len(faces))]
        axis = load // 2
        w = Workload(**base, load_axis=axis, load_value=0.0 if load % 2 == 0 else base["size"][axis],
                        traction_kind=1, traction_scale=1.0, traction_seed=seed, load_case=i)
        out.append((t, w))
    return out
