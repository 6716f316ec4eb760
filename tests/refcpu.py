"""ctypes binding of the oracle (oracle/librefcpu.so): TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it, as the checker.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "oracle" / "librefcpu.so"

u32, u64, i32, i64, f64 = C.c_uint32, C.c_uint64, C.c_int32, C.c_int64, C.c_double
P, vp = C.POINTER, C.c_void_p


class RefWorkload(C.Structure):
    # the same layout as sf_workload
    from solidfem_amd._capi import Workload as _W
    _fields_ = _W._fields_


class RefOptions(C.Structure):
    _fields_ = [("neighbour_reach", f64), ("coarse_dofs", i64), ("max_levels", i64), ("smoother_degree", i64),
                ("smoother_range", f64), ("precision", i32), ("reserved", i32), ("threads", i64)]


class RefResult(C.Structure):
    _fields_ = [("status", i32), ("iterations", u64), ("relative_residual", f64), ("at_rounding_floor", i32),
                ("fingerprint", u64), ("total_dofs", i64), ("free_dofs", i64), ("nonzeros", i64), ("levels", i64),
                ("coarse_dofs", i64), ("operator_complexity", f64), ("memory_bytes", i64),
                ("assembly_seconds", f64), ("setup_seconds", f64), ("solve_seconds", f64)]


_SIG = {
    "ref_default_workload": (None, [P(RefWorkload)]),
    "ref_default_mg_options": (None, [P(RefOptions)]),
    "ref_problem_create": (vp, [P(RefWorkload)]),
    "ref_problem_destroy": (None, [vp]),
    "ref_problem_create_prescribed": (vp, [P(RefWorkload), P(u32), P(u32), P(f64), i64]),
    "ref_problem_node_count": (i64, [vp]),
    "ref_problem_vertex_count": (i64, [vp]),
    "ref_problem_tet_count": (i64, [vp]),
    "ref_problem_free_dofs": (i64, [vp]),
    "ref_problem_nonzeros": (i64, [vp]),
    "ref_problem_boundary_face_count": (i64, [vp]),
    "ref_problem_nodes": (None, [vp, P(f64)]),
    "ref_problem_tets": (None, [vp, P(u32)]),
    "ref_problem_free_map": (None, [vp, P(u32)]),
    "ref_problem_boundary_faces": (None, [vp, P(u32)]),
    "ref_problem_loaded_face_count": (i64, [vp]),
    "ref_problem_loaded_faces": (None, [vp, P(u32), P(f64)]),
    "ref_problem_csr": (None, [vp, P(i64), P(u32), P(f64)]),
    "ref_problem_rhs": (None, [vp, P(f64)]),
    "ref_problem_rigid_modes": (None, [vp, P(f64)]),
    "ref_problem_element_stiffness": (i32, [vp, i64, P(f64)]),
    "ref_recover": (i32, [vp, P(f64), P(f64), P(f64), P(i64)]),
    "ref_error_estimate": (i32, [vp, P(f64), P(f64), P(f64), P(i64)]),
    "ref_compare": (i32, [vp, P(f64), u64, P(f64), u64, P(C.c_uint8), P(C.c_uint8), vp, P(f64), u64, P(f64), u64,
                          P(C.c_uint8), P(C.c_uint8), P(f64), P(i64)]),
    "ref_problem_quadrature": (i32, [vp]),
    "ref_peak": (i32, [P(f64), u64, P(f64), u64, P(f64), u64, P(C.c_uint8), P(C.c_uint8), P(i32), P(u64), P(f64), P(f64)]),
    "ref_nodal_rotations": (i32, [vp, P(f64), P(f64), P(i64)]),
    "ref_multigrid_create": (vp, [vp, P(RefOptions), P(i32)]),
    "ref_multigrid_destroy": (None, [vp]),
    "ref_multigrid_levels": (i64, [vp]),
    "ref_multigrid_threads": (i64, [vp]),
    "ref_default_threads": (i64, []),
    "ref_multigrid_setup_seconds": (f64, [vp]),
    "ref_multigrid_level_info": (None, [vp, i64, P(i64)]),
    "ref_multigrid_level_matrix": (None, [vp, i64, P(i64), P(u32), P(f64)]),
    "ref_multigrid_level_prolongator": (None, [vp, i64, P(i64), P(u32), P(f64)]),
    "ref_multigrid_level_inverse_diagonal": (None, [vp, i64, P(f64)]),
    "ref_multigrid_level_spectrum": (None, [vp, i64, P(f64)]),
    "ref_multigrid_level_aggregates": (None, [vp, i64, P(u32)]),
    "ref_multigrid_coarse_factor": (i32, [vp, P(f64)]),
    "ref_multigrid_apply": (None, [vp, P(f64), P(f64)]),
    "ref_multigrid_smooth": (None, [vp, i64, P(f64), P(f64)]),
    "ref_multigrid_stats": (None, [vp, P(i64), P(i64), P(f64), P(i64)]),
    "ref_solve": (i32, [P(RefWorkload), P(RefOptions), f64, P(f64), P(RefResult)]),
    "ref_solve_with": (i32, [vp, vp, f64, u64, P(f64), P(RefResult)]),
    "ref_solve_rhs": (i32, [vp, vp, P(f64), f64, P(f64), P(RefResult)]),
    "ref_fingerprint": (u64, [P(f64), i64]),
    "ref_pcg_jacobi": (i32, [i64, P(i64), P(u32), P(f64), P(f64), f64, u64, P(f64), P(RefResult)]),
    "ref_csr_multiply": (None, [i64, P(i64), P(u32), P(f64), P(f64), P(f64)]),
    "ref_csr_rounding_bound": (f64, [i64, P(i64), P(u32), P(f64), P(f64)]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"{LIB} is missing: run make -C oracle")
        l = C.CDLL(str(LIB))
        for name, (res, args) in _SIG.items():
            f = getattr(l, name)
            f.restype, f.argtypes = res, args
        _lib = l
    return _lib


def ptr(a, ctype):
    return a.ctypes.data_as(P(ctype))


def workload(w) -> RefWorkload:
    """A solidfem_amd.Workload as the oracle's struct."""
    return w.c(RefWorkload)


def options(precision=1, coarse_dofs=1500, degree=3, reach=3.0, max_levels=12, smoother_range=30.0, threads=1) -> RefOptions:
    """MultigridOptions; `threads` 0 takes the reference's default, min(logical / 2, 8)."""
    o = RefOptions()
    lib().ref_default_mg_options(C.byref(o))
    o.precision, o.coarse_dofs, o.smoother_degree, o.threads = precision, coarse_dofs, degree, threads
    o.neighbour_reach, o.max_levels, o.smoother_range = reach, max_levels, smoother_range
    return o


class Problem:
    def __init__(self, w, prescribed=None):
        """`prescribed`: (node, axis, value) components (Model::prescribe_displacement), eliminated
        once assembled as crates/solidfem-elastic/src/study.rs:567-593 does."""
        self.w = workload(w)
        if prescribed:
            nodes = np.array([c[0] for c in prescribed], dtype=np.uint32)
            axes = np.array([c[1] for c in prescribed], dtype=np.uint32)
            values = np.array([c[2] for c in prescribed], dtype=np.float64)
            self._p = lib().ref_problem_create_prescribed(C.byref(self.w), ptr(nodes, u32), ptr(axes, u32), ptr(values, f64),
                                                          len(nodes))
        else:
            self._p = lib().ref_problem_create(C.byref(self.w))
        if not self._p:
            raise RuntimeError("the oracle refused the workload")

    def __del__(self):
        if getattr(self, "_p", None):
            lib().ref_problem_destroy(self._p)
            self._p = None

    @property
    def node_count(self):
        return lib().ref_problem_node_count(self._p)

    @property
    def free_dofs(self):
        return lib().ref_problem_free_dofs(self._p)

    def nodes(self):
        out = np.empty((self.node_count, 3))
        lib().ref_problem_nodes(self._p, ptr(out, f64))
        return out

    def tets(self):
        out = np.empty((lib().ref_problem_tet_count(self._p), 10), dtype=np.uint32)
        lib().ref_problem_tets(self._p, ptr(out, u32))
        return out

    def free_map(self):
        out = np.empty(self.node_count, dtype=np.uint32)
        lib().ref_problem_free_map(self._p, ptr(out, u32))
        return out

    def boundary_faces(self):
        out = np.empty((lib().ref_problem_boundary_face_count(self._p), 8), dtype=np.uint32)
        lib().ref_problem_boundary_faces(self._p, ptr(out, u32))
        return out

    def loaded_faces(self):
        n = lib().ref_problem_loaded_face_count(self._p)
        faces = np.empty((n, 2), dtype=np.uint32)
        t = np.empty((n, 3))
        lib().ref_problem_loaded_faces(self._p, ptr(faces, u32), ptr(t, f64))
        return faces, t

    def csr(self):
        n = self.free_dofs
        rs = np.empty(n + 1, dtype=np.int64)
        nnz = lib().ref_problem_nonzeros(self._p)
        cols = np.empty(nnz, dtype=np.uint32)
        vals = np.empty(nnz)
        lib().ref_problem_csr(self._p, ptr(rs, i64), ptr(cols, u32), ptr(vals, f64))
        return rs, cols, vals

    def rhs(self):
        out = np.empty(self.free_dofs)
        lib().ref_problem_rhs(self._p, ptr(out, f64))
        return out

    def rigid_modes(self):
        out = np.empty(6 * self.free_dofs)
        lib().ref_problem_rigid_modes(self._p, ptr(out, f64))
        return out

    @property
    def vertex_count(self):
        return lib().ref_problem_vertex_count(self._p)

    def recover(self, u):
        """Stress recovery of nodal displacements: (points (elements, points, 10), nodal (nodes, 7))."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        q = lib().ref_problem_quadrature(self._p)
        points = np.empty((lib().ref_problem_tet_count(self._p), q, 10))
        nodal = np.empty((self.node_count, 7))
        fault = C.c_int64(-1)
        if lib().ref_recover(self._p, ptr(u, f64), ptr(points, f64), ptr(nodal, f64), C.byref(fault)) != 0:
            raise RuntimeError(f"element {fault.value} is inverted or has an invalid material")
        return points, nodal

    def error_estimate(self, u):
        """The error estimate of displacements u (error_estimate.rs:121-171): a dict with status (0,
        or the refusal's status code), a, b, the per-element errors done and the three sums."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        errors = np.empty(lib().ref_problem_tet_count(self._p))
        sums = np.zeros(3)
        payload = np.zeros(3, dtype=np.int64)
        status = lib().ref_error_estimate(self._p, ptr(u, f64), ptr(errors, f64), ptr(sums, f64), ptr(payload, i64))
        return dict(status=status, a=int(payload[0]), b=int(payload[1]), element_errors=errors[:payload[2]].copy(),
                    error_norm_squared=sums[0], solution_norm_squared=sums[1], relative_error=sums[2])

    def nodal_rotations(self, u):
        """nodal_rotations (results.rs:262-305): (nodes, 3)."""
        u = np.ascontiguousarray(u, dtype=np.float64).reshape(-1)
        out = np.empty((self.node_count, 3))
        fault = C.c_int64(-1)
        if lib().ref_nodal_rotations(self._p, ptr(u, f64), ptr(out, f64), C.byref(fault)) != 0:
            raise RuntimeError(f"element {fault.value} is inverted at a node")
        return out

    def element_stiffness(self, tet):
        ke = np.empty((30, 30))
        if lib().ref_problem_element_stiffness(self._p, tet, ptr(ke, f64)) != 0:
            raise RuntimeError("element fault")
        return ke

    def free_positions(self):
        fm = self.free_map()
        return self.nodes()[fm != 0xFFFFFFFF]


class Multigrid:
    def __init__(self, problem: Problem, opts: RefOptions):
        st = i32()
        self.problem = problem
        self._m = lib().ref_multigrid_create(problem._p, C.byref(opts), C.byref(st))
        if not self._m:
            raise RuntimeError(f"the oracle's setup failed ({st.value})")

    def __del__(self):
        if getattr(self, "_m", None):
            lib().ref_multigrid_destroy(self._m)
            self._m = None

    @property
    def levels(self):
        return lib().ref_multigrid_levels(self._m)

    @property
    def threads(self):
        return lib().ref_multigrid_threads(self._m)

    def info(self, level):
        out = np.empty(5, dtype=np.int64)
        lib().ref_multigrid_level_info(self._m, level, ptr(out, i64))
        return dict(rows=int(out[0]), rb=int(out[1]), blocks=int(out[2]), p_blocks=int(out[3]), factor=bool(out[4]))

    def matrix(self, level):
        inf = self.info(level)
        starts = np.empty(inf["rows"] + 1, dtype=np.int64)
        cols = np.empty(inf["blocks"], dtype=np.uint32)
        vals = np.empty(inf["blocks"] * inf["rb"] ** 2)
        lib().ref_multigrid_level_matrix(self._m, level, ptr(starts, i64), ptr(cols, u32), ptr(vals, f64))
        return starts, cols, vals

    def prolongator(self, level):
        inf = self.info(level)
        starts = np.empty(inf["rows"] + 1, dtype=np.int64)
        cols = np.empty(inf["p_blocks"], dtype=np.uint32)
        vals = np.empty(inf["p_blocks"] * inf["rb"] * 6)
        lib().ref_multigrid_level_prolongator(self._m, level, ptr(starts, i64), ptr(cols, u32), ptr(vals, f64))
        return starts, cols, vals

    def inverse_diagonal(self, level):
        inf = self.info(level)
        out = np.empty(inf["rows"] * inf["rb"] ** 2)
        lib().ref_multigrid_level_inverse_diagonal(self._m, level, ptr(out, f64))
        return out

    def spectrum(self, level):
        out = np.empty(2)
        lib().ref_multigrid_level_spectrum(self._m, level, ptr(out, f64))
        return float(out[0]), float(out[1])

    def aggregates(self, level):
        out = np.empty(self.info(level)["rows"], dtype=np.uint32)
        lib().ref_multigrid_level_aggregates(self._m, level, ptr(out, u32))
        return out

    def coarse_factor(self):
        inf = self.info(self.levels - 1)
        n = inf["rows"] * inf["rb"]
        out = np.empty((n, n))
        return out if lib().ref_multigrid_coarse_factor(self._m, ptr(out, f64)) else None

    def apply(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        lib().ref_multigrid_apply(self._m, ptr(r, f64), ptr(z, f64))
        return z

    def smooth(self, level, b):
        b = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(b)
        lib().ref_multigrid_smooth(self._m, level, ptr(b, f64), ptr(x, f64))
        return x

    def solve_rhs(self, rhs, tolerance):
        """The problem's matrix solved for another right-hand side over the free unknowns."""
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        x = np.empty_like(b)
        r = RefResult()
        lib().ref_solve_rhs(self.problem._p, self._m, ptr(b, f64), tolerance, ptr(x, f64), C.byref(r))
        return x, r

    def solve(self, tolerance, iteration_limit=0):
        disp = np.empty(3 * self.problem.node_count)
        r = RefResult()
        lib().ref_solve_with(self.problem._p, self._m, tolerance, iteration_limit, ptr(disp, f64), C.byref(r))
        return disp.reshape(-1, 3), r


def solve(w, precision=0, coarse_dofs=1500, tolerance=1e-9, degree=3, threads=1):
    disp = None
    p = Problem(w)
    disp = np.empty(3 * p.node_count)
    r = RefResult()
    lib().ref_solve(C.byref(p.w), C.byref(options(precision, coarse_dofs, degree, threads=threads)), tolerance, ptr(disp, f64),
                    C.byref(r))
    return disp.reshape(-1, 3), r


def fingerprint(values: np.ndarray) -> int:
    v = np.ascontiguousarray(values, dtype=np.float64).ravel()
    return int(lib().ref_fingerprint(ptr(v, f64), len(v)))


def csr_multiply(rs, cols, vals, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(len(rs) - 1)
    lib().ref_csr_multiply(len(rs) - 1, ptr(rs, i64), ptr(cols, u32), ptr(vals, f64), ptr(x, f64), ptr(y, f64))
    return y


def csr_rounding_bound(rs, cols, vals, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().ref_csr_rounding_bound(len(rs) - 1, ptr(rs, i64), ptr(cols, u32), ptr(vals, f64), ptr(x, f64))


def pcg_jacobi(rs, cols, vals, rhs, tolerance=1e-9, slice_iterations=0):
    n = len(rs) - 1
    x = np.empty(n)
    r = RefResult()
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    lib().ref_pcg_jacobi(n, ptr(rs, i64), ptr(cols, u32), ptr(vals, f64), ptr(rhs, f64), tolerance, slice_iterations,
                         ptr(x, f64), C.byref(r))
    return x, r


def compare(coarse, fine, inside=None):
    """MeshComparison::between (comparison.rs:101-141) in the oracle. Each side is (Problem,
    nodal_values (nodes, 7), point_values (elements, points, 10)); `inside` is evaluated here into
    masks. A dict with status (0, or the refusal's status code) and per side the peak and counts."""
    sides = []
    for problem, nodal, points in (coarse, fine):
        nodal = np.ascontiguousarray(nodal, dtype=np.float64)
        pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 10)
        x = problem.nodes()
        nm = None if inside is None else np.ascontiguousarray([bool(inside(p)) for p in x], dtype=np.uint8)
        pm = None if inside is None else np.ascontiguousarray([bool(inside(p[:3])) for p in pts], dtype=np.uint8)
        vm = np.ascontiguousarray(nodal[:, 6])
        sides.append((problem, vm, pts, nm, pm))
    values = np.zeros(13)
    integers = np.zeros(8, dtype=np.int64)
    u8 = C.c_uint8
    args = []
    for problem, vm, pts, nm, pm in sides:
        args += [problem._p, ptr(vm, f64), 1, ptr(pts, f64), pts.shape[0], None if nm is None else ptr(nm, u8),
                 None if pm is None else ptr(pm, u8)]
    status = lib().ref_compare(*args, ptr(values, f64), ptr(integers, i64))
    out = dict(status=status, a=int(integers[6]), b=int(integers[7]), value=values[12], peak_change=values[10],
               peak_distance=values[11])
    for w, name in enumerate(("coarse", "fine")):
        out[name] = dict(von_mises=values[5 * w], position=values[5 * w + 1:5 * w + 4].copy(), volume=values[5 * w + 4],
                         elements=int(integers[3 * w]), at_point=bool(integers[3 * w + 1]), index=int(integers[3 * w + 2]))
    return out


def peak(positions, nodal_von_mises, point_values, node_inside=None, point_inside=None):
    """peak_with_holder (results.rs:184-226) over the candidates the boolean masks keep (None: all):
    None, or (von_mises, position (3,), at_point, index)."""
    x = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    vm = np.ascontiguousarray(nodal_von_mises, dtype=np.float64).reshape(-1)
    pts = np.ascontiguousarray(point_values, dtype=np.float64).reshape(-1, 10)
    masks = []
    for m, n in ((node_inside, x.shape[0]), (point_inside, pts.shape[0])):
        masks.append(None if m is None else np.ascontiguousarray(np.broadcast_to(m, (n,)), dtype=np.uint8))
    at_point, index, value, position = C.c_int32(), C.c_uint64(), C.c_double(), np.empty(3)
    found = lib().ref_peak(ptr(x, f64), x.shape[0], ptr(vm, f64), 1, ptr(pts, f64), pts.shape[0],
                           None if masks[0] is None else ptr(masks[0], C.c_uint8),
                           None if masks[1] is None else ptr(masks[1], C.c_uint8),
                           C.byref(at_point), C.byref(index), C.byref(value), ptr(position, f64))
    if not found:
        return None
    return value.value, position, bool(at_point.value), index.value
