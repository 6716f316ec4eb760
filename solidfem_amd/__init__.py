"""solidfem_amd: the MI355X-native linear-static solve path of solidfem.

A Python mirror of the reference's Rust API (crates/solidfem-solve/src/lib.rs:28-34,
crates/solidfem-elastic/src/lib.rs:44-66) over the C ABI in include/solidfem_hip.h. Every
compute call runs in libsolidfem_hip.so on the GPU: Tet10 stiffness assembly into 3x3 blocks,
the block smoothed-aggregation multigrid V-cycle with its Chebyshev smoother, and the
resumable, cancellable, slice-invariant conjugate gradient solve. The multigrid setup's phase
machine and aggregation run on host threads inside the same library, and the numeric phases of
its large levels on the GPU.
"""
from __future__ import annotations

import ctypes as C
import enum
import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import _capi

__all__ = [
    "Context", "Precision", "DotOrder", "PcgOptions", "MultigridOptions", "Workload", "TetMesh", "QuadraticTetMesh",
    "box_tet_mesh", "Model", "StudyBuilder", "StaticStudy", "CsrMatrix", "MultigridSetup", "Multigrid", "Pcg",
    "StepOutcome", "SolidfemError", "PcgError", "MultigridError", "ModelError", "MeshError", "library",
    "SetupDevice", "Recovery", "PeakStress", "peak_stress", "peak_von_mises",
]


def library():
    return _capi.load()


class Precision(enum.IntEnum):
    """crates/solidfem-solve/src/multigrid/mod.rs:126-141"""
    DOUBLE = 0
    SINGLE_STORAGE = 1
    SINGLE = 2


class SetupDevice(enum.IntEnum):
    """Where a level's numeric setup phases run (power iteration, smoothed prolongator, A P,
    P^T A P): AUTO on the device for levels of at least 4096 block rows when the setup has a
    device context, HOST never, DEVICE for every level. The hierarchy's bits are the same."""
    AUTO = 0
    HOST = 1
    DEVICE = 2


class DotOrder(enum.IntEnum):
    TREE = 0
    REFERENCE = 1


class SolidfemError(RuntimeError):
    def __init__(self, status: int, message: str, a: int = 0, b: int = 0, value: float = 0.0):
        super().__init__(f"[{status}] {message}")
        self.status, self.message, self.a, self.b, self.value = status, message, a, b, value


class PcgError(SolidfemError):
    """pcg.rs:59-77: NotPositiveDefinite (10) or NotConverged (11)."""


class MultigridError(SolidfemError):
    """multigrid/mod.rs:197-246"""


class ModelError(SolidfemError):
    """crates/solidfem-elastic/src/error.rs:23-95"""


class MeshError(SolidfemError):
    """crates/solidfem-mesh/src/error.rs:7-27"""


def _raise(status: int):
    if status == 0:
        return
    lib = library()
    a, b, v = C.c_int64(), C.c_int64(), C.c_double()
    lib.sf_last_error_payload(C.byref(a), C.byref(b), C.byref(v))
    msg = lib.sf_last_error().decode()
    cls = SolidfemError
    if 10 <= status < 20 or status == 29:  # 29: a caller's preconditioner failed inside a step
        cls = PcgError
    elif 20 <= status < 30:
        cls = MultigridError
    elif 30 <= status < 40:
        cls = MeshError
    elif 40 <= status < 60 or 70 <= status < 80:
        cls = ModelError
    raise cls(status, msg, a.value, b.value, v.value)


def _ptr(array: np.ndarray, ctype):
    return array.ctypes.data_as(C.POINTER(ctype))


def _cancel(cancel: Optional[Callable[[], bool]]):
    if cancel is None:
        return _capi.CANCEL_FN()  # NULL: nothing to poll, no call per unit of work
    return _capi.CANCEL_FN(lambda _u: 1 if cancel() else 0)


@dataclass
class PcgOptions:
    """pcg.rs:14-33"""
    tolerance: float = 1e-9
    max_iterations: int = 0  # 0: 10 n + 1000
    dot_order: DotOrder = DotOrder.TREE

    def c(self):
        return _capi.PcgOptions(self.tolerance, self.max_iterations, int(self.dot_order), 0)


@dataclass
class MultigridOptions:
    """multigrid/mod.rs:143-182"""
    neighbour_reach: float = 3.0
    coarse_dofs: int = 1500
    max_levels: int = 12
    smoother_degree: int = 3
    smoother_range: float = 30.0
    precision: Precision = Precision.SINGLE_STORAGE
    threads: int = 0
    coarse_solve: int = 0  # 0: reference order, 1: explicit inverse
    setup_device: "SetupDevice" = None  # where the numeric setup phases run (default AUTO)

    def c(self):
        placement = SetupDevice.AUTO if self.setup_device is None else self.setup_device
        return _capi.MultigridOptions(self.neighbour_reach, self.coarse_dofs, self.max_levels, self.smoother_degree,
                                      self.smoother_range, int(self.precision), self.threads, self.coarse_solve, int(placement))


@dataclass
class Workload:
    """A structured box or tube problem (see include/solidfem_hip.h: sf_workload)."""
    size: tuple = (1.0, 1.0, 1.0)
    divisions: tuple = (1, 1, 1)
    origin: tuple = (0.0, 0.0, 0.0)
    jitter_fraction: float = 0.0
    jitter_seed: int = 0
    clamp_axis: int = 0
    clamp_value: float = 0.0
    load_axis: int = 0
    load_value: float = 1.0
    traction_kind: int = 0
    traction: tuple = (0.0, 0.0, 0.0)
    traction_scale: float = 0.0
    traction_seed: int = 0
    load_case: int = 0
    axis_u: float = 0.0
    axis_v: float = 0.0
    material_kind: int = 0
    youngs: float = 200000.0
    poisson: float = 0.3
    core_youngs: float = 1.0
    core_poisson: float = 0.0
    shell_depth: float = 0.0
    quadrature_points: int = 4
    wall_cells: int = 0  # a tube when positive: cells through each wall across the longest axis
    wall_thickness: float = 0.0

    def fields(self):
        return dict(origin=self.origin, size=self.size, divisions=self.divisions, jitter_fraction=self.jitter_fraction,
                    jitter_seed=self.jitter_seed, clamp_axis=self.clamp_axis, clamp_value=self.clamp_value,
                    load_axis=self.load_axis, load_value=self.load_value, traction_kind=self.traction_kind,
                    traction=self.traction, traction_scale=self.traction_scale, traction_seed=self.traction_seed,
                    load_case=self.load_case, axis_u=self.axis_u, axis_v=self.axis_v, material_kind=self.material_kind,
                    youngs=self.youngs, poisson=self.poisson, core_youngs=self.core_youngs,
                    core_poisson=self.core_poisson, shell_depth=self.shell_depth,
                    quadrature_points=self.quadrature_points, wall_cells=self.wall_cells,
                    wall_thickness=self.wall_thickness)

    def c(self, struct=None):
        w = (struct or _capi.Workload)()
        for name, value in self.fields().items():
            if isinstance(value, tuple):
                arr = getattr(w, name)
                for i, x in enumerate(value):
                    arr[i] = x
            else:
                setattr(w, name, value)
        return w


class _Handle:
    _destroy = ""

    def __init__(self, ptr):
        self._ptr = ptr

    def close(self):
        if getattr(self, "_ptr", None):
            getattr(library(), self._destroy)(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Context(_Handle):
    """One GPU and one stream."""
    _destroy = "sf_context_destroy"

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        _raise(library().sf_context_create(device, C.byref(h)))
        super().__init__(h)

    def synchronize(self):
        _raise(library().sf_context_synchronize(self._ptr))

    def stream_copy(self, nbytes: int, repeats: int = 10) -> float:
        """Average seconds of a plain device copy of `nbytes` (read + write)."""
        t = C.c_double()
        _raise(library().sf_stream_copy(self._ptr, nbytes, repeats, C.byref(t)))
        return t.value

    def stream_read(self, nbytes: int, repeats: int = 10) -> float:
        """Average seconds of a plain device read of `nbytes`."""
        t = C.c_double()
        _raise(library().sf_stream_read(self._ptr, nbytes, repeats, C.byref(t)))
        return t.value


class TetMesh(_Handle):
    """crates/solidfem-mesh/src/tet.rs:50-130"""
    _destroy = "sf_tet_mesh_destroy"

    @classmethod
    def new(cls, positions, tets):
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        tt = np.ascontiguousarray(tets, dtype=np.uint32).reshape(-1, 4)
        h = C.c_void_p()
        _raise(library().sf_tet_mesh_create(_ptr(pos, C.c_double), len(pos), _ptr(tt, C.c_uint32), len(tt), C.byref(h)))
        return cls(h)

    @classmethod
    def read_text(cls, text):
        """A mesh from its text form (`vertex-count tet-count`, x y z per vertex, a b c d per
        tetrahedron): the on-wire form of a mesh request."""
        data = text.encode() if isinstance(text, str) else bytes(text)
        h = C.c_void_p()
        _raise(library().sf_tet_mesh_read_text(data, len(data), C.byref(h)))
        return cls(h)

    def positions(self) -> np.ndarray:
        out = np.empty((library().sf_tet_mesh_node_count(self._ptr), 3))
        library().sf_tet_mesh_positions(self._ptr, _ptr(out, C.c_double))
        return out

    def tets(self) -> np.ndarray:
        out = np.empty((library().sf_tet_mesh_tet_count(self._ptr), 4), dtype=np.uint32)
        library().sf_tet_mesh_tets(self._ptr, _ptr(out, C.c_uint32))
        return out

    @staticmethod
    def text_of(positions, tets) -> str:
        """The text form of a mesh, with every coordinate written exactly (17 significant digits)."""
        pos = np.asarray(positions, dtype=np.float64).reshape(-1, 3)
        tt = np.asarray(tets, dtype=np.uint32).reshape(-1, 4)
        lines = [f"{len(pos)} {len(tt)}"]
        lines += [f"{x:.17g} {y:.17g} {z:.17g}" for x, y, z in pos]
        lines += [f"{a} {b} {c} {d}" for a, b, c, d in tt]
        return "\n".join(lines) + "\n"


def box_tet_mesh(origin, size, divisions) -> TetMesh:
    """crates/solidfem-mesh/src/box_mesh.rs:22-80"""
    o = (C.c_double * 3)(*origin)
    s = (C.c_double * 3)(*size)
    d = (C.c_int32 * 3)(*divisions)
    h = C.c_void_p()
    _raise(library().sf_box_tet_mesh(o, s, d, C.byref(h)))
    return TetMesh(h)


class QuadraticTetMesh(_Handle):
    """crates/solidfem-mesh/src/quadratic.rs:13-125"""
    _destroy = "sf_quad_mesh_destroy"

    def __init__(self, ptr, owned=True):
        super().__init__(ptr)
        self._owned = owned

    def close(self):
        if self._owned:
            super().close()

    @classmethod
    def from_linear(cls, mesh: TetMesh):
        h = C.c_void_p()
        _raise(library().sf_quad_mesh_from_linear(mesh._ptr, C.byref(h)))
        return cls(h)

    def nodes(self) -> np.ndarray:
        n = library().sf_quad_mesh_node_count(self._ptr)
        out = np.empty((n, 3), dtype=np.float64)
        library().sf_quad_mesh_nodes(self._ptr, _ptr(out, C.c_double))
        return out

    def tets(self) -> np.ndarray:
        n = library().sf_quad_mesh_tet_count(self._ptr)
        out = np.empty((n, 10), dtype=np.uint32)
        library().sf_quad_mesh_tets(self._ptr, _ptr(out, C.c_uint32))
        return out

    def vertex_count(self) -> int:
        return library().sf_quad_mesh_vertex_count(self._ptr)

    def boundary_faces(self) -> np.ndarray:
        """(tet, face, 6 nodes) per face, in the reference's order."""
        n = library().sf_quad_mesh_boundary_faces(self._ptr, None)
        out = np.empty((n, 8), dtype=np.uint32)
        library().sf_quad_mesh_boundary_faces(self._ptr, _ptr(out, C.c_uint32))
        return out


class Model(_Handle):
    """crates/solidfem-elastic/src/study.rs:45-105"""
    _destroy = "sf_model_destroy"

    @classmethod
    def new(cls, mesh: QuadraticTetMesh, youngs_modulus: float, poisson_ratio: float):
        h = C.c_void_p()
        _raise(library().sf_model_create(mesh._ptr, youngs_modulus, poisson_ratio, C.byref(h)))
        return cls(h)

    @classmethod
    def from_workload(cls, workload: Workload):
        h = C.c_void_p()
        w = workload.c()
        _raise(library().sf_model_from_workload(C.byref(w), C.byref(h)))
        return cls(h)

    def mesh(self) -> QuadraticTetMesh:
        """The model's mesh, a view that keeps the model alive."""
        m = QuadraticTetMesh(library().sf_model_mesh(self._ptr), owned=False)
        m._parent = self
        return m

    def set_quadrature(self, points: int):
        _raise(library().sf_model_set_quadrature(self._ptr, points))

    def fix_node(self, node: int):
        _raise(library().sf_model_fix_node(self._ptr, node))

    def add_nodal_force(self, node: int, force):
        _raise(library().sf_model_add_nodal_force(self._ptr, node, (C.c_double * 3)(*force)))

    def prescribe_displacement(self, node: int, axis: int, value: float):
        _raise(library().sf_model_prescribe_displacement(self._ptr, node, axis, value))

    def add_traction(self, face, traction):
        _raise(library().sf_model_add_traction(self._ptr, (C.c_uint32 * 8)(*[int(v) for v in face]),
                                               (C.c_double * 3)(*traction)))

    def set_material_field(self, at: Callable[[np.ndarray], tuple]):
        """Model::new with a caller's MaterialField (material.rs:35-50): `at(position)` gives
        (youngs_modulus, poisson_ratio); it is evaluated once at every integration point of both
        rules, and the model keeps the values."""
        failure = []

        def field(p, out, _user):
            try:
                e, nu = at(np.array(p[:3]))
                out[0], out[1] = float(e), float(nu)
            except BaseException as exc:  # the callback cannot raise through C
                failure.append(exc)
                out[0], out[1] = math.nan, math.nan

        status = library().sf_model_set_material_field(self._ptr, _capi.MATERIAL_FN(field), None)
        if failure:
            raise failure[0]
        _raise(status)

    def set_shell_core(self, shell_depth: float, core_youngs: float, core_poisson: float):
        _raise(library().sf_model_set_shell_core(self._ptr, shell_depth, core_youngs, core_poisson))

    def element_stiffness(self, ctx: Context, first: int, count: int) -> np.ndarray:
        """element.rs:98-133 on the device: (count, 30, 30)."""
        out = np.empty((count, 30, 30), dtype=np.float64)
        _raise(library().sf_model_element_stiffness(ctx._ptr, self._ptr, first, count, _ptr(out, C.c_double)))
        return out

    def assemble_stiffness(self, ctx: Context, cancel: Optional[Callable[[], bool]] = None) -> "CsrMatrix":
        """assemble_stiffness (assembly.rs:17-27) on the device: the full stiffness with no node
        eliminated, three unknowns per node in node order, as a device CsrMatrix."""
        h = C.c_void_p()
        _raise(library().sf_assemble_stiffness(ctx._ptr, self._ptr, _cancel(cancel), None, C.byref(h)))
        m = CsrMatrix(h)
        n = C.c_uint64()
        _raise(library().sf_csr_read(h, C.byref(n), None, None, None, None))
        m._ctx, m.size = ctx, n.value
        return m

    def points_per_element(self) -> int:
        return library().sf_model_points_per_element(self._ptr)

    def begin_recovery(self, ctx: "Context", displacements, converged: bool = False) -> "Recovery":
        """A resumable recovery of `displacements` (3 per mesh node) over this model, none of it done
        (Recovery::new, results.rs:440-467)."""
        u = self._displacements(displacements)
        h = C.c_void_p()
        _raise(library().sf_recovery_create(ctx._ptr, self._ptr, _ptr(u, C.c_double), 1 if converged else 0, C.byref(h)))
        return Recovery(h)

    def begin_error_estimate(self, recovery: "Recovery") -> "ErrorEstimation":
        """An error estimate of `recovery` over this model's mesh and material, none of it done."""
        h = C.c_void_p()
        _raise(library().sf_error_estimation_create(self._ptr, recovery._ptr, C.byref(h)))
        e = ErrorEstimation(h)
        e._recovery = recovery
        return e

    def nodal_rotations(self, ctx: "Context", displacements) -> np.ndarray:
        """nodal_rotations (results.rs:262-305): the local rotation, half the curl of the
        displacement, in radians at every mesh node, (nodes, 3)."""
        u = self._displacements(displacements)
        out = np.empty((u.size // 3, 3))
        _raise(library().sf_nodal_rotations(ctx._ptr, self._ptr, _ptr(u, C.c_double), _ptr(out, C.c_double)))
        return out

    def _displacements(self, displacements) -> np.ndarray:
        u = np.ascontiguousarray(displacements, dtype=np.float64).reshape(-1)
        if u.size != 3 * self.mesh().nodes().shape[0]:
            # results.rs:268-273 (ModelError::DisplacementFieldLength)
            raise ValueError("displacements must hold 3 values per mesh node")
        return u

    def recover(self, ctx: "Context", displacements):
        """Stress recovery on the device (results.rs:307-576): (point_values (elements, points, 10):
        position, stress [xx, yy, zz, xy, yz, zx], von Mises; nodal_values (nodes, 7): recovered
        stress and von Mises)."""
        u = np.ascontiguousarray(displacements, dtype=np.float64).reshape(-1)
        nodes = self.mesh().nodes().shape[0]
        if u.size != 3 * nodes:
            raise ValueError("displacements must hold 3 values per mesh node")
        q = self.points_per_element()
        elements = self.mesh().tets().shape[0]
        points = np.empty((elements, q, 10))
        nodal = np.empty((nodes, 7))
        _raise(library().sf_recover_stress(ctx._ptr, self._ptr, _ptr(u, C.c_double), _ptr(points, C.c_double),
                                           _ptr(nodal, C.c_double)))
        return points, nodal

    def load_vector(self) -> np.ndarray:
        """The right-hand side over the free unknowns of this model's loads (study.rs:444-449,
        :696-735), without building the matrix pattern."""
        lib = library()
        nf = C.c_uint64()
        _raise(lib.sf_model_free_problem(self._ptr, C.byref(nf), None, None, None, None, None, None))
        rhs = np.empty(3 * nf.value)
        _raise(lib.sf_model_free_problem(self._ptr, None, None, None, _ptr(rhs, C.c_double), None, None, None))
        return rhs

    def free_problem(self):
        """(free_map, rhs, block_starts, block_columns, rigid_modes) on the host."""
        lib = library()
        nf, nb = C.c_uint64(), C.c_uint64()
        _raise(lib.sf_model_free_problem(self._ptr, C.byref(nf), C.byref(nb), None, None, None, None, None))
        nn = self.mesh().nodes().shape[0]
        free_map = np.empty(nn, dtype=np.uint32)
        rhs = np.empty(3 * nf.value)
        starts = np.empty(nf.value + 1, dtype=np.int64)
        cols = np.empty(nb.value, dtype=np.uint32)
        modes = np.empty(18 * nf.value)
        _raise(lib.sf_model_free_problem(self._ptr, None, None, _ptr(free_map, C.c_uint32), _ptr(rhs, C.c_double),
                                         _ptr(starts, C.c_int64), _ptr(cols, C.c_uint32), _ptr(modes, C.c_double)))
        return free_map, rhs, starts, cols, modes


class StepOutcome(enum.IntEnum):
    """pcg.rs:47-57"""
    RUNNING = 0
    CONVERGED = 1
    CANCELLED = 2


@dataclass
class Progress:
    iteration: int
    relative_residual: float
    at_rounding_floor: bool
    outcome: StepOutcome

    @classmethod
    def of(cls, p):
        return cls(p.iteration, p.relative_residual, bool(p.at_rounding_floor), StepOutcome(p.outcome))


class StaticStudy(_Handle):
    """crates/solidfem-elastic/src/study.rs:178-372"""
    _destroy = "sf_static_study_destroy"

    @classmethod
    def new(cls, ctx: Context, model: Model, pcg: PcgOptions = None, multigrid: Optional[MultigridOptions] = None,
            cancel=None):
        pcg = pcg or PcgOptions()
        h = C.c_void_p()
        po = pcg.c()
        mo = multigrid.c() if multigrid else None
        cb = _cancel(cancel)
        _raise(library().sf_static_study_create(ctx._ptr, model._ptr, C.byref(po), 1 if multigrid else 0,
                                                C.byref(mo) if mo else None, cb, None, C.byref(h)))
        s = cls(h)
        s._ctx = ctx
        return s

    def step(self, max_iterations: int, cancel=None) -> Progress:
        p = _capi.Progress()
        cb = _cancel(cancel)
        _raise(library().sf_static_study_step(self._ptr, max_iterations, cb, None, C.byref(p)))
        return Progress.of(p)

    def solve(self, slice_iterations: int = 2**62) -> Progress:
        while True:
            p = self.step(slice_iterations)
            if p.outcome != StepOutcome.RUNNING:
                return p

    def displacements(self) -> np.ndarray:
        n = self.report().total_dofs // 3
        out = np.empty((n, 3))
        _raise(library().sf_static_study_displacements(self._ptr, _ptr(out, C.c_double)))
        return out

    def report(self) -> _capi.Report:
        r = _capi.Report()
        _raise(library().sf_static_study_report(self._ptr, C.byref(r)))
        return r

    def matrix(self):
        """The assembled free stiffness as scalar CSR (row_starts, columns, values)."""
        lib = library()
        n, nnz = C.c_uint64(), C.c_uint64()
        _raise(lib.sf_static_study_matrix(self._ptr, C.byref(n), C.byref(nnz), None, None, None))
        rs = np.empty(n.value + 1, dtype=np.int64)
        cols = np.empty(nnz.value, dtype=np.uint32)
        vals = np.empty(nnz.value)
        _raise(lib.sf_static_study_matrix(self._ptr, None, None, _ptr(rs, C.c_int64), _ptr(cols, C.c_uint32),
                                          _ptr(vals, C.c_double)))
        return rs, cols, vals

    def recover(self):
        """Stress recovery of the current displacements on the device (results.rs:307-576):
        (point_values (elements, points, 10), nodal_values (nodes, 7)), as Model.recover."""
        nodes, elements, points_per = C.c_uint64(), C.c_uint64(), C.c_uint64()
        _raise(library().sf_static_study_counts(self._ptr, C.byref(nodes), C.byref(elements), C.byref(points_per)))
        points = np.empty((elements.value, points_per.value, 10))
        nodal = np.empty((nodes.value, 7))
        _raise(library().sf_static_study_recover(self._ptr, _ptr(points, C.c_double), _ptr(nodal, C.c_double)))
        return points, nodal

    def begin_recovery(self) -> "Recovery":
        """StaticStudy::begin_recovery (study.rs:263-273): a resumable recovery of the current
        displacements, the converged answer or a preview of a solve in progress."""
        h = C.c_void_p()
        _raise(library().sf_static_study_begin_recovery(self._ptr, C.byref(h)))
        return Recovery(h)

    def continue_recovery(self, recovery: "Recovery", max_work: int, cancel=None) -> bool:
        """StaticStudy::continue_recovery (study.rs:275-292): up to max_work more units; True once
        all are done. The time joins the report's recovery_seconds."""
        done = C.c_int32()
        _raise(library().sf_static_study_continue_recovery(self._ptr, recovery._ptr, int(max_work), _cancel(cancel), None,
                                                           C.byref(done)))
        return bool(done.value)

    def finish_recovery(self, recovery: "Recovery"):
        """StaticStudy::finish_recovery (study.rs:294-300): the finished field; refuses one with
        units left (ModelError NOT_FINISHED)."""
        return recovery.field()

    def begin_error_estimate(self, recovery: "Recovery") -> "ErrorEstimation":
        """StaticStudy::begin_error_estimate (study.rs:329-333)."""
        h = C.c_void_p()
        _raise(library().sf_static_study_begin_error_estimate(self._ptr, recovery._ptr, C.byref(h)))
        e = ErrorEstimation(h)
        e._recovery = recovery
        return e

    def continue_error_estimate(self, estimation: "ErrorEstimation", max_elements: int, cancel=None) -> bool:
        """StaticStudy::continue_error_estimate (study.rs:335-352): True once every element is done;
        the time joins the report's recovery_seconds."""
        done = C.c_int32()
        _raise(library().sf_static_study_continue_error_estimate(self._ptr, estimation._ptr, int(max_elements),
                                                                 _cancel(cancel), None, C.byref(done)))
        return bool(done.value)

    def finish_error_estimate(self, estimation: "ErrorEstimation") -> ErrorEstimate:
        """StaticStudy::finish_error_estimate (study.rs:354-360)."""
        return estimation.estimate()

    def solution_with_error_estimate(self, cancel=None):
        """StaticStudy::solution_with_error_estimate (study.rs:304-327): refuses until the solve has
        converged (a preview's estimate is refused NOT_FINISHED); ((point_values, nodal_values),
        ErrorEstimate)."""
        recovery = self.begin_recovery()
        self.continue_recovery(recovery, 2 ** 63 - 1, cancel)
        estimation = self.begin_error_estimate(recovery)
        self.continue_error_estimate(estimation, 2 ** 63 - 1, cancel)
        estimate = self.finish_error_estimate(estimation)
        return self.finish_recovery(estimation.into_recovery()), estimate

    def rhs(self) -> np.ndarray:
        out = np.empty(self.report().free_dofs)
        _raise(library().sf_static_study_rhs(self._ptr, _ptr(out, C.c_double)))
        return out

    def system(self) -> "CsrMatrix":
        """The study's system matrix on the device (owned by the study)."""
        m = CsrMatrix(library().sf_static_study_system(self._ptr))
        m._owned, m._parent, m._ctx, m.size = False, self, self._ctx, self.report().free_dofs
        return m

    def multigrid(self) -> Optional["Multigrid"]:
        h = library().sf_static_study_multigrid(self._ptr)
        if not h:
            return None
        m = Multigrid(h, owned=False)
        m._parent = self
        return m


class StudyBuilder(_Handle):
    """crates/solidfem-elastic/src/study.rs:389-656"""
    _destroy = "sf_study_builder_destroy"

    @classmethod
    def new(cls, ctx: Context, model: Model, pcg: PcgOptions = None, multigrid: Optional[MultigridOptions] = None):
        pcg = pcg or PcgOptions()
        h = C.c_void_p()
        po = pcg.c()
        mo = multigrid.c() if multigrid else None
        _raise(library().sf_study_builder_create(ctx._ptr, model._ptr, C.byref(po), 1 if multigrid else 0,
                                                 C.byref(mo) if mo else None, C.byref(h)))
        b = cls(h)
        b._ctx = ctx
        return b

    def step(self, max_elements: int, cancel=None):
        p = _capi.BuildProgress()
        cb = _cancel(cancel)
        _raise(library().sf_study_builder_step(self._ptr, max_elements, cb, None, C.byref(p)))
        return p

    def finish(self) -> StaticStudy:
        h = C.c_void_p()
        _raise(library().sf_study_builder_finish(self._ptr, C.byref(h)))
        self._ptr = None  # consumed
        s = StaticStudy(h)
        s._ctx = self._ctx
        return s


class CsrMatrix(_Handle):
    """crates/solidfem-solve/src/csr.rs:6-168 (on the device when ctx is given)."""
    _destroy = "sf_csr_destroy"
    _owned = True

    def close(self):
        if self._owned:
            super().close()

    @classmethod
    def new(cls, ctx: Optional[Context], row_starts, columns, values):
        rs = np.ascontiguousarray(row_starts, dtype=np.int64)
        cols = np.ascontiguousarray(columns, dtype=np.uint32)
        vals = np.ascontiguousarray(values, dtype=np.float64)
        h = C.c_void_p()
        _raise(library().sf_csr_create(ctx._ptr if ctx else None, len(rs) - 1, _ptr(rs, C.c_int64),
                                       _ptr(cols, C.c_uint32), _ptr(vals, C.c_double), C.byref(h)))
        m = cls(h)
        m._ctx, m.size = ctx, len(rs) - 1
        return m

    def read(self):
        """The matrix as scalar CSR (row_starts, columns, values), read back from the device."""
        lib = library()
        n, nnz = C.c_uint64(), C.c_uint64()
        _raise(lib.sf_csr_read(self._ptr, C.byref(n), C.byref(nnz), None, None, None))
        rs = np.empty(n.value + 1, dtype=np.int64)
        cols = np.empty(nnz.value, dtype=np.uint32)
        vals = np.empty(nnz.value)
        _raise(lib.sf_csr_read(self._ptr, None, None, _ptr(rs, C.c_int64), _ptr(cols, C.c_uint32),
                               _ptr(vals, C.c_double)))
        return rs, cols, vals

    def nonzeros(self) -> int:
        """CsrMatrix::nonzeros (csr.rs:79-82)."""
        nnz = C.c_uint64()
        _raise(library().sf_csr_read(self._ptr, None, C.byref(nnz), None, None, None))
        return nnz.value

    def row(self, row: int):
        """CsrMatrix::row (csr.rs:157-161): (columns, values) of a scalar row."""
        rs, cols, vals = self.read()
        return cols[rs[row]:rs[row + 1]], vals[rs[row]:rs[row + 1]]

    def get(self, row: int, column: int) -> float:
        """CsrMatrix::get (csr.rs:99-110): the stored value, 0 where there is none."""
        cols, vals = self.row(row)
        at = np.searchsorted(cols, column)
        return float(vals[at]) if at < len(cols) and cols[at] == column else 0.0

    def diagonal(self) -> np.ndarray:
        """CsrMatrix::diagonal (csr.rs:163-168)."""
        rs, cols, vals = self.read()
        out = np.zeros(len(rs) - 1)
        for i in range(len(rs) - 1):
            hit = np.nonzero(cols[rs[i]:rs[i + 1]] == i)[0]
            if hit.size:
                out[i] = vals[rs[i] + hit[0]]
        return out

    def multiply(self, x) -> np.ndarray:
        xx = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.size)
        _raise(library().sf_csr_multiply(self._ptr, _ptr(xx, C.c_double), _ptr(y, C.c_double)))
        return y

    def tuned_depth(self) -> int:
        """The batch depth the tuning chose for this matrix's conjugate-gradient product."""
        d = C.c_int32()
        _raise(library().sf_csr_tuned_depth(self._ptr, C.byref(d)))
        return d.value

    def multiply_rounding_bound(self, x, dot_order=DotOrder.REFERENCE) -> float:
        xx = np.ascontiguousarray(x, dtype=np.float64)
        b = C.c_double()
        _raise(library().sf_csr_rounding_bound(self._ptr, _ptr(xx, C.c_double), int(dot_order), C.byref(b)))
        return b.value


class _Levels:
    """Shared level views of MultigridSetup (finished) and Multigrid."""
    _prefix = ""

    def _fn(self, name):
        return getattr(library(), self._prefix + name)

    def level_matrix(self, level: int, prolongator: bool = False):
        rows, bc, rb, cb, nb = (C.c_uint64() for _ in range(5))
        f = self._fn("level_matrix")
        _raise(f(self._ptr, level, int(prolongator), C.byref(rows), C.byref(bc), C.byref(rb), C.byref(cb), C.byref(nb),
                 None, None, None))
        starts = np.empty(rows.value + 1, dtype=np.int64)
        cols = np.empty(nb.value, dtype=np.uint32)
        vals = np.empty(nb.value * rb.value * cb.value)
        _raise(f(self._ptr, level, int(prolongator), None, None, None, None, None, _ptr(starts, C.c_int64),
                 _ptr(cols, C.c_uint32), _ptr(vals, C.c_double)))
        return dict(rows=rows.value, block_columns=bc.value, row_block=rb.value, col_block=cb.value, starts=starts,
                    columns=cols, values=vals)

    def level_inverse_diagonal(self, level: int) -> np.ndarray:
        m = self.level_matrix(level)
        out = np.empty(m["rows"] * m["row_block"] ** 2)
        _raise(self._fn("level_inverse_diagonal")(self._ptr, level, _ptr(out, C.c_double)))
        return out

    def smoother_spectrum(self, level: int):
        lo, hi = C.c_double(), C.c_double()
        _raise(self._fn("smoother_spectrum")(self._ptr, level, C.byref(lo), C.byref(hi)))
        return lo.value, hi.value

    def aggregates(self, level: int) -> np.ndarray:
        n = C.c_uint64()
        f = self._fn("aggregates")
        _raise(f(self._ptr, level, None, C.byref(n)))
        out = np.empty(n.value, dtype=np.uint32)
        _raise(f(self._ptr, level, _ptr(out, C.c_uint32), None))
        return out


class MultigridSetup(_Handle, _Levels):
    """crates/solidfem-solve/src/multigrid/setup.rs:81-240"""
    _destroy = "sf_multigrid_setup_destroy"
    _prefix = "sf_multigrid_setup_"

    @classmethod
    def new(cls, ctx: Optional[Context], matrix: CsrMatrix, near_null, positions, options: MultigridOptions = None):
        options = options or MultigridOptions()
        nn = np.ascontiguousarray(near_null, dtype=np.float64).ravel()
        pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
        o = options.c()
        h = C.c_void_p()
        _raise(library().sf_multigrid_setup_create(ctx._ptr if ctx else None, matrix._ptr, _ptr(nn, C.c_double), len(nn),
                                                   _ptr(pos, C.c_double), len(pos), C.byref(o), C.byref(h)))
        s = cls(h)
        s._ctx = ctx
        return s

    def step(self, budget: int, cancel=None) -> bool:
        done = C.c_int32()
        cb = _cancel(cancel)
        _raise(library().sf_multigrid_setup_step(self._ptr, budget, cb, None, C.byref(done)))
        return bool(done.value)

    def level_count(self) -> int:
        n = C.c_uint64()
        _raise(library().sf_multigrid_setup_level_count(self._ptr, C.byref(n)))
        return n.value

    def coarse_factor(self) -> Optional[np.ndarray]:
        n = C.c_uint64()
        _raise(library().sf_multigrid_setup_coarse_factor(self._ptr, C.byref(n), None))
        if n.value == 0:
            return None
        out = np.empty((n.value, n.value))
        _raise(library().sf_multigrid_setup_coarse_factor(self._ptr, C.byref(n), _ptr(out, C.c_double)))
        return out

    def finish(self) -> "Multigrid":
        h = C.c_void_p()
        _raise(library().sf_multigrid_setup_finish(self._ptr, C.byref(h)))
        m = Multigrid(h)
        m._ctx = self._ctx
        return m


class Multigrid(_Handle, _Levels):
    """crates/solidfem-solve/src/multigrid/cycle.rs:253-458"""
    _destroy = "sf_multigrid_destroy"
    _prefix = "sf_multigrid_"

    def __init__(self, ptr, owned=True):
        super().__init__(ptr)
        self._owned = owned

    def close(self):
        if getattr(self, "_owned", True):
            super().close()

    def tuned_depth(self, level: int, kind: int = 0) -> int:
        """The batch depth of `level`'s fused Chebyshev step (kind 0) or residual (kind 1)."""
        d = C.c_int32()
        _raise(library().sf_multigrid_tuned_depth(self._ptr, level, kind, C.byref(d)))
        return d.value

    def stats(self) -> _capi.MultigridStats:
        s = _capi.MultigridStats()
        _raise(library().sf_multigrid_stats_get(self._ptr, C.byref(s)))
        return s

    def precision(self) -> Precision:
        return Precision(library().sf_multigrid_precision(self._ptr))

    def threads(self) -> int:
        """Multigrid::threads (cycle.rs:310-316): one host thread enqueues the device's products."""
        return library().sf_multigrid_threads(self._ptr)

    def level_sizes(self):
        n = C.c_uint64()
        out = np.zeros(3 * 64, dtype=np.uint64)
        _raise(library().sf_multigrid_level_sizes(self._ptr, _ptr(out, C.c_uint64), 64, C.byref(n)))
        return [tuple(int(v) for v in out[3 * i:3 * i + 3]) for i in range(n.value)]

    def with_precision(self, precision: Precision, coarse_solve: int = -1) -> "Multigrid":
        """The same hierarchy in another precision (cycle.rs:265-308 over the same levels); the coarse
        solve mode is kept unless given."""
        h = C.c_void_p()
        _raise(library().sf_multigrid_with_precision(self._ptr, int(precision), coarse_solve, C.byref(h)))
        m = Multigrid(h)
        m._parent = self
        return m

    def cycle_bytes(self) -> float:
        return library().sf_multigrid_cycle_bytes(self._ptr)

    def apply(self, residual) -> np.ndarray:
        r = np.ascontiguousarray(residual, dtype=np.float64)
        z = np.empty_like(r)
        _raise(library().sf_multigrid_apply(self._ptr, _ptr(r, C.c_double), _ptr(z, C.c_double)))
        return z

    def set_graph(self, enable: bool = True):
        """Replay each V-cycle as one captured HIP graph (the same bits); see sf_multigrid_set_graph."""
        _raise(library().sf_multigrid_set_graph(self._ptr, 1 if enable else 0))

    def apply_batch(self, residuals) -> np.ndarray:
        """apply() of each row of `residuals` (columns x n), all through one batched V-cycle."""
        r = np.ascontiguousarray(residuals, dtype=np.float64)
        if r.ndim != 2:
            raise ValueError("residuals must be (columns, n)")
        z = np.empty_like(r)
        _raise(library().sf_multigrid_apply_batch(self._ptr, _ptr(r, C.c_double), r.shape[0], _ptr(z, C.c_double)))
        return z

    def smooth(self, level: int, b) -> np.ndarray:
        bb = np.ascontiguousarray(b, dtype=np.float64)
        x = np.empty_like(bb)
        _raise(library().sf_multigrid_smooth(self._ptr, level, _ptr(bb, C.c_double), _ptr(x, C.c_double)))
        return x

    def solve_coarsest(self, values) -> np.ndarray:
        v = np.array(values, dtype=np.float64)
        _raise(library().sf_multigrid_solve_coarsest(self._ptr, _ptr(v, C.c_double), len(v)))
        return v


class Pcg(_Handle):
    """crates/solidfem-solve/src/pcg.rs:113-283, preconditioned by a Multigrid or by Jacobi."""
    _destroy = "sf_pcg_destroy"

    @classmethod
    def new(cls, ctx: Context, matrix: CsrMatrix, preconditioner: Optional[Multigrid], rhs,
            options: PcgOptions = None):
        options = options or PcgOptions()
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        o = options.c()
        h = C.c_void_p()
        _raise(library().sf_pcg_create(ctx._ptr, matrix._ptr, preconditioner._ptr if preconditioner else None,
                                       _ptr(b, C.c_double), C.byref(o), C.byref(h)))
        p = cls(h)
        p._keep = (ctx, matrix, preconditioner)
        return p

    @classmethod
    def with_preconditioner(cls, ctx: Context, matrix: CsrMatrix, apply, rhs, options: PcgOptions = None,
                            device: bool = False):
        """Pcg::new with a caller's Preconditioner (preconditioner.rs:9-20). Host form: apply(r, z)
        gets two numpy arrays of n values and writes M^-1 r into z. Device form (device=True):
        apply(r_ptr, z_ptr, n, stream) gets device addresses and the solve's stream, and enqueues its
        work there. An exception in apply fails the step (PcgError) and is raised after it."""
        options = options or PcgOptions()
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        o = options.c()
        failures = []

        def trampoline(r, z, n, stream, _user):
            try:
                if device:
                    apply(C.cast(r, C.c_void_p).value, C.cast(z, C.c_void_p).value, int(n), stream)
                else:
                    apply(np.ctypeslib.as_array(r, shape=(int(n),)), np.ctypeslib.as_array(z, shape=(int(n),)))
                return 0
            except BaseException as e:  # the callback cannot raise through C
                failures.append(e)
                return 1

        fn = _capi.APPLY_FN(trampoline)
        pre = _capi.Preconditioner(fn, None, 1 if device else 0, 0)
        h = C.c_void_p()
        status = library().sf_pcg_create_with(ctx._ptr, matrix._ptr, C.byref(pre), _ptr(b, C.c_double), C.byref(o),
                                              C.byref(h))
        if failures:
            raise failures[0]
        _raise(status)
        p = cls(h)
        p._keep = (ctx, matrix, fn, pre)
        p._failures = failures
        return p

    def step(self, max_iterations: int, cancel=None) -> Progress:
        p = _capi.Progress()
        cb = _cancel(cancel)
        status = library().sf_pcg_step(self._ptr, max_iterations, cb, None, C.byref(p))
        failures = getattr(self, "_failures", None)
        if failures:
            raise failures.pop(0)
        _raise(status)
        return Progress.of(p)

    def restart(self, rhs=None):
        """Pcg::new again on the resident buffers (optionally a new right-hand side)."""
        if rhs is None:
            _raise(library().sf_pcg_restart(self._ptr, None))
        else:
            b = np.ascontiguousarray(rhs, dtype=np.float64)
            _raise(library().sf_pcg_restart(self._ptr, _ptr(b, C.c_double)))

    def estimate(self) -> np.ndarray:
        n = library().sf_pcg_size(self._ptr)
        out = np.empty(n)
        _raise(library().sf_pcg_estimate(self._ptr, _ptr(out, C.c_double)))
        return out

    def progress(self) -> Progress:
        p = _capi.Progress()
        _raise(library().sf_pcg_progress(self._ptr, C.byref(p)))
        return Progress.of(p)

    def is_converged(self) -> bool:
        return bool(library().sf_pcg_is_converged(self._ptr))


@dataclass
class ColumnProgress:
    """One column of a batched solve: pcg.rs:35-45 plus its outcome; `error` is the PcgError
    status (10 or 11) of a failed column, with its iteration and value."""
    iteration: int
    relative_residual: float
    at_rounding_floor: bool
    outcome: int
    error: int = 0
    error_iteration: int = 0
    error_value: float = 0.0

    @classmethod
    def of(cls, p):
        return cls(p.iteration, p.relative_residual, bool(p.at_rounding_floor), p.outcome, p.error, p.error_iteration,
                   p.error_value)


class PcgBatch(_Handle):
    """Up to 64 load cases over one matrix and one preconditioner, each column the recurrence of
    pcg.rs:113-283 on its own scalars, in lockstep: column m's iterates are bit for bit those of
    Pcg.new with that right-hand side alone."""
    _destroy = "sf_pcg_batch_destroy"
    FAILED = 3

    @classmethod
    def new(cls, ctx: Context, matrix: CsrMatrix, preconditioner: Optional[Multigrid], rhs, options: PcgOptions = None):
        options = options or PcgOptions()
        b = np.ascontiguousarray(rhs, dtype=np.float64)
        if b.ndim != 2:
            raise ValueError("rhs must be (columns, n)")
        o = options.c()
        h = C.c_void_p()
        _raise(library().sf_pcg_batch_create(ctx._ptr, matrix._ptr, preconditioner._ptr if preconditioner else None,
                                             _ptr(b, C.c_double), b.shape[0], C.byref(o), C.byref(h)))
        p = cls(h)
        p._keep = (ctx, matrix, preconditioner)
        p.columns, p.n = b.shape
        return p

    def step(self, max_iterations: int, cancel=None):
        """Runs up to `max_iterations` lockstep iterations; returns (running columns, [ColumnProgress])."""
        progress = (_capi.ColumnProgress * self.columns)()
        running = C.c_uint64()
        _raise(library().sf_pcg_batch_step(self._ptr, max_iterations, _cancel(cancel), None, progress, C.byref(running)))
        return running.value, [ColumnProgress.of(p) for p in progress]

    def progress(self):
        progress = (_capi.ColumnProgress * self.columns)()
        _raise(library().sf_pcg_batch_progress(self._ptr, progress))
        return [ColumnProgress.of(p) for p in progress]

    def restart(self, rhs=None):
        if rhs is None:
            _raise(library().sf_pcg_batch_restart(self._ptr, None))
        else:
            b = np.ascontiguousarray(rhs, dtype=np.float64)
            if b.shape != (self.columns, self.n):
                raise ValueError("rhs must keep the batch's shape")
            _raise(library().sf_pcg_batch_restart(self._ptr, _ptr(b, C.c_double)))

    def estimate(self) -> np.ndarray:
        out = np.empty((self.columns, self.n))
        _raise(library().sf_pcg_batch_estimate(self._ptr, _ptr(out, C.c_double)))
        return out


class ServeState(enum.IntEnum):
    QUEUED = 0
    BUILDING = 1
    SOLVING = 2
    DONE = 3
    FAILED = 4


@dataclass
class ServerOptions:
    """Serving options (include/solidfem_hip.h: sf_server_options)."""
    hbm_budget: int = 200 << 30
    precision: Precision = Precision.SINGLE
    coarse_solve: int = 1
    tolerance: float = 1e-8
    max_columns: int = 32
    slice_iterations: int = 8
    builders: int = 4
    solvers: int = 8
    dot_order: DotOrder = DotOrder.TREE

    def c(self):
        return _capi.ServerOptions(self.hbm_budget, int(self.precision), self.coarse_solve, self.tolerance,
                                   self.max_columns, self.slice_iterations, self.builders, self.solvers,
                                   int(self.dot_order), 0)


@dataclass
class ServeStatus:
    state: ServeState
    error: int
    cache_hit: bool
    at_rounding_floor: bool
    iterations: int
    relative_residual: float
    queued_seconds: float
    build_seconds: float
    solve_seconds: float
    latency_seconds: float
    nodes: int
    free_dofs: int

    @classmethod
    def of(cls, s):
        return cls(ServeState(s.state), s.error, bool(s.cache_hit), bool(s.at_rounding_floor), s.iterations,
                   s.relative_residual, s.queued_seconds, s.build_seconds, s.solve_seconds, s.latency_seconds,
                   s.nodes, s.free_dofs)


class Server(_Handle):
    """A solve server on one GPU: submit box problems from any thread, collect displacements.
    Requests sharing mesh, constraints and material share one resident system and are solved as
    the columns of one batch; each answer is bit for bit the request's answer alone."""
    _destroy = "sf_server_destroy"

    @classmethod
    def new(cls, device: int = 0, options: ServerOptions = None):
        o = (options or ServerOptions()).c()
        h = C.c_void_p()
        _raise(library().sf_server_create(device, C.byref(o), C.byref(h)))
        return cls(h)

    def submit(self, workload: "Workload") -> int:
        w = workload.c(_capi.Workload)
        rid = C.c_uint64()
        _raise(library().sf_server_submit(self._ptr, C.byref(w), C.byref(rid)))
        return rid.value

    def submit_model(self, model: "Model") -> int:
        """A model of any mesh (copied); models with the same mesh, constraints and material share
        one system."""
        rid = C.c_uint64()
        _raise(library().sf_server_submit_model(self._ptr, model._ptr, C.byref(rid)))
        return rid.value

    def status(self, rid: int) -> ServeStatus:
        s = _capi.ServeStatus()
        _raise(library().sf_server_status(self._ptr, rid, C.byref(s)))
        return ServeStatus.of(s)

    def wait(self, rid: int, timeout: float = -1.0) -> ServeStatus:
        s = _capi.ServeStatus()
        _raise(library().sf_server_wait(self._ptr, rid, timeout, C.byref(s)))
        return ServeStatus.of(s)

    def result(self, rid: int) -> np.ndarray:
        """Displacements, (nodes, 3); raises the request's error if it failed."""
        s = self.status(rid)
        out = np.empty((s.nodes, 3))
        _raise(library().sf_server_result(self._ptr, rid, _ptr(out, C.c_double), out.size))
        return out

    def release(self, rid: int):
        _raise(library().sf_server_release(self._ptr, rid))

    def stats(self) -> dict:
        s = _capi.ServerStats()
        _raise(library().sf_server_stats_get(self._ptr, C.byref(s)))
        return {name: getattr(s, name) for name, _ in s._fields_}


class Kernel(enum.IntEnum):
    FINE_CHEBYSHEV = 0
    SYSTEM_PRODUCT = 1
    FINE_RESIDUAL = 2
    BATCH_FINE_CHEBYSHEV = 3
    BATCH_SYSTEM_PRODUCT = 4


def force_variant(code: int = 0):
    """Every later tuning that has `code` among its candidates takes it (0: time them)."""
    _raise(library().sf_force_variant(int(code)))


def setup_slot_limit(blocks: int = 0):
    """Device setup rows of more than `blocks` output blocks take the host path (0: no extra limit)."""
    _raise(library().sf_setup_slot_limit(int(blocks)))


def profile_enable(kernel: Kernel, enable: bool = True):
    _raise(library().sf_profile_enable(int(kernel), 1 if enable else 0))


def profile_reset():
    _raise(library().sf_profile_reset())


def profile_read(kernel: Kernel):
    """(launches, summed seconds, summed algorithmic bytes) since the last reset."""
    n, t, b = C.c_uint64(), C.c_double(), C.c_double()
    _raise(library().sf_profile_read(int(kernel), C.byref(n), C.byref(t), C.byref(b)))
    return n.value, t.value, b.value


class Recovery(_Handle):
    """A resumable stress recovery on the device (Recovery, results.rs:103-124 and :440-536): every
    element in order, then one patch fit per mesh vertex in order, in slices of any size; the
    finished field has the bits of a recovery in one call."""
    _destroy = "sf_recovery_destroy"

    def advance(self, max_work: int, cancel=None) -> bool:
        """Up to max_work more units (an element, then a patch fit), `cancel` read before each;
        True once all are done."""
        done = C.c_int32()
        _raise(library().sf_recovery_advance(self._ptr, int(max_work), _cancel(cancel), None, C.byref(done)))
        return bool(done.value)

    def state(self) -> _capi.RecoveryState:
        st = _capi.RecoveryState()
        _raise(library().sf_recovery_state_of(self._ptr, C.byref(st)))
        return st

    def elements_done(self) -> int:
        return self.state().elements_done

    def is_finished(self) -> bool:
        return bool(self.state().finished)

    def is_converged(self) -> bool:
        return bool(self.state().converged)

    def displacements(self) -> np.ndarray:
        out = np.empty((self.state().nodes, 3))
        _raise(library().sf_recovery_displacements(self._ptr, _ptr(out, C.c_double)))
        return out

    def field(self):
        """(point_values (elements, points, 10), nodal_values (nodes, 7)), as Model.recover; refuses
        while units are left (ModelError NOT_FINISHED)."""
        st = self.state()
        points = np.empty((st.elements, st.points_per_element, 10))
        nodal = np.empty((st.nodes, 7))
        _raise(library().sf_recovery_field(self._ptr, _ptr(points, C.c_double), _ptr(nodal, C.c_double)))
        return points, nodal


@dataclass
class ErrorEstimate:
    """ErrorEstimate (error_estimate.rs:62-78)."""
    element_errors: np.ndarray
    error_norm_squared: float
    solution_norm_squared: float
    relative_error: float


class ErrorEstimation(_Handle):
    """A resumable error estimate of a finished recovery, on the device (ErrorEstimation,
    error_estimate.rs:80-238): one element per unit in element order; any slicing gives one run's
    bits. It borrows the recovery, which it keeps alive."""
    _destroy = "sf_error_estimation_destroy"

    def advance(self, max_elements: int, cancel=None) -> bool:
        done = C.c_int32()
        _raise(library().sf_error_estimation_advance(self._ptr, int(max_elements), _cancel(cancel), None, C.byref(done)))
        return bool(done.value)

    def elements_done(self) -> int:
        return library().sf_error_estimation_elements_done(self._ptr)

    def estimate(self) -> ErrorEstimate:
        """ErrorEstimation::estimate; refuses one with elements left (ModelError NOT_FINISHED)."""
        out = _capi.ErrorEstimate()
        _raise(library().sf_error_estimation_finish(self._ptr, C.byref(out), None))
        errors = np.empty(out.elements)
        _raise(library().sf_error_estimation_finish(self._ptr, None, _ptr(errors, C.c_double)))
        return ErrorEstimate(errors, out.error_norm_squared, out.solution_norm_squared, out.relative_error)

    def into_recovery(self) -> "Recovery":
        return self._recovery


@dataclass
class PeakStress:
    """PeakStress (results.rs:141-147), with what holds it (PeakHolder, results.rs:178-182)."""
    von_mises: float
    position: np.ndarray
    at_point: bool
    index: int


def peak_stress(positions, nodal_von_mises, point_values, inside: Optional[Callable[[np.ndarray], bool]] = None):
    """peak_stress (results.rs:164-176): the peak von Mises over the mesh nodes (positions (nodes, 3),
    nodal_von_mises (nodes,)), then the integration points (point_values (..., 10) as recovered),
    whose positions `inside` accepts (None: everywhere); None when none is inside."""
    x = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    vm = np.ascontiguousarray(nodal_von_mises, dtype=np.float64).reshape(-1)
    pts = np.ascontiguousarray(point_values, dtype=np.float64).reshape(-1, 10)
    if vm.size != x.shape[0]:
        raise ValueError("one nodal von Mises per node")
    failure = []

    def region(p, _user):
        try:
            return 1 if inside(np.array(p[:3])) else 0
        except BaseException as e:  # the callback cannot raise through C
            failure.append(e)
            return 0

    fn = _capi.REGION_FN(region) if inside is not None else _capi.REGION_FN()
    out = _capi.Peak()
    _raise(library().sf_peak_stress(_ptr(x, C.c_double), x.shape[0], _ptr(vm, C.c_double), 1, _ptr(pts, C.c_double),
                                    pts.shape[0], fn, None, C.byref(out)))
    if failure:
        raise failure[0]
    if not out.found:
        return None
    return PeakStress(out.von_mises, np.array(out.position[:]), bool(out.at_point), out.index)


def _doubles(values, count):
    a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
    if a.size != count:
        raise ValueError(f"{count} values expected")
    return a


def shape_values(natural) -> np.ndarray:
    """shape_values (shape.rs:35-45): the 10 Tet10 shape functions at natural (r, s, t)."""
    out = np.empty(10)
    library().sf_shape_values(_ptr(_doubles(natural, 3), C.c_double), _ptr(out, C.c_double))
    return out


def shape_derivatives(natural) -> np.ndarray:
    """shape_derivatives (shape.rs:47-67): (10, 3)."""
    out = np.empty((10, 3))
    library().sf_shape_derivatives(_ptr(_doubles(natural, 3), C.c_double), _ptr(out, C.c_double))
    return out


def quadrature_points(rule: int) -> np.ndarray:
    """Quadrature::points (quadrature.rs:44-77) of the 4- or 14-point rule: (points, 4) rows of
    r, s, t and the weight."""
    n = library().sf_quadrature_points(rule, None)
    if n == 0:
        raise ValueError("the rules have 4 or 14 points")
    out = np.empty((n, 4))
    library().sf_quadrature_points(rule, _ptr(out, C.c_double))
    return out


@dataclass
class PointGeometry:
    """PointGeometry (element.rs:14-23)."""
    gradients: np.ndarray  # (10, 3)
    det_j: float
    position: np.ndarray
    _raw: object = None


def point_geometry(nodes, natural) -> PointGeometry:
    """point_geometry (element.rs:25-80) of a Tet10 (nodes (10, 3)); ModelError 45 when inverted."""
    g = _capi.Geometry()
    _raise(library().sf_point_geometry(_ptr(_doubles(nodes, 30), C.c_double), _ptr(_doubles(natural, 3), C.c_double),
                                       C.byref(g)))
    return PointGeometry(np.array(g.gradients[:]).reshape(10, 3), g.det_j, np.array(g.position[:]), g)


def strain_at(geometry: PointGeometry, displacements) -> np.ndarray:
    """strain_at (element.rs:160-171): [xx, yy, zz, 2xy, 2yz, 2zx]."""
    out = np.empty(6)
    library().sf_strain_at(C.byref(geometry._raw), _ptr(_doubles(displacements, 30), C.c_double), _ptr(out, C.c_double))
    return out


def stress_at(youngs_modulus: float, poisson_ratio: float, strain) -> np.ndarray:
    """stress_at (element.rs:174-186): [xx, yy, zz, xy, yz, zx]."""
    out = np.empty(6)
    library().sf_stress_at(youngs_modulus, poisson_ratio, _ptr(_doubles(strain, 6), C.c_double), _ptr(out, C.c_double))
    return out


def von_mises(stress) -> float:
    """von_mises (element.rs:189-192)."""
    return library().sf_von_mises(_ptr(_doubles(stress, 6), C.c_double))


class Verdict(enum.IntEnum):
    """Verdict (comparison.rs:68-80)."""
    SETTLED = 0
    STILL_MOVING = 1
    MESHES_TOO_ALIKE = 2


MIN_ELEMENT_RATIO = 4.0  # comparison.rs:53


@dataclass
class ComparedPeak:
    """ComparedPeak (comparison.rs:55-66)."""
    peak: PeakStress
    elements: int
    peak_element_volume: float


class MeshComparison:
    """MeshComparison (comparison.rs:83-223): how much the peak von Mises moved between a coarser and
    a finer solve of one part and load, over a region. A solve is (mesh, nodal_values (nodes, 7) or
    the nodal von Mises (nodes,), point_values (elements, points, 10)) as a recovery gives them."""

    def __init__(self, raw):
        self._raw = raw

    @classmethod
    def between(cls, coarse, fine, inside: Optional[Callable[[np.ndarray], bool]] = None) -> "MeshComparison":
        keep = []

        def solve(side):
            mesh, nodal, points = side
            nodal = np.ascontiguousarray(nodal, dtype=np.float64)
            pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 10)
            keep.extend([nodal, pts])
            if nodal.ndim == 1:
                vm, stride = nodal.ctypes.data, 1
            else:  # recovered nodal values: the von Mises is the last of 7 columns
                if nodal.shape[1] != 7:
                    raise ValueError("nodal values are 7 per node")
                vm, stride = nodal.ctypes.data + 6 * 8, 7
            return _capi.ComparedSolve(mesh._ptr, C.cast(vm, C.POINTER(C.c_double)), nodal.shape[0], stride,
                                       _ptr(pts, C.c_double), pts.shape[0])

        failure = []

        def region(p, _user):
            try:
                return 1 if inside(np.array(p[:3])) else 0
            except BaseException as e:  # the callback cannot raise through C
                failure.append(e)
                return 0

        fn = _capi.REGION_FN(region) if inside is not None else _capi.REGION_FN()
        c, f = solve(coarse), solve(fine)
        out = _capi.MeshComparison()
        status = library().sf_mesh_comparison_between(C.byref(c), C.byref(f), fn, None, C.byref(out))
        if failure:
            raise failure[0]
        _raise(status)
        return cls(out)

    def _peak(self, p) -> ComparedPeak:
        return ComparedPeak(PeakStress(p.peak.von_mises, np.array(p.peak.position[:]), bool(p.peak.at_point), p.peak.index),
                            p.elements, p.peak_element_volume)

    def coarse(self) -> ComparedPeak:
        return self._peak(self._raw.coarse)

    def fine(self) -> ComparedPeak:
        return self._peak(self._raw.fine)

    def peak_change(self) -> float:
        return self._raw.peak_change

    def peak_distance(self) -> float:
        return self._raw.peak_distance

    def element_ratio(self) -> float:
        return library().sf_mesh_comparison_element_ratio(C.byref(self._raw))

    def peak_element_ratio(self) -> float:
        return library().sf_mesh_comparison_peak_element_ratio(C.byref(self._raw))

    def safety_factors(self, yield_strength: float):
        a, b = C.c_double(), C.c_double()
        _raise(library().sf_mesh_comparison_safety_factors(C.byref(self._raw), yield_strength, C.byref(a), C.byref(b)))
        return a.value, b.value

    def safety_factor_change(self, yield_strength: float) -> float:
        out = C.c_double()
        _raise(library().sf_mesh_comparison_safety_factor_change(C.byref(self._raw), yield_strength, C.byref(out)))
        return out.value

    def verdict(self, tolerance: float) -> Verdict:
        return Verdict(library().sf_mesh_comparison_verdict(C.byref(self._raw), tolerance))


def peak_von_mises(positions, nodal_von_mises, point_values, inside=None) -> Optional[float]:
    """peak_von_mises (results.rs:155-162): peak_stress's value."""
    peak = peak_stress(positions, nodal_von_mises, point_values, inside)
    return None if peak is None else peak.von_mises
