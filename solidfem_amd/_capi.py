"""ctypes declarations of include/solidfem_hip.h and the loader of libsolidfem_hip.so.

The library is loaded from this package's directory only. There is no fallback: a missing
library raises at import of the binding, so nothing can silently run elsewhere.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIBRARY = Path(__file__).resolve().parent / "libsolidfem_hip.so"

u32, u64, i32, i64, f64 = C.c_uint32, C.c_uint64, C.c_int32, C.c_int64, C.c_double
P = C.POINTER
vp = C.c_void_p

CANCEL_FN = C.CFUNCTYPE(i32, vp)
REGION_FN = C.CFUNCTYPE(i32, P(f64), vp)
MATERIAL_FN = C.CFUNCTYPE(None, P(f64), P(f64), vp)
APPLY_FN = C.CFUNCTYPE(i32, P(f64), P(f64), u64, vp, vp)


class Preconditioner(C.Structure):
    _fields_ = [("apply", APPLY_FN), ("user", vp), ("where", i32), ("reserved", i32)]


class Progress(C.Structure):
    _fields_ = [("iteration", u64), ("relative_residual", f64), ("at_rounding_floor", i32), ("outcome", i32)]


class ColumnProgress(C.Structure):
    _fields_ = [("iteration", u64), ("relative_residual", f64), ("at_rounding_floor", i32), ("outcome", i32),
                ("error", i32), ("reserved", i32), ("error_iteration", u64), ("error_value", f64)]


class ServerOptions(C.Structure):
    _fields_ = [("hbm_budget", u64), ("precision", i32), ("coarse_solve", i32), ("tolerance", f64), ("max_columns", u32),
                ("slice_iterations", u32), ("builders", i32), ("solvers", i32), ("dot_order", i32), ("reserved", i32)]


class ServeStatus(C.Structure):
    _fields_ = [("state", i32), ("error", i32), ("cache_hit", i32), ("at_rounding_floor", i32), ("iterations", u64),
                ("relative_residual", f64), ("queued_seconds", f64), ("build_seconds", f64), ("solve_seconds", f64),
                ("latency_seconds", f64), ("nodes", u64), ("free_dofs", u64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", u64), ("completed", u64), ("failed", u64), ("builds", u64), ("cache_hits", u64),
                ("evictions", u64), ("resident_systems", u64), ("hbm_reserved", u64), ("hbm_peak", u64),
                ("p50_seconds", f64), ("p99_seconds", f64), ("mean_seconds", f64), ("max_seconds", f64)]


class PcgOptions(C.Structure):
    _fields_ = [("tolerance", f64), ("max_iterations", u64), ("dot_order", i32), ("reserved", i32)]


class MultigridOptions(C.Structure):
    _fields_ = [
        ("neighbour_reach", f64),
        ("coarse_dofs", u64),
        ("max_levels", u64),
        ("smoother_degree", u64),
        ("smoother_range", f64),
        ("precision", i32),
        ("threads", i32),
        ("coarse_solve", i32),
        ("setup_device", i32),
    ]


class MultigridStats(C.Structure):
    _fields_ = [("levels", u64), ("coarse_dofs", u64), ("operator_complexity", f64), ("memory_bytes", u64)]


class Report(C.Structure):
    _fields_ = [
        ("total_dofs", u64),
        ("free_dofs", u64),
        ("nonzeros", u64),
        ("at_rounding_floor", i32),
        ("has_multigrid", i32),
        ("iterations", u64),
        ("relative_residual", f64),
        ("assembly_seconds", f64),
        ("preconditioner_seconds", f64),
        ("multigrid", MultigridStats),
        ("solve_seconds", f64),
        ("recovery_seconds", f64),
    ]


class ErrorEstimate(C.Structure):
    _fields_ = [("error_norm_squared", f64), ("solution_norm_squared", f64), ("relative_error", f64), ("elements", u64)]


class RecoveryState(C.Structure):
    _fields_ = [("elements", u64), ("elements_done", u64), ("vertices", u64), ("fits_done", u64), ("nodes", u64),
                ("points_per_element", u64), ("converged", i32), ("finished", i32)]


class Peak(C.Structure):
    _fields_ = [("found", i32), ("at_point", i32), ("index", u64), ("von_mises", f64), ("position", f64 * 3)]


class Geometry(C.Structure):
    _fields_ = [("gradients", f64 * 30), ("det_j", f64), ("position", f64 * 3)]


class ComparedSolve(C.Structure):
    _fields_ = [("mesh", vp), ("nodal_von_mises", P(f64)), ("nodes", u64), ("nodal_stride", u64),
                ("point_values", P(f64)), ("points", u64)]


class ComparedPeak(C.Structure):
    _fields_ = [("peak", Peak), ("elements", u64), ("peak_element_volume", f64)]


class MeshComparison(C.Structure):
    _fields_ = [("coarse", ComparedPeak), ("fine", ComparedPeak), ("peak_change", f64), ("peak_distance", f64)]


class BuildProgress(C.Structure):
    _fields_ = [("elements_done", u64), ("elements_total", u64), ("setup_finished", i32), ("reserved", i32)]


class Workload(C.Structure):
    _fields_ = [
        ("origin", f64 * 3),
        ("size", f64 * 3),
        ("divisions", i32 * 3),
        ("jitter_fraction", f64),
        ("jitter_seed", u64),
        ("clamp_axis", i32),
        ("clamp_value", f64),
        ("load_axis", i32),
        ("load_value", f64),
        ("traction_kind", i32),
        ("traction", f64 * 3),
        ("traction_scale", f64),
        ("traction_seed", u64),
        ("load_case", i32),
        ("axis_u", f64),
        ("axis_v", f64),
        ("material_kind", i32),
        ("youngs", f64),
        ("poisson", f64),
        ("core_youngs", f64),
        ("core_poisson", f64),
        ("shell_depth", f64),
        ("quadrature_points", i32),
        ("wall_cells", i32),
        ("wall_thickness", f64),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "sf_last_error": (C.c_char_p, []),
    "sf_last_error_payload": (None, [P(i64), P(i64), P(f64)]),
    "sf_version": (C.c_char_p, []),
    "sf_device_count": (i32, [P(i32)]),
    "sf_context_create": (i32, [i32, P(vp)]),
    "sf_context_destroy": (None, [vp]),
    "sf_context_synchronize": (i32, [vp]),
    "sf_tet_mesh_create": (i32, [P(f64), u64, P(u32), u64, P(vp)]),
    "sf_box_tet_mesh": (i32, [P(f64), P(f64), P(i32), P(vp)]),
    "sf_tet_mesh_destroy": (None, [vp]),
    "sf_tet_mesh_positions": (None, [vp, P(f64)]),
    "sf_tet_mesh_tets": (None, [vp, P(u32)]),
    "sf_tet_mesh_node_count": (u64, [vp]),
    "sf_tet_mesh_tet_count": (u64, [vp]),
    "sf_quad_mesh_from_linear": (i32, [vp, P(vp)]),
    "sf_quad_mesh_destroy": (None, [vp]),
    "sf_quad_mesh_node_count": (u64, [vp]),
    "sf_quad_mesh_tet_count": (u64, [vp]),
    "sf_quad_mesh_vertex_count": (u64, [vp]),
    "sf_quad_mesh_nodes": (None, [vp, P(f64)]),
    "sf_quad_mesh_tets": (None, [vp, P(u32)]),
    "sf_quad_mesh_boundary_faces": (u64, [vp, P(u32)]),
    "sf_model_create": (i32, [vp, f64, f64, P(vp)]),
    "sf_model_set_shell_core": (i32, [vp, f64, f64, f64]),
    "sf_model_set_quadrature": (i32, [vp, i32]),
    "sf_model_fix_node": (i32, [vp, u32]),
    "sf_model_add_nodal_force": (i32, [vp, u32, P(f64)]),
    "sf_model_prescribe_displacement": (i32, [vp, u32, u32, f64]),
    "sf_model_add_traction": (i32, [vp, P(u32), P(f64)]),
    "sf_model_from_workload": (i32, [P(Workload), P(vp)]),
    "sf_model_destroy": (None, [vp]),
    "sf_model_mesh": (vp, [vp]),
    "sf_model_element_stiffness": (i32, [vp, vp, u64, u64, P(f64)]),
    "sf_model_free_problem": (i32, [vp, P(u64), P(u64), P(u32), P(f64), P(i64), P(u32), P(f64)]),
    "sf_study_builder_create": (i32, [vp, vp, P(PcgOptions), i32, P(MultigridOptions), P(vp)]),
    "sf_study_builder_step": (i32, [vp, u64, CANCEL_FN, vp, P(BuildProgress)]),
    "sf_study_builder_finish": (i32, [vp, P(vp)]),
    "sf_study_builder_destroy": (None, [vp]),
    "sf_static_study_create": (i32, [vp, vp, P(PcgOptions), i32, P(MultigridOptions), CANCEL_FN, vp, P(vp)]),
    "sf_static_study_step": (i32, [vp, u64, CANCEL_FN, vp, P(Progress)]),
    "sf_static_study_displacements": (i32, [vp, P(f64)]),
    "sf_static_study_report": (i32, [vp, P(Report)]),
    "sf_static_study_multigrid": (vp, [vp]),
    "sf_static_study_pcg": (vp, [vp]),
    "sf_static_study_matrix": (i32, [vp, P(u64), P(u64), P(i64), P(u32), P(f64)]),
    "sf_static_study_rhs": (i32, [vp, P(f64)]),
    "sf_static_study_destroy": (None, [vp]),
    "sf_csr_create": (i32, [vp, u64, P(i64), P(u32), P(f64), P(vp)]),
    "sf_csr_destroy": (None, [vp]),
    "sf_csr_read": (i32, [vp, P(u64), P(u64), P(i64), P(u32), P(f64)]),
    "sf_assemble_stiffness": (i32, [vp, vp, CANCEL_FN, vp, P(vp)]),
    "sf_csr_multiply": (i32, [vp, P(f64), P(f64)]),
    "sf_csr_rounding_bound": (i32, [vp, P(f64), i32, P(f64)]),
    "sf_multigrid_setup_create": (i32, [vp, vp, P(f64), u64, P(f64), u64, P(MultigridOptions), P(vp)]),
    "sf_multigrid_setup_step": (i32, [vp, u64, CANCEL_FN, vp, P(i32)]),
    "sf_multigrid_setup_finish": (i32, [vp, P(vp)]),
    "sf_multigrid_setup_destroy": (None, [vp]),
    "sf_multigrid_setup_level_count": (i32, [vp, P(u64)]),
    "sf_multigrid_setup_level_matrix": (i32, [vp, u64, i32, P(u64), P(u64), P(u64), P(u64), P(u64), P(i64), P(u32), P(f64)]),
    "sf_multigrid_setup_level_inverse_diagonal": (i32, [vp, u64, P(f64)]),
    "sf_multigrid_setup_smoother_spectrum": (i32, [vp, u64, P(f64), P(f64)]),
    "sf_multigrid_setup_aggregates": (i32, [vp, u64, P(u32), P(u64)]),
    "sf_multigrid_setup_coarse_factor": (i32, [vp, P(u64), P(f64)]),
    "sf_multigrid_destroy": (None, [vp]),
    "sf_multigrid_stats_get": (i32, [vp, P(MultigridStats)]),
    "sf_multigrid_precision": (i32, [vp]),
    "sf_multigrid_threads": (u64, [vp]),
    "sf_multigrid_apply": (i32, [vp, P(f64), P(f64)]),
    "sf_multigrid_smooth": (i32, [vp, u64, P(f64), P(f64)]),
    "sf_multigrid_level_sizes": (i32, [vp, P(u64), u64, P(u64)]),
    "sf_multigrid_level_matrix": (i32, [vp, u64, i32, P(u64), P(u64), P(u64), P(u64), P(u64), P(i64), P(u32), P(f64)]),
    "sf_multigrid_level_inverse_diagonal": (i32, [vp, u64, P(f64)]),
    "sf_multigrid_smoother_spectrum": (i32, [vp, u64, P(f64), P(f64)]),
    "sf_multigrid_aggregates": (i32, [vp, u64, P(u32), P(u64)]),
    "sf_multigrid_solve_coarsest": (i32, [vp, P(f64), u64]),
    "sf_pcg_create": (i32, [vp, vp, vp, P(f64), P(PcgOptions), P(vp)]),
    "sf_pcg_create_with": (i32, [vp, vp, P(Preconditioner), P(f64), P(PcgOptions), P(vp)]),
    "sf_pcg_step": (i32, [vp, u64, CANCEL_FN, vp, P(Progress)]),
    "sf_pcg_estimate": (i32, [vp, P(f64)]),
    "sf_pcg_progress": (i32, [vp, P(Progress)]),
    "sf_pcg_is_converged": (i32, [vp]),
    "sf_pcg_size": (u64, [vp]),
    "sf_pcg_destroy": (None, [vp]),
    "sf_static_study_system": (vp, [vp]),
    "sf_multigrid_with_precision": (i32, [vp, i32, i32, P(vp)]),
    "sf_multigrid_cycle_bytes": (f64, [vp]),
    "sf_pcg_restart": (i32, [vp, P(f64)]),
    "sf_force_variant": (i32, [i32]),
    "sf_multigrid_tuned_depth": (i32, [vp, u64, i32, P(i32)]),
    "sf_csr_tuned_depth": (i32, [vp, P(i32)]),
    "sf_csr_rounding_bound_noalloc": (i32, [vp, vp, i32, vp, vp]),
    "sf_noalloc_max_rows": (u64, []),
    "sf_setup_slot_limit": (i32, [u32]),
    "sf_profile_enable": (i32, [i32, i32]),
    "sf_profile_reset": (i32, []),
    "sf_profile_read": (i32, [i32, P(u64), P(f64), P(f64)]),
    "sf_stream_copy": (i32, [vp, u64, u32, P(f64)]),
    "sf_stream_read": (i32, [vp, u64, u32, P(f64)]),
    "sf_multigrid_apply_batch": (i32, [vp, P(f64), u64, P(f64)]),
    "sf_multigrid_set_graph": (i32, [vp, i32]),
    "sf_recover_stress": (i32, [vp, vp, P(f64), P(f64), P(f64)]),
    "sf_static_study_recover": (i32, [vp, P(f64), P(f64)]),
    "sf_recovery_create": (i32, [vp, vp, P(f64), i32, P(vp)]),
    "sf_static_study_begin_recovery": (i32, [vp, P(vp)]),
    "sf_recovery_advance": (i32, [vp, u64, CANCEL_FN, vp, P(i32)]),
    "sf_error_estimation_create": (i32, [vp, vp, P(vp)]),
    "sf_static_study_begin_error_estimate": (i32, [vp, vp, P(vp)]),
    "sf_error_estimation_advance": (i32, [vp, u64, CANCEL_FN, vp, P(i32)]),
    "sf_static_study_continue_error_estimate": (i32, [vp, vp, u64, CANCEL_FN, vp, P(i32)]),
    "sf_error_estimation_elements_done": (u64, [vp]),
    "sf_error_estimation_finish": (i32, [vp, P(ErrorEstimate), P(f64)]),
    "sf_error_estimation_destroy": (None, [vp]),
    "sf_model_set_material_field": (i32, [vp, MATERIAL_FN, vp]),
    "sf_shape_values": (None, [P(f64), P(f64)]),
    "sf_shape_derivatives": (None, [P(f64), P(f64)]),
    "sf_quadrature_points": (u64, [i32, P(f64)]),
    "sf_point_geometry": (i32, [P(f64), P(f64), P(Geometry)]),
    "sf_strain_at": (None, [P(Geometry), P(f64), P(f64)]),
    "sf_stress_at": (None, [f64, f64, P(f64), P(f64)]),
    "sf_von_mises": (f64, [P(f64)]),
    "sf_mesh_comparison_between": (i32, [P(ComparedSolve), P(ComparedSolve), REGION_FN, vp, P(MeshComparison)]),
    "sf_mesh_comparison_element_ratio": (f64, [P(MeshComparison)]),
    "sf_mesh_comparison_peak_element_ratio": (f64, [P(MeshComparison)]),
    "sf_mesh_comparison_safety_factors": (i32, [P(MeshComparison), f64, P(f64), P(f64)]),
    "sf_mesh_comparison_safety_factor_change": (i32, [P(MeshComparison), f64, P(f64)]),
    "sf_mesh_comparison_verdict": (i32, [P(MeshComparison), f64]),
    "sf_static_study_continue_recovery": (i32, [vp, vp, u64, CANCEL_FN, vp, P(i32)]),
    "sf_recovery_state_of": (i32, [vp, P(RecoveryState)]),
    "sf_recovery_displacements": (i32, [vp, P(f64)]),
    "sf_recovery_field": (i32, [vp, P(f64), P(f64)]),
    "sf_recovery_destroy": (None, [vp]),
    "sf_peak_stress": (i32, [P(f64), u64, P(f64), u64, P(f64), u64, REGION_FN, vp, P(Peak)]),
    "sf_nodal_rotations": (i32, [vp, vp, P(f64), P(f64)]),
    "sf_model_points_per_element": (u64, [vp]),
    "sf_static_study_counts": (i32, [vp, P(u64), P(u64), P(u64)]),
    "sf_pcg_batch_create": (i32, [vp, vp, vp, P(f64), u64, P(PcgOptions), P(vp)]),
    "sf_pcg_batch_step": (i32, [vp, u64, CANCEL_FN, vp, P(ColumnProgress), P(u64)]),
    "sf_pcg_batch_progress": (i32, [vp, P(ColumnProgress)]),
    "sf_pcg_batch_estimate": (i32, [vp, P(f64)]),
    "sf_pcg_batch_restart": (i32, [vp, P(f64)]),
    "sf_pcg_batch_columns": (u64, [vp]),
    "sf_pcg_batch_destroy": (None, [vp]),
    "sf_server_create": (i32, [i32, P(ServerOptions), P(vp)]),
    "sf_server_submit": (i32, [vp, P(Workload), P(u64)]),
    "sf_server_submit_model": (i32, [vp, vp, P(u64)]),
    "sf_tet_mesh_read_text": (i32, [C.c_char_p, u64, P(vp)]),
    "sf_server_status": (i32, [vp, u64, P(ServeStatus)]),
    "sf_server_wait": (i32, [vp, u64, f64, P(ServeStatus)]),
    "sf_server_result": (i32, [vp, u64, P(f64), u64]),
    "sf_server_release": (i32, [vp, u64]),
    "sf_server_stats_get": (i32, [vp, P(ServerStats)]),
    "sf_server_destroy": (None, [vp]),
}

_lib = None


def load() -> C.CDLL:
    """The HIP library, loaded once. Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIBRARY.exists():
            raise ImportError(
                f"{LIBRARY} is missing: build it with __graft_entry__.build() or python solidfem_amd/build.py"
            )
        lib = C.CDLL(os.fspath(LIBRARY), mode=C.RTLD_GLOBAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
