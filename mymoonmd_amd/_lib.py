"""ctypes binding of libpmx.so (the C ABI declared in include/pmx.h).

The library is the product: this module only loads it and declares the
signatures.  Loading fails loudly when the library has not been built.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpmx.so")

c_int = ctypes.c_int
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p
P = ctypes.POINTER

PMX_OK = 0
ERRORS = {1: "PMX_E_ARG", 2: "PMX_E_HIP", 3: "PMX_E_ZERO_DIAG", 4: "PMX_E_DIVERGED",
          5: "PMX_E_BREAKDOWN", 6: "PMX_E_STATE", 7: "PMX_E_NOMEM", 8: "PMX_E_INTERNAL"}

BC2D_FN = ctypes.CFUNCTYPE(c_int, c_int, c_double, c_void_p)
BC3D_FN = ctypes.CFUNCTYPE(c_int, c_double, c_double, c_double, c_void_p)


class SpaceInfo(ctypes.Structure):
    _fields_ = [("n_cells", c_int), ("ndof_cell", c_int), ("n_dof", c_int),
                ("active_bound", c_int), ("n_inner", c_int), ("n_dirichlet", c_int),
                ("n_bound", c_int * 9), ("dim", c_int)]


class CDParams(ctypes.Structure):
    _fields_ = [("disc", c_int), ("example", c_int), ("hexa_quad_rule", c_int),
                ("eps", c_double), ("b", c_double * 3), ("c", c_double), ("f0", c_double),
                ("delta0", c_double), ("poly_degree", c_double),
                ("dev_coeff_table", c_void_p), ("host_dirichlet", c_void_p),
                ("time", c_double), ("theta1", c_double), ("tau", c_double),
                ("assembly_variant", c_int), ("fp32_assembly", c_int), ("overwrite", c_int)]


class BatchProblem(ctypes.Structure):
    _fields_ = [("n_side", c_int), ("eps", c_double), ("b", c_double * 2), ("c", c_double)]


class BatchParams(ctypes.Structure):
    _fields_ = [("example", c_int), ("scheme", c_int), ("tau", c_double), ("t0", c_double),
                ("inner_solver", c_int), ("gmres_restart", c_int), ("max_inner", c_int),
                ("max_outer", c_int), ("inner_reduction", c_double),
                ("residual_reduction", c_double), ("residual_tolerance", c_double), ("block", c_int),
                ("max_groups", c_int)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("n_problems", c_int), ("total_dof", c_int), ("total_active", c_int),
                ("total_cells", c_int), ("total_nnz", ctypes.c_longlong), ("block", c_int),
                ("time", c_double), ("n_workgroups", c_int)]


class BatchReport(ctypes.Structure):
    _fields_ = [("seconds", c_double), ("total_inner_iterations", ctypes.c_longlong),
                ("total_outer_iterations", ctypes.c_longlong), ("max_inner_iterations", c_int),
                ("failed_solves", c_int), ("block", c_int)]


class NseInfo(ctypes.Structure):
    _fields_ = [("n_cells", c_int), ("n_u", c_int), ("n_u_active", c_int), ("n_p", c_int),
                ("nnz_a", ctypes.c_longlong), ("nnz_b", ctypes.c_longlong), ("n_combined", c_int),
                ("nnz_combined", ctypes.c_longlong), ("pressure_space", c_int), ("n_p_cell", c_int),
                ("enclosed", c_int)]


class NseParams(ctypes.Structure):
    _fields_ = [("example", c_int), ("nu", c_double)]


class SolverParams(ctypes.Structure):
    _fields_ = [("iterative_solver_type", c_int), ("preconditioner", c_int),
                ("max_n_iterations", c_int), ("min_n_iterations", c_int),
                ("residual_tolerance", c_double), ("residual_reduction", c_double),
                ("divergence_factor", c_double), ("gmres_restart", c_int),
                ("damping", c_double), ("mg_cycle", c_int), ("mg_smoother", c_int),
                ("mg_smoother_coarse", c_int), ("mg_correction_damp", c_double),
                ("mg_n_pre_smooth", c_int), ("mg_n_post_smooth", c_int),
                ("mg_coarse_residual", c_double), ("mg_coarse_max_n_iterations", c_int),
                ("jacobi_damp", c_double), ("precision", c_int), ("vanka_damp", c_double),
                ("zero_start_shortcut", c_int), ("renumber_levels", c_int),
                ("transfer_sell", c_int), ("vanka_inverse", c_int), ("refinement", c_int),
                ("refinement_inner_reduction", c_double), ("refinement_inner_bf16", c_int),
                ("graph_cycle", c_int)]


class PodParams(ctypes.Structure):
    _fields_ = [("pod_rank", c_int), ("pod_fluctuations_only", c_int), ("pod_inner_product", c_int)]


class SolveReport(ctypes.Structure):
    _fields_ = [("iterations", c_int), ("initial_residual", c_double),
                ("final_residual", c_double), ("converged", c_int)]


# name -> argtypes (all functions return int unless listed in _RESTYPES)
SIGNATURES = {
    "pmx_last_error": [],
    "pmx_version": [],
    "pmx_ctx_create": [c_int, P(c_void_p)],
    "pmx_ctx_destroy": [c_void_p],
    "pmx_ctx_synchronize": [c_void_p],
    "pmx_domain_create": [c_int, c_int, P(c_void_p)],
    "pmx_domain_create_structured_tri": [c_int, P(c_void_p)],
    "pmx_domain_destroy": [c_void_p],
    "pmx_domain_level_info": [c_void_p, c_int, P(c_int), P(c_int), P(c_int)],
    "pmx_domain_get_vertices": [c_void_p, c_int, P(c_double)],
    "pmx_space_create": [c_void_p, c_void_p, c_int, c_int, BC2D_FN, BC3D_FN, c_void_p, c_int,
                         P(c_void_p)],
    "pmx_space_destroy": [c_void_p],
    "pmx_space_get_info": [c_void_p, P(SpaceInfo)],
    "pmx_space_get_global_numbers": [c_void_p, P(c_int)],
    "pmx_space_get_dof_coords": [c_void_p, P(c_double)],
    "pmx_space_get_structure": [c_void_p, P(c_int), P(c_int), P(c_int)],
    "pmx_space_get_prolongation": [c_void_p, c_void_p, P(c_int), P(c_int), P(c_int), P(c_double)],
    "pmx_matrix_create": [c_void_p, c_void_p, P(c_void_p)],
    "pmx_matrix_create_csr": [c_void_p, c_int, c_int, P(c_int), P(c_int), P(c_double), c_int,
                              P(c_void_p)],
    "pmx_matrix_destroy": [c_void_p],
    "pmx_matrix_get_info": [c_void_p, P(c_int), P(c_int), P(c_int)],
    "pmx_matrix_get_structure": [c_void_p, P(c_int), P(c_int)],
    "pmx_matrix_get_entries": [c_void_p, c_void_p, P(c_double)],
    "pmx_matrix_set_entries": [c_void_p, c_void_p, P(c_double)],
    "pmx_matrix_entries_device": [c_void_p, P(c_void_p)],
    "pmx_matrix_reset": [c_void_p, c_void_p],
    "pmx_matrix_set_spmv_lanes": [c_void_p, c_int],
    "pmx_matrix_apply_scaled_add": [c_void_p, c_void_p, c_void_p, c_void_p, c_double],
    "pmx_matrix_apply": [c_void_p, c_void_p, c_void_p, c_void_p],
    "pmx_matrix_time_kernel": [c_void_p, c_void_p, c_int, c_int, P(c_double)],
    "pmx_assembly_n_quad": [c_void_p, c_int, P(c_int)],
    "pmx_assembly_quad_points": [c_void_p, c_int, P(c_double)],
    "pmx_assemble_cd": [c_void_p, c_void_p, P(CDParams), c_void_p, c_void_p],
    "pmx_compute_errors": [c_void_p, c_void_p, P(CDParams), c_void_p, P(c_double)],
    "pmx_compute_errors_noalloc": [c_void_p, c_void_p, P(CDParams), c_void_p, P(c_double)],
    "pmx_vector_dot": [c_void_p, c_int, c_void_p, c_void_p, P(c_double)],
    "pmx_vector_norm": [c_void_p, c_int, c_void_p, P(c_double)],
    "pmx_vector_add_scaled": [c_void_p, c_int, c_double, c_void_p, c_void_p],
    "pmx_vector_scale": [c_void_p, c_int, c_double, c_void_p],
    "pmx_vector_copy": [c_void_p, c_int, c_void_p, c_void_p],
    "pmx_assemble_tcd": [c_void_p, c_void_p, P(CDParams), c_void_p, c_void_p, c_void_p],
    "pmx_interpolate": [c_void_p, c_void_p, P(CDParams), c_void_p],
    "pmx_matrix_scale_active": [c_void_p, c_void_p, c_double],
    "pmx_matrix_add_scaled_active": [c_void_p, c_void_p, c_void_p, c_double],
    "pmx_matrix_multiply_active": [c_void_p, c_void_p, c_void_p, c_void_p, c_double],
    "pmx_time_prepare_rhs": [c_void_p, c_int, c_int, c_double, c_void_p, c_void_p, c_void_p, c_void_p,
                             c_void_p, c_void_p],
    "pmx_time_prepare_system": [c_void_p, c_int, c_int, c_double, c_void_p, c_void_p],
    "pmx_time_reset_system": [c_void_p, c_int, c_int, c_double, c_void_p, c_void_p],
    "pmx_batch_default_params": [P(BatchParams)],
    "pmx_batch_create": [c_void_p, c_int, P(BatchProblem), P(BatchParams), P(c_void_p)],
    "pmx_batch_destroy": [c_void_p],
    "pmx_batch_get_info": [c_void_p, P(BatchInfo)],
    "pmx_batch_problem_size": [c_void_p, c_int, P(c_int), P(c_int), P(c_int)],
    "pmx_batch_initialize": [c_void_p],
    "pmx_batch_run": [c_void_p, c_int, P(BatchReport)],
    "pmx_batch_get_solution": [c_void_p, c_int, P(c_double)],
    "pmx_batch_get_iterations": [c_void_p, P(c_int), P(c_int), P(c_int)],
    "pmx_batch_get_structure": [c_void_p, c_int, P(c_int), P(c_int), P(c_int)],
    "pmx_batch_get_matrices": [c_void_p, c_int, P(c_double), P(c_double), P(c_double)],
    "pmx_batch_errors": [c_void_p, c_int, P(c_double)],
    "pmx_nse_create": [c_void_p, c_void_p, c_int, P(c_void_p)],
    "pmx_nse_create_pair": [c_void_p, c_void_p, c_int, c_int, c_int, P(c_void_p)],
    "pmx_nse_destroy": [c_void_p],
    "pmx_nse_get_info": [c_void_p, P(NseInfo)],
    "pmx_nse_set_velocity_sell_rows": [c_void_p, c_int],
    "pmx_nse_assemble": [c_void_p, c_void_p, P(NseParams), c_void_p, c_void_p],
    "pmx_nse_apply": [c_void_p, c_void_p, c_void_p, c_void_p],
    "pmx_nse_apply_scaled_add": [c_void_p, c_void_p, c_void_p, c_void_p, c_double],
    "pmx_nse_get_structure": [c_void_p, c_int, P(c_int), P(c_int), P(c_int)],
    "pmx_nse_get_values": [c_void_p, c_void_p, c_int, P(c_double)],
    "pmx_nse_get_global_numbers": [c_void_p, P(c_int), P(c_int)],
    "pmx_nse_interpolate": [c_void_p, c_void_p, c_int, c_void_p],
    "pmx_nse_project_pressure": [c_void_p, c_void_p, c_void_p],
    "pmx_nse_restrict_velocity": [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p],
    "pmx_nse_errors": [c_void_p, c_void_p, c_int, c_void_p, P(c_double)],
    "pmx_write_vtk": [c_void_p, c_int, P(c_void_p), P(ctypes.c_char_p), c_double, ctypes.c_char_p],
    "pmx_snapshots_create": [c_void_p, ctypes.c_char_p, c_int, c_int, c_int, P(c_void_p)],
    "pmx_snapshots_destroy": [c_void_p],
    "pmx_snapshots_write_data": [c_void_p, c_void_p, c_int, P(c_int)],
    "pmx_snapshots_get": [c_void_p, P(c_int), P(c_void_p)],
    "pmx_pod_default_params": [P(PodParams)],
    "pmx_pod_create": [c_void_p, P(PodParams), P(c_void_p)],
    "pmx_pod_destroy": [c_void_p],
    "pmx_pod_set_gramian": [c_void_p, c_void_p],
    "pmx_pod_read_snapshots": [c_void_p, ctypes.c_char_p],
    "pmx_pod_set_snapshots": [c_void_p, c_void_p, c_int, c_int, c_int],
    "pmx_pod_compute_basis": [c_void_p],
    "pmx_pod_get_info": [c_void_p, P(c_int), P(c_int), P(c_int), P(c_int)],
    "pmx_pod_get_eigs": [c_void_p, P(c_double)],
    "pmx_pod_get_basis": [c_void_p, c_void_p],
    "pmx_pod_get_mean": [c_void_p, c_void_p],
    "pmx_pod_write": [c_void_p, ctypes.c_char_p, ctypes.c_char_p],
    "pmx_batch_snapshots_reserve": [c_void_p, c_int],
    "pmx_batch_write_snapshot": [c_void_p],
    "pmx_batch_compute_pod": [c_void_p, P(PodParams)],
    "pmx_batch_get_pod": [c_void_p, c_int, P(c_int), P(c_double), P(c_double), P(c_double)],
    "pmx_solver_default_params": [P(SolverParams)],
    "pmx_solver_create": [c_void_p, P(SolverParams), P(c_void_p)],
    "pmx_solver_destroy": [c_void_p],
    "pmx_solver_update": [c_void_p, P(c_void_p), P(c_void_p), c_int],
    "pmx_solver_update_nse": [c_void_p, P(c_void_p), c_int],
    "pmx_solver_solve": [c_void_p, c_void_p, c_void_p, P(SolveReport)],
    "pmx_solver_get_history": [c_void_p, P(c_double), c_int, P(c_int)],
    "pmx_solver_apply_preconditioner": [c_void_p, c_void_p, c_void_p],
    "pmx_solver_time_level_kernel": [c_void_p, c_int, c_int, c_int, P(c_double)],
}
_RESTYPES = {"pmx_last_error": ctypes.c_char_p}

_lib = None


class PmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (ERRORS.get(code, code), msg))
        self.code = code


def lib():
    """Load libpmx.so once; raise if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libpmx.so is not built: run python -m mymoonmd_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, argtypes in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = argtypes
            f.restype = _RESTYPES.get(name, c_int)
        _lib = L
    return _lib


def call(name, *args):
    L = lib()
    rc = getattr(L, name)(*args)
    if rc != PMX_OK:
        raise PmxError(rc, L.pmx_last_error().decode())
    return rc
