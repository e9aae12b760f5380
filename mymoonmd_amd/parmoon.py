"""Host-side mirror of the reference's interface for the solve path.

Names and argument meaning follow ParMooN (nahmedwias/MyMooNMD):
TDomain + refine_and_get_hierarchy_of_collections, TFESpace, FEMatrix /
BlockFEMatrix::apply(_scaled_add), Solver (default_solver_database,
update_matrix, solve) and ConvectionDiffusion (assemble, solve, output,
get_L2_error ...).  Every operation calls libpmx.so through the C ABI of
include/pmx.h; device vectors are torch tensors on the HIP device (torch is
only used as the device-memory allocator).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (BC2D_FN, BC3D_FN, BatchInfo, BatchParams, BatchProblem, BatchReport, CDParams,
                   NseInfo, NseParams, PodParams, SolveReport, SolverParams, SpaceInfo, call, c_double,
                   c_int, c_void_p)

GEOMETRY = {"UnitSquare": 0, "TwoTriangles": 1, "Default_UnitCube_Hexa": 2}
EXAMPLES = {"const": 0, "sine_laplace_2d": 1, "p0_3d": 2, "p1_3d": 3, "p2_3d": 4,
            "exp_sin_3d": 5, "sine_2d": 6, "tcd_exp": 7, "tcd_linear": 8}
TIME_SCHEME = {"backward_euler": 0, "crank_nicolson": 1, "bdf_two": 2}
IT_SOLVER = {"fgmres": 0, "cg": 1, "richardson": 2, "jacobi": 3}
PRECOND = {"no_preconditioner": 0, "jacobi": 1, "multigrid": 2}
SMOOTHER = {"jacobi": 0, "direct_solve": 1, "cell_vanka_jacobi": 2, "cell_vanka": 3}
CYCLE = {"V": 0, "W": 1, "F": 2}
BC_BUILTIN = {"all_dirichlet": 0, "sine_laplace": 1}


def _destroy(name, h):
    try:
        getattr(_lib.lib(), name)(h)
    except Exception:  # interpreter shutdown
        pass


def _ptr(t, n, device, what="vector"):
    """Device address of a float64 CUDA tensor that the C ABI will read or
    write over `n` entries.  Raises ValueError before any GPU call when the
    tensor is on the host or another device, is not float64, is not
    contiguous or is shorter than n: the kernels would otherwise dereference a
    host pointer or run past the end of the buffer."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError("%s must be a torch tensor, got %s" % (what, type(t).__name__))
    if not t.is_cuda:
        raise ValueError("%s must be a CUDA (HIP device) tensor, got device %s" % (what, t.device))
    if t.device.index != device:
        raise ValueError("%s is on device %s but the context is on device %d" % (what, t.device, device))
    if t.dtype != torch.float64:
        raise ValueError("%s must be float64, got %s" % (what, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % what)
    if t.numel() < n:
        raise ValueError("%s has %d entries, %d needed" % (what, t.numel(), n))
    return ctypes.c_void_p(t.data_ptr())


class Context:
    """One HIP device + stream (pmx_ctx)."""

    def __init__(self, device=0):
        self.h = c_void_p()
        call("pmx_ctx_create", device, ctypes.byref(self.h))
        self.device = device

    def synchronize(self):
        call("pmx_ctx_synchronize", self.h)

    def __del__(self, _d=_destroy):
        if getattr(self, "h", None) and self.h.value:
            _d("pmx_ctx_destroy", self.h)
            self.h = None


class TDomain:
    """Default geometry refined `n_refine` times; level k = collection after k refinements."""

    def __init__(self, geo_file="UnitSquare", n_refine=0, structured_n=None):
        self.h = c_void_p()
        if structured_n is not None:
            # n x n triangulation in one level (mesh of the batched solves for n != 2^k)
            call("pmx_domain_create_structured_tri", structured_n, ctypes.byref(self.h))
            self.n_refine = 0
        else:
            call("pmx_domain_create", GEOMETRY[geo_file], n_refine, ctypes.byref(self.h))
            self.n_refine = n_refine

    def level_info(self, level):
        n, nv, dim = c_int(), c_int(), c_int()
        call("pmx_domain_level_info", self.h, level, ctypes.byref(n), ctypes.byref(nv), ctypes.byref(dim))
        return n.value, nv.value, dim.value

    def vertices(self, level):
        n, nv, dim = self.level_info(level)
        out = np.zeros(n * nv * dim)
        call("pmx_domain_get_vertices", self.h, level, out.ctypes.data_as(ctypes.POINTER(c_double)))
        return out.reshape(n, nv, dim)

    def __del__(self, _d=_destroy):
        if getattr(self, "h", None) and self.h.value:
            _d("pmx_domain_destroy", self.h)
            self.h = None


class TFESpace:
    """Conforming Lagrange space (ConstructSpace numbering)."""

    def __init__(self, domain, level, order, ctx=None, bc="all_dirichlet", bc2d=None, bc3d=None):
        self.domain = domain
        self.ctx = ctx
        self.level = level
        self._bc2d = BC2D_FN(bc2d) if bc2d else BC2D_FN()
        self._bc3d = BC3D_FN(bc3d) if bc3d else BC3D_FN()
        self.h = c_void_p()
        call("pmx_space_create", ctx.h if ctx else None, domain.h, level, order, self._bc2d,
             self._bc3d, None, BC_BUILTIN[bc], ctypes.byref(self.h))
        inf = SpaceInfo()
        call("pmx_space_get_info", self.h, ctypes.byref(inf))
        self.info = inf
        self.n_dof = inf.n_dof
        self.active_bound = inf.active_bound
        self.n_cells = inf.n_cells
        self.ndof_cell = inf.ndof_cell

    def global_numbers(self):
        out = np.zeros(self.n_cells * self.ndof_cell, dtype=np.int32)
        call("pmx_space_get_global_numbers", self.h, out.ctypes.data_as(ctypes.POINTER(c_int)))
        return out.reshape(self.n_cells, self.ndof_cell)

    def n_quad(self, hexa_quad_rule=0):
        """quadrature points per cell of the assembly (pmx_assembly_n_quad)"""
        n = c_int()
        call("pmx_assembly_n_quad", self.h, hexa_quad_rule, ctypes.byref(n))
        return n.value

    def quad_points(self, hexa_quad_rule=0):
        """physical quadrature points [n_cells, n_quad, 3] at which a
        coefficient table is sampled (pmx_assembly_quad_points)"""
        nq = self.n_quad(hexa_quad_rule)
        out = np.zeros(self.n_cells * nq * 3)
        call("pmx_assembly_quad_points", self.h, hexa_quad_rule, out.ctypes.data_as(ctypes.POINTER(c_double)))
        return out.reshape(self.n_cells, nq, 3)

    def dof_coords(self):
        out = np.zeros(self.n_dof * 3)
        call("pmx_space_get_dof_coords", self.h, out.ctypes.data_as(ctypes.POINTER(c_double)))
        return out.reshape(self.n_dof, 3)

    def structure(self):
        nnz = c_int()
        call("pmx_space_get_structure", self.h, ctypes.byref(nnz), None, None)
        rp = np.zeros(self.n_dof + 1, dtype=np.int32)
        col = np.zeros(nnz.value, dtype=np.int32)
        call("pmx_space_get_structure", self.h, ctypes.byref(nnz),
             rp.ctypes.data_as(ctypes.POINTER(c_int)), col.ctypes.data_as(ctypes.POINTER(c_int)))
        return rp, col

    def prolongation_from(self, coarse):
        nnz = c_int()
        call("pmx_space_get_prolongation", coarse.h, self.h, ctypes.byref(nnz), None, None, None)
        rp = np.zeros(self.n_dof + 1, dtype=np.int32)
        col = np.zeros(nnz.value, dtype=np.int32)
        val = np.zeros(nnz.value)
        call("pmx_space_get_prolongation", coarse.h, self.h, ctypes.byref(nnz),
             rp.ctypes.data_as(ctypes.POINTER(c_int)), col.ctypes.data_as(ctypes.POINTER(c_int)),
             val.ctypes.data_as(ctypes.POINTER(c_double)))
        return rp, col, val

    def __del__(self, _d=_destroy):
        if getattr(self, "h", None) and self.h.value:
            _d("pmx_space_destroy", self.h)
            self.h = None


class FEMatrix:
    """Square FE matrix on the device (TStructure pattern + values).  With
    `csr=(row_ptr, col, val, active_bound)` a plain CSR matrix is uploaded."""

    def __init__(self, ctx, space=None, csr=None):
        self.ctx = ctx
        self.space = space
        self.h = c_void_p()
        if csr is not None:
            rp, col, val, ab = csr
            rp = np.ascontiguousarray(rp, np.int32)
            col = np.ascontiguousarray(col, np.int32)
            val = np.ascontiguousarray(val, np.float64)
            call("pmx_matrix_create_csr", ctx.h, len(rp) - 1, len(col),
                 rp.ctypes.data_as(ctypes.POINTER(c_int)), col.ctypes.data_as(ctypes.POINTER(c_int)),
                 val.ctypes.data_as(ctypes.POINTER(c_double)), ab, ctypes.byref(self.h))
        else:
            call("pmx_matrix_create", ctx.h, space.h, ctypes.byref(self.h))
        n, nnz, ab = c_int(), c_int(), c_int()
        call("pmx_matrix_get_info", self.h, ctypes.byref(n), ctypes.byref(nnz), ctypes.byref(ab))
        self.n_rows, self.nnz, self.active_bound = n.value, nnz.value, ab.value

    def _v(self, t, what):
        return _ptr(t, self.n_rows, self.ctx.device, what)

    def reset(self):
        call("pmx_matrix_reset", self.ctx.h, self.h)

    def get_entries(self):
        out = np.zeros(self.nnz)
        call("pmx_matrix_get_entries", self.ctx.h, self.h, out.ctypes.data_as(ctypes.POINTER(c_double)))
        return out

    def set_entries(self, vals):
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        call("pmx_matrix_set_entries", self.ctx.h, self.h, vals.ctypes.data_as(ctypes.POINTER(c_double)))

    def get_structure(self):
        rp = np.zeros(self.n_rows + 1, dtype=np.int32)
        col = np.zeros(self.nnz, dtype=np.int32)
        call("pmx_matrix_get_structure", self.h, rp.ctypes.data_as(ctypes.POINTER(c_int)),
             col.ctypes.data_as(ctypes.POINTER(c_int)))
        return rp, col

    def time_kernel(self, kernel, reps=20):
        """mean duration (ms) of one launch of an operator kernel (HIP events)"""
        out = c_double()
        call("pmx_matrix_time_kernel", self.ctx.h, self.h, {"apply_fp64": 0, "jacobi_fp64": 1,
             "apply_bf16": 2, "jacobi_bf16": 3}[kernel], reps, ctypes.byref(out))
        return out.value

    def set_spmv_lanes(self, lanes):
        call("pmx_matrix_set_spmv_lanes", self.h, lanes)

    def apply(self, x, y):
        """BlockFEMatrix::apply: y = A x (identity on non-active rows)."""
        import torch
        torch.cuda.synchronize()
        call("pmx_matrix_apply", self.ctx.h, self.h, self._v(x, "x"), self._v(y, "y"))
        self.ctx.synchronize()

    def apply_scaled_add(self, x, y, a=1.0):
        """BlockFEMatrix::apply_scaled_add: y += a A x (y += a x on non-active rows)."""
        import torch
        torch.cuda.synchronize()
        call("pmx_matrix_apply_scaled_add", self.ctx.h, self.h, self._v(x, "x"), self._v(y, "y"), a)
        self.ctx.synchronize()

    def multiply_active(self, x, y, a=1.0):
        """FEMatrix::multiplyActive: y[i] += a (A x)[i] on the active rows."""
        import torch
        torch.cuda.synchronize()
        call("pmx_matrix_multiply_active", self.ctx.h, self.h, self._v(x, "x"), self._v(y, "y"), a)
        self.ctx.synchronize()

    def scale_active(self, factor):
        """FEMatrix::scaleActive"""
        call("pmx_matrix_scale_active", self.ctx.h, self.h, factor)

    def add_scaled_active(self, other, factor=1.0):
        """BlockFEMatrix::add_blockfe_matrix (active entries += factor * other)"""
        call("pmx_matrix_add_scaled_active", self.ctx.h, self.h, other.h, factor)

    def __del__(self, _d=_destroy):
        if getattr(self, "h", None) and self.h.value:
            _d("pmx_matrix_destroy", self.h)
            self.h = None


def cd_params(example="const", disc="galerkin", eps=1.0, b=(0.0, 0.0, 0.0), c=0.0, f0=0.0,
              hexa_quad_rule=0, delta0=1.0, poly_degree=1.0, time=0.0, theta1=1.0, tau=0.0,
              coeff_table=None, dirichlet=None, assembly_variant=0, fp32_assembly=False, overwrite=False):
    """pmx_cd_params.  `coeff_table`: optional float64 device tensor
    [n_cells, n_quad, 8] with the coefficients (eps, b1, b2, b3, c, f, -, -) at
    the points of TFESpace.quad_points(hexa_quad_rule) -- what a user
    CoeffFct3D returns there (LocalAssembling::GetParameters,
    LocalAssembling.C:766-840); it replaces the constants and the example's
    right-hand side.  `dirichlet`: optional float64 host array with the
    boundary values of the Dirichlet DOFs (DOFs active_bound.. in order),
    replacing the example's.  Both are checked against the space when the
    parameters are used.  `assembly_variant` / `fp32_assembly`: the Q2
    hexahedron assembly route (0: element GEMM + row gather where it applies,
    1: the quadrature kernel) and FP32 matrix cores for its element
    matrices.  `overwrite`: assemble_cd sets the matrix and rhs instead of
    adding to them."""
    p = CDParams()
    p._table = coeff_table
    p._dirichlet = None
    if dirichlet is not None:
        p._dirichlet = np.ascontiguousarray(dirichlet, dtype=np.float64)
    p.disc = {"galerkin": 0, "supg": 1, "gls": 2}[disc]
    p.example = EXAMPLES[example]
    p.hexa_quad_rule = hexa_quad_rule
    p.eps = eps
    for k in range(3):
        p.b[k] = b[k] if k < len(b) else 0.0
    p.c = c
    p.f0 = f0
    p.delta0 = delta0
    p.poly_degree = poly_degree
    p.time, p.theta1, p.tau = time, theta1, tau
    p.assembly_variant = assembly_variant
    p.fp32_assembly = 1 if fp32_assembly else 0
    p.overwrite = 1 if overwrite else 0
    return p


def _bind_user_data(ctx, space, params):
    """Set the coefficient-table and Dirichlet-value pointers of `params` for
    this space after checking their sizes (ValueError before any GPU call)."""
    table = getattr(params, "_table", None)
    params.dev_coeff_table = None
    params.host_dirichlet = None
    if table is not None:
        nq = space.n_quad(params.hexa_quad_rule)
        params.dev_coeff_table = _ptr(table, space.n_cells * nq * 8, ctx.device, "coeff_table").value
    d = getattr(params, "_dirichlet", None)
    if d is not None:
        n_dir = space.n_dof - space.active_bound
        if d.ndim != 1 or d.size < n_dir:
            raise ValueError("dirichlet has %d entries, %d needed" % (d.size, n_dir))
        params.host_dirichlet = d.ctypes.data


def assemble_cd(ctx, space, params, matrix, rhs):
    """Assemble3D/Assemble2D + LocalAssembling(ConvDiff): adds into matrix and
    rhs (sets them with params.overwrite)."""
    import torch
    rp = _ptr(rhs, space.n_dof, ctx.device, "rhs")
    _bind_user_data(ctx, space, params)
    torch.cuda.synchronize()
    call("pmx_assemble_cd", ctx.h, space.h, ctypes.byref(params), matrix.h, rp)
    ctx.synchronize()


def assemble_tcd(ctx, space, params, stiffness=None, mass=None, rhs=None):
    """Assemble2D/3D + LocalAssembling(TCDStiffMassRhs / TCDStiffRhs): adds into
    the given stiffness matrix, mass matrix and rhs (each optional)."""
    import torch
    rp = _ptr(rhs, space.n_dof, ctx.device, "rhs") if rhs is not None else None
    _bind_user_data(ctx, space, params)
    torch.cuda.synchronize()
    call("pmx_assemble_tcd", ctx.h, space.h, ctypes.byref(params), stiffness.h if stiffness else None,
         mass.h if mass else None, rp)
    ctx.synchronize()


def interpolate(ctx, space, params, u):
    """TFEFunction::Interpolate of the example's exact solution at params.time."""
    import torch
    torch.cuda.synchronize()
    call("pmx_interpolate", ctx.h, space.h, ctypes.byref(params), _ptr(u, space.n_dof, ctx.device, "u"))
    ctx.synchronize()


def vector_copy(ctx, x, y):
    """y = x for two equally long (sub)vectors (BlockVector::copy / copy_nonactive)."""
    call("pmx_vector_copy", ctx.h, x.numel(), _ptr(x, x.numel(), ctx.device, "x"), _ptr(y, x.numel(), ctx.device, "y"))


def vector_dot(ctx, x, y):
    """BlockVector::dot (src/Matrix/BlockVector.C:491-530) of two equally long (sub)vectors."""
    import torch
    if x.numel() != y.numel():
        raise ValueError("dot of vectors of lengths %d and %d" % (x.numel(), y.numel()))
    torch.cuda.synchronize()
    out = c_double()
    call("pmx_vector_dot", ctx.h, x.numel(), _ptr(x, x.numel(), ctx.device, "x"),
         _ptr(y, x.numel(), ctx.device, "y"), ctypes.byref(out))
    return out.value


def vector_norm(ctx, x):
    """BlockVector::norm (src/Matrix/BlockVector.C:293-386): Euclidean norm."""
    import torch
    torch.cuda.synchronize()
    out = c_double()
    call("pmx_vector_norm", ctx.h, x.numel(), _ptr(x, x.numel(), ctx.device, "x"), ctypes.byref(out))
    return out.value


def vector_add_scaled(ctx, x, y, a):
    """BlockVector::add_scaled: y += a x."""
    import torch
    if x.numel() != y.numel():
        raise ValueError("add_scaled of vectors of lengths %d and %d" % (x.numel(), y.numel()))
    torch.cuda.synchronize()
    call("pmx_vector_add_scaled", ctx.h, x.numel(), a, _ptr(x, x.numel(), ctx.device, "x"),
         _ptr(y, x.numel(), ctx.device, "y"))
    ctx.synchronize()


def vector_scale(ctx, x, a):
    call("pmx_vector_scale", ctx.h, x.numel(), a, _ptr(x, x.numel(), ctx.device, "x"))


def compute_errors(ctx, space, params, u):
    import torch
    torch.cuda.synchronize()
    out = (c_double * 4)()
    up = _ptr(u, space.n_dof, ctx.device, "u")
    _bind_user_data(ctx, space, params)
    call("pmx_compute_errors", ctx.h, space.h, ctypes.byref(params), up, out)
    return list(out)


def default_solver_database():
    p = SolverParams()
    call("pmx_solver_default_params", ctypes.byref(p))
    return p


def solver_params(**kw):
    """Solver parameters by the reference's database keys."""
    p = default_solver_database()
    keymap = {"max_n_iterations": "max_n_iterations", "min_n_iterations": "min_n_iterations",
              "residual_tolerance": "residual_tolerance", "residual_reduction": "residual_reduction",
              "divergence_factor": "divergence_factor", "gmres_restart": "gmres_restart",
              "damping": "damping", "multigrid_correction_damp_factor": "mg_correction_damp",
              "multigrid_n_pre_smooth": "mg_n_pre_smooth", "multigrid_n_post_smooth": "mg_n_post_smooth",
              "multigrid_coarse_residual": "mg_coarse_residual",
              "multigrid_coarse_max_n_iterations": "mg_coarse_max_n_iterations",
              "jacobi_damp": "jacobi_damp", "multigrid_vanka_damp_factor": "vanka_damp",
              "zero_start_shortcut": "zero_start_shortcut", "renumber_levels": "renumber_levels",
              "transfer_sell": "transfer_sell", "vanka_inverse": "vanka_inverse", "refinement": "refinement",
              "refinement_inner_reduction": "refinement_inner_reduction",
              "refinement_inner_bf16": "refinement_inner_bf16", "graph_cycle": "graph_cycle"}
    for k, v in kw.items():
        if k == "iterative_solver_type":
            p.iterative_solver_type = IT_SOLVER[v]
        elif k == "preconditioner":
            p.preconditioner = PRECOND[v]
        elif k == "multigrid_smoother":
            p.mg_smoother = SMOOTHER[v]
        elif k == "multigrid_smoother_coarse":
            p.mg_smoother_coarse = SMOOTHER[v]
        elif k == "multigrid_cycle_type":
            p.mg_cycle = CYCLE[v]
        elif k == "precision":
            p.precision = {"fp64": 0, "mixed": 1}[v]
        else:
            setattr(p, keymap[k], v)
    return p


class Solver:
    """Solver<BlockFEMatrix, BlockVector> with an iterative method (+ multigrid)."""

    def __init__(self, ctx, params):
        self.ctx = ctx
        self.params = params
        self.h = c_void_p()
        self.n = None
        call("pmx_solver_create", ctx.h, ctypes.byref(params), ctypes.byref(self.h))

    def _v(self, t, what):
        if self.n is None:
            raise ValueError("update the solver with its matrices before passing vectors")
        return _ptr(t, self.n, self.ctx.device, what)

    def update_matrix(self, matrices, spaces=None):
        """matrices/spaces of all multigrid levels, coarsest first (or one matrix)."""
        if not isinstance(matrices, (list, tuple)):
            matrices = [matrices]
        mats = (c_void_p * len(matrices))(*[m.h.value for m in matrices])
        sps = (c_void_p * len(matrices))(*[s.h.value for s in spaces]) if spaces else None
        import torch
        torch.cuda.synchronize()
        call("pmx_solver_update", self.h, mats, sps, len(matrices))
        self.n = matrices[-1].n_rows

    def update_saddle(self, systems):
        """Solver::update_matrix for NSE systems (NSESystem), coarsest first."""
        if not isinstance(systems, (list, tuple)):
            systems = [systems]
        arr = (c_void_p * len(systems))(*[x.h.value for x in systems])
        import torch
        torch.cuda.synchronize()
        call("pmx_solver_update_nse", self.h, arr, len(systems))
        self.n = systems[-1].n

    def solve(self, rhs, sol):
        import torch
        torch.cuda.synchronize()
        rep = SolveReport()
        call("pmx_solver_solve", self.h, self._v(rhs, "rhs"), self._v(sol, "sol"), ctypes.byref(rep))
        return rep

    def history(self):
        n = c_int()
        call("pmx_solver_get_history", self.h, None, 0, ctypes.byref(n))
        out = np.zeros(max(n.value, 1))
        call("pmx_solver_get_history", self.h, out.ctypes.data_as(ctypes.POINTER(c_double)), n.value,
             ctypes.byref(n))
        return out[: n.value]

    def apply_preconditioner(self, r, z):
        import torch
        torch.cuda.synchronize()
        call("pmx_solver_apply_preconditioner", self.h, self._v(r, "r"), self._v(z, "z"))

    def time_level_kernel(self, level, kernel, reps=20):
        """mean duration (ms) of a level's smoothing sweep as the cycle runs it"""
        out = c_double()
        call("pmx_solver_time_level_kernel", self.h, level, {"jacobi_fp64": 1, "jacobi_bf16": 3}[kernel], reps,
             ctypes.byref(out))
        return out.value

    def __del__(self, _d=_destroy):
        if getattr(self, "h", None) and self.h.value:
            _d("pmx_solver_destroy", self.h)
            self.h = None


class ConvectionDiffusion:
    """ConvectionDiffusion<d> (src/System/ConvectionDiffusion.C): one system per
    multigrid level, assemble() / solve() / output() and the error getters."""

    def __init__(self, ctx, domain, order, params, solver, n_mg_levels=1, bc="all_dirichlet",
                 bc2d=None, bc3d=None):
        import torch
        self.ctx = ctx
        self.params = params
        nref = domain.n_refine
        levels = list(range(nref - n_mg_levels + 1, nref + 1))
        self.spaces = [TFESpace(domain, l, order, ctx, bc=bc, bc2d=bc2d, bc3d=bc3d) for l in levels]
        self.matrices = [FEMatrix(ctx, s) for s in self.spaces]
        dev = torch.device("cuda", ctx.device)
        self.rhs = [torch.zeros(s.n_dof, dtype=torch.float64, device=dev) for s in self.spaces]
        self.solution = torch.zeros(self.spaces[-1].n_dof, dtype=torch.float64, device=dev)
        self.solver = Solver(ctx, solver)
        self.errors = None
        self.report = None

    def assemble(self):
        # the reference resets matrix and rhs and adds the cell contributions
        # (ConvectionDiffusion.C:329-340); pmx_assemble_cd with overwrite
        # gives the same result without the separate reset
        ow = self.params.overwrite
        self.params.overwrite = 1
        try:
            for s, A, f in zip(self.spaces, self.matrices, self.rhs):
                assemble_cd(self.ctx, s, self.params, A, f)
        finally:
            self.params.overwrite = ow
        # solution.copy_nonactive(rhs)
        ab = self.spaces[-1].active_bound
        vector_copy(self.ctx, self.rhs[-1][ab:], self.solution[ab:])
        self.ctx.synchronize()

    def solve(self):
        if self.solver.params.preconditioner == PRECOND["multigrid"]:
            self.solver.update_matrix(self.matrices, self.spaces)
        else:
            self.solver.update_matrix(self.matrices[-1])
        self.report = self.solver.solve(self.rhs[-1], self.solution)
        return self.report

    def output(self):
        self.errors = compute_errors(self.ctx, self.spaces[-1], self.params, self.solution)
        return self.errors

    def get_L2_error(self):
        return self.errors[0]

    def get_H1_semi_error(self):
        return self.errors[1]

    def get_SD_error(self):
        return self.errors[2]

    def get_L_inf_error(self):
        return self.errors[3]


class TimeDiscretization:
    """TimeDiscretization (src/TimeDiscRoutines/TimeDiscretizations.C): the
    scheme, the time and the step counter, with the rhs / system-matrix steps
    delegated to the C ABI."""

    def __init__(self, ctx, time_discretization="backward_euler", time_step_length=0.05,
                 time_start=0.0, time_end=1.0):
        self.ctx = ctx
        self.scheme = TIME_SCHEME[time_discretization]
        self.tau = time_step_length
        self.start_time, self.end_time = time_start, time_end
        self.current_time_ = time_start
        self.current_step_ = 0

    def n_old_solutions(self):
        return 2 if self.scheme == TIME_SCHEME["bdf_two"] else 1

    def theta1(self):
        """TDatabase::TimeDB->THETA1 as set_time_disc_parameters sets it (:341-381)."""
        if self.scheme == TIME_SCHEME["crank_nicolson"]:
            return 0.5
        if self.scheme == TIME_SCHEME["bdf_two"] and self.current_step_ >= 2:
            return 2.0 / 3.0
        return 1.0

    def reached_final_time_step(self):
        return self.current_time_ >= self.end_time - 1e-10

    def prepare_rhs_from_time_disc(self, stiffness, mass, rhs, old_rhs, u_m1, u_m2=None):
        def v(t, what):
            return _ptr(t, stiffness.n_rows, self.ctx.device, what)
        call("pmx_time_prepare_rhs", self.ctx.h, self.scheme, self.current_step_, self.tau,
             stiffness.h, mass.h, v(rhs, "rhs"), v(old_rhs, "old_rhs") if old_rhs is not None else None,
             v(u_m1, "u_m1"), v(u_m2, "u_m2") if u_m2 is not None else None)

    def prepare_system_matrix(self, stiffness, mass):
        call("pmx_time_prepare_system", self.ctx.h, self.scheme, self.current_step_, self.tau,
             stiffness.h, mass.h)

    def reset_linear_matrices(self, stiffness, mass):
        call("pmx_time_reset_system", self.ctx.h, self.scheme, self.current_step_, self.tau,
             stiffness.h, mass.h)


class TimeConvectionDiffusion:
    """TimeConvectionDiffusion<d> (src/System/TimeConvectionDiffusion.C):
    systems per multigrid level with stiffness, mass, rhs;
    assemble_initial_time / assemble / solve / output as in the reference."""

    def __init__(self, ctx, domain, order, example, time_disc, solver, disc="galerkin", eps=1.0,
                 b=(0.0, 0.0, 0.0), c=0.0, n_mg_levels=1, bc="all_dirichlet"):
        import torch
        self.ctx = ctx
        self.example, self.disc, self.eps, self.b, self.c = example, disc, eps, b, c
        self.tss = time_disc
        nref = domain.n_refine
        levels = list(range(nref - n_mg_levels + 1, nref + 1))
        self.spaces = [TFESpace(domain, l, order, ctx, bc=bc) for l in levels]
        self.stiffness = [FEMatrix(ctx, s) for s in self.spaces]
        self.mass = [FEMatrix(ctx, s) for s in self.spaces]
        dev = torch.device("cuda", ctx.device)
        z = lambda s: torch.zeros(s.n_dof, dtype=torch.float64, device=dev)  # noqa: E731
        self.rhs = [z(s) for s in self.spaces]
        fs = self.spaces[-1]
        self.solution, self.solution_m1, self.solution_m2 = z(fs), z(fs), z(fs)
        self.old_rhs, self.rhs_from_time_disc = z(fs), z(fs)
        self.solver = Solver(ctx, solver)
        self.multigrid = solver.preconditioner == PRECOND["multigrid"]
        self.report = None
        # initial condition: Interpolate(example.get_initial_cond(0)) at time_start
        interpolate(ctx, fs, self.params(), self.solution)

    def params(self):
        return cd_params(self.example, disc=self.disc, eps=self.eps, b=self.b, c=self.c,
                         time=self.tss.current_time_, theta1=self.tss.theta1(), tau=self.tss.tau)

    def _assemble(self, with_mass):
        p = self.params()
        for s, A, M, f in zip(self.spaces, self.stiffness, self.mass, self.rhs):
            A.reset()
            f.zero_()
            if with_mass:
                M.reset()
            assemble_tcd(self.ctx, s, p, A, M if with_mass else None, f)

    def assemble_initial_time(self):
        self._assemble(True)
        vector_copy(self.ctx, self.rhs[-1], self.old_rhs)
        vector_copy(self.ctx, self.solution, self.solution_m1)
        vector_copy(self.ctx, self.solution_m1, self.solution_m2)
        self.ctx.synchronize()

    def assemble(self):
        self._assemble(self.disc == "supg")
        ab = self.spaces[-1].active_bound
        f, rtd = self.rhs[-1], self.rhs_from_time_disc
        vector_copy(self.ctx, f, rtd)
        self.tss.prepare_rhs_from_time_disc(self.stiffness[-1], self.mass[-1], rtd, self.old_rhs,
                                            self.solution_m1, self.solution_m2)
        vector_copy(self.ctx, f, self.old_rhs)
        vector_copy(self.ctx, f[ab:], rtd[ab:])
        for A, M in zip(self.stiffness, self.mass):
            self.tss.prepare_system_matrix(A, M)
        vector_copy(self.ctx, f[ab:], self.solution[ab:])
        self.ctx.synchronize()

    def solve(self):
        if self.multigrid:
            self.solver.update_matrix(self.stiffness, self.spaces)
        else:
            self.solver.update_matrix(self.stiffness[-1])
        self.report = self.solver.solve(self.rhs_from_time_disc, self.solution)
        for A, M in zip(self.stiffness, self.mass):
            self.tss.reset_linear_matrices(A, M)
        vector_copy(self.ctx, self.solution_m1, self.solution_m2)
        vector_copy(self.ctx, self.solution, self.solution_m1)
        self.ctx.synchronize()
        return self.report

    def output(self):
        """L2, H1-semi, SD, L_inf errors at the current time."""
        return compute_errors(self.ctx, self.spaces[-1], self.params(), self.solution)


class TimeConvectionDiffusionPOD:
    """TimeConvectionDiffusionPOD<d> (src/ReducedOrderModels/
    TimeConvectionDiffusionPOD.C:37-273): the POD of TCD snapshots on a
    space of the domain; with pod_inner_product "L2" the Gramian is the
    Galerkin mass matrix (assemble_gramian, TCDMassOnly), and
    compute_pod_basis writes <pod_basename>.<inner product>.{pod,mean,eigs}."""

    def __init__(self, ctx, domain, order, example="tcd_exp", pod_directory=".", pod_basename="parmoon_pod",
                 pod_rank=0, pod_fluctuations_only=True, pod_inner_product="euclidean", eps=1.0,
                 b=(0.0, 0.0, 0.0), c=0.0, bc="all_dirichlet"):
        import os
        import torch
        self.ctx = ctx
        self.space = TFESpace(domain, domain.n_refine, order, ctx, bc=bc)
        self.pod_directory, self.pod_basename = pod_directory, pod_basename
        self.inner_product = pod_inner_product
        self.pod = POD(ctx, pod_rank, pod_fluctuations_only, pod_inner_product)
        self.gramian = FEMatrix(ctx, self.space)
        self._params = cd_params(example, disc="galerkin", eps=eps, b=b, c=c)
        self._rhs = torch.zeros(self.space.n_dof, dtype=torch.float64, device="cuda:%d" % ctx.device)
        os.makedirs(pod_directory, exist_ok=True)

    def assemble_gramian(self):
        self.gramian.reset()
        assemble_tcd(self.ctx, self.space, self._params, None, self.gramian, None)

    def compute_pod_basis(self, snapshots):
        """snapshots: a SnapshotsCollector or the path of a snapshot file"""
        if self.inner_product == "L2":
            self.assemble_gramian()
            self.pod.set_gramian(self.gramian)
        if isinstance(snapshots, str):
            self.pod.read_snapshots(snapshots)
        else:
            self.pod.set_snapshots(snapshots)
        self.pod.compute_basis()
        self.pod.write_pod(self.pod_directory, self.pod_basename + "." + self.inner_product + ".")


class BatchTCD:
    """Many independent TCD problems (P1 on n x n unit-square triangulations,
    own eps, b, c) advanced together by backward Euler (pmx_batch_*).
    `problems` = list of (n_side, eps, (b1, b2), c)."""

    def __init__(self, ctx, problems, example="tcd_exp", tau=0.01, t0=0.0, inner_solver="bi_cgstab",
                 gmres_restart=30, max_inner=2000, max_outer=50, inner_reduction=1e-2,
                 residual_reduction=1e-10, residual_tolerance=0.0, block=0, max_groups=16):
        self.ctx = ctx
        prm = BatchParams()
        call("pmx_batch_default_params", ctypes.byref(prm))
        prm.example = EXAMPLES[example]
        prm.scheme = TIME_SCHEME["backward_euler"]
        prm.tau, prm.t0 = tau, t0
        prm.inner_solver = {"gmres": 0, "bi_cgstab": 1}[inner_solver]
        prm.gmres_restart, prm.max_inner, prm.max_outer = gmres_restart, max_inner, max_outer
        prm.inner_reduction, prm.residual_reduction = inner_reduction, residual_reduction
        prm.residual_tolerance, prm.block = residual_tolerance, block
        prm.max_groups = max_groups
        arr = (BatchProblem * len(problems))()
        for k, (n, eps, b, c) in enumerate(problems):
            arr[k].n_side, arr[k].eps, arr[k].c = n, eps, c
            arr[k].b[0], arr[k].b[1] = b[0], b[1]
        self.params = prm
        self.h = c_void_p()
        call("pmx_batch_create", ctx.h, len(problems), arr, ctypes.byref(prm), ctypes.byref(self.h))
        self.n_problems = len(problems)

    def info(self):
        inf = BatchInfo()
        call("pmx_batch_get_info", self.h, ctypes.byref(inf))
        return inf

    def problem_size(self, p):
        n, ab, nnz = c_int(), c_int(), c_int()
        call("pmx_batch_problem_size", self.h, p, ctypes.byref(n), ctypes.byref(ab), ctypes.byref(nnz))
        return n.value, ab.value, nnz.value

    def initialize(self):
        call("pmx_batch_initialize", self.h)

    def run(self, n_steps):
        rep = BatchReport()
        call("pmx_batch_run", self.h, n_steps, ctypes.byref(rep))
        return rep

    def solution(self, p):
        n, _, _ = self.problem_size(p)
        out = np.zeros(n)
        call("pmx_batch_get_solution", self.h, p, out.ctypes.data_as(ctypes.POINTER(c_double)))
        return out

    def iterations(self):
        inner = np.zeros(self.n_problems, np.int32)
        outer = np.zeros(self.n_problems, np.int32)
        failed = np.zeros(self.n_problems, np.int32)
        P = ctypes.POINTER(c_int)
        call("pmx_batch_get_iterations", self.h, inner.ctypes.data_as(P), outer.ctypes.data_as(P),
             failed.ctypes.data_as(P))
        return inner, outer, failed

    def structure(self, p):
        n, ab, nnz = self.problem_size(p)
        rp = np.zeros(n + 1, np.int32)
        col = np.zeros(nnz, np.int32)
        P = ctypes.POINTER(c_int)
        call("pmx_batch_get_structure", self.h, p, rp.ctypes.data_as(P), col.ctypes.data_as(P), None)
        return rp, col

    def global_numbers(self, p, n_cells):
        g = np.zeros(n_cells * 3, np.int32)
        call("pmx_batch_get_structure", self.h, p, None, None, g.ctypes.data_as(ctypes.POINTER(c_int)))
        return g.reshape(n_cells, 3)

    def matrices(self, p):
        _, _, nnz = self.problem_size(p)
        a, m, s = np.zeros(nnz), np.zeros(nnz), np.zeros(nnz)
        P = ctypes.POINTER(c_double)
        call("pmx_batch_get_matrices", self.h, p, a.ctypes.data_as(P), m.ctypes.data_as(P), s.ctypes.data_as(P))
        return a, m, s

    def errors(self, p):
        out = (c_double * 4)()
        call("pmx_batch_errors", self.h, p, out)
        return list(out)

    # snapshots and POD of every problem (pmx_batch_snapshots_* / _pod)
    def reserve_snapshots(self, max_snaps):
        call("pmx_batch_snapshots_reserve", self.h, max_snaps)
        self.n_snaps = 0

    def write_snapshot(self):
        call("pmx_batch_write_snapshot", self.h)
        self.n_snaps += 1

    def compute_pod(self, pod_rank=0, pod_fluctuations_only=True, pod_inner_product="euclidean"):
        call("pmx_batch_compute_pod", self.h, ctypes.byref(pod_params(pod_rank, pod_fluctuations_only,
                                                                       pod_inner_product)))

    def pod(self, p):
        """(eigenvalues largest first, basis n_dof x rank, mean) of problem p"""
        n, _, _ = self.problem_size(p)
        r = c_int()
        call("pmx_batch_get_pod", self.h, p, ctypes.byref(r), None, None, None)
        eigs = np.zeros(self.n_snaps)
        basis = np.zeros((max(r.value, 1), n))
        mean = np.zeros(n)
        P = ctypes.POINTER(c_double)
        call("pmx_batch_get_pod", self.h, p, ctypes.byref(r), eigs.ctypes.data_as(P), basis.ctypes.data_as(P),
             mean.ctypes.data_as(P))
        return eigs, basis[:r.value].T.copy(), mean

    def __del__(self, _d=_destroy):
        if getattr(self, "h", None) and self.h.value:
            _d("pmx_batch_destroy", self.h)
            self.h = None


def write_vtk(space, functions, path, time=0.0):
    """DataWriter::writeVtk for scalar FE functions: functions = {name: CUDA
    tensor of the space's DOF values}"""
    import torch
    names = list(functions)
    dev = space.ctx.device if space.ctx is not None else 0
    ptrs = (c_void_p * max(1, len(names)))(*[_ptr(functions[k], space.n_dof, dev, k).value for k in names])
    cnames = (ctypes.c_char_p * max(1, len(names)))(*[k.encode() for k in names])
    torch.cuda.synchronize()
    call("pmx_write_vtk", space.h, len(names), ptrs, cnames, time, path.encode())


POD_INNER_PRODUCT = {"euclidean": 0, "L2": 1}


def pod_params(pod_rank=0, pod_fluctuations_only=True, pod_inner_product="euclidean"):
    """POD::default_pod_database (src/ReducedOrderModels/POD.C:20-60) with overrides"""
    p = PodParams()
    call("pmx_pod_default_params", ctypes.byref(p))
    p.pod_rank = pod_rank
    p.pod_fluctuations_only = 1 if pod_fluctuations_only else 0
    p.pod_inner_product = POD_INNER_PRODUCT[pod_inner_product]
    return p


class SnapshotsCollector:
    """SnapshotsCollector (src/ReducedOrderModels/SnapshotsCollector.C):
    write_data(solution, time_step) keeps every steps_per_snap-th solution in
    device memory and, with write_snaps, appends it to
    snaps_directory/snaps_basename in the reference's text format."""

    def __init__(self, ctx, length, max_snaps, snaps_directory=".", snaps_basename="parmoon_snapshots",
                 steps_per_snap=5, write_snaps=True):
        import os
        self.ctx = ctx
        self.length = length
        self.filename = None
        if write_snaps:
            os.makedirs(snaps_directory, exist_ok=True)
            self.filename = os.path.join(snaps_directory, snaps_basename)
        self.h = c_void_p()
        call("pmx_snapshots_create", ctx.h, self.filename.encode() if self.filename else None, steps_per_snap,
             length, max_snaps, ctypes.byref(self.h))

    def write_data(self, solution, time_step=0):
        import torch
        torch.cuda.synchronize()   # torch's stream has produced the tensor
        w = c_int()
        call("pmx_snapshots_write_data", self.h, _ptr(solution, self.length, self.ctx.device, "solution"), time_step,
             ctypes.byref(w))
        return bool(w.value)

    def count(self):
        n = c_int()
        call("pmx_snapshots_get", self.h, ctypes.byref(n), None)
        return n.value

    def device_matrix(self):
        """(pointer, n_snaps): column-major length x n_snaps in device memory"""
        n, d = c_int(), c_void_p()
        call("pmx_snapshots_get", self.h, ctypes.byref(n), ctypes.byref(d))
        return d, n.value

    def __del__(self, _d=_destroy):
        if getattr(self, "h", None) and self.h.value:
            _d("pmx_snapshots_destroy", self.h)
            self.h = None


class POD:
    """POD (include/ReducedOrderModels/POD.h): method of snapshots with the
    Gram matrix on the device.  set_gramian(FEMatrix) for the L2 inner product
    (TimeConvectionDiffusionPOD::assemble_gramian assembles the mass matrix)."""

    def __init__(self, ctx, pod_rank=0, pod_fluctuations_only=True, pod_inner_product="euclidean"):
        self.ctx = ctx
        self.h = c_void_p()
        call("pmx_pod_create", ctx.h, ctypes.byref(pod_params(pod_rank, pod_fluctuations_only, pod_inner_product)),
             ctypes.byref(self.h))

    def set_gramian(self, matrix):
        self._gram = matrix
        call("pmx_pod_set_gramian", self.h, matrix.h)

    def read_snapshots(self, path):
        call("pmx_pod_read_snapshots", self.h, path.encode())

    def set_snapshots(self, snaps):
        """snaps: collector, or a CUDA tensor (n_snaps, length) (row j = snapshot j)"""
        if isinstance(snaps, SnapshotsCollector):
            d, m = snaps.device_matrix()
            call("pmx_pod_set_snapshots", self.h, d, snaps.length, m, snaps.length)
        else:
            import torch
            if snaps.dim() != 2:
                raise ValueError("snapshots must be a (n_snaps, length) tensor")
            torch.cuda.synchronize()
            call("pmx_pod_set_snapshots", self.h, _ptr(snaps, snaps.numel(), self.ctx.device, "snapshots"), snaps.shape[1],
                 snaps.shape[0], snaps.shape[1])

    def compute_basis(self):
        call("pmx_pod_compute_basis", self.h)

    def info(self):
        n, m, r, v = c_int(), c_int(), c_int(), c_int()
        call("pmx_pod_get_info", self.h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(r), ctypes.byref(v))
        return n.value, m.value, r.value, v.value

    def get_pod_eigs(self):
        _, m, _, _ = self.info()
        out = np.zeros(m)
        call("pmx_pod_get_eigs", self.h, out.ctypes.data_as(ctypes.POINTER(c_double)))
        return out

    def get_basis(self):
        """CUDA tensor (rank, length): row k = mode k"""
        import torch
        n, _, r, _ = self.info()
        out = torch.zeros((max(r, 1), n), dtype=torch.float64, device="cuda:%d" % self.ctx.device)
        torch.cuda.synchronize()   # the fill on torch's stream ends before the copy on the context's
        call("pmx_pod_get_basis", self.h, _ptr(out, out.numel(), self.ctx.device, "basis"))
        return out[:r]

    def get_snaps_avr(self):
        import torch
        n, _, _, _ = self.info()
        out = torch.zeros(n, dtype=torch.float64, device="cuda:%d" % self.ctx.device)
        torch.cuda.synchronize()
        call("pmx_pod_get_mean", self.h, _ptr(out, n, self.ctx.device, "mean"))
        return out

    def write_pod(self, pod_directory, basename):
        call("pmx_pod_write", self.h, pod_directory.encode(), basename.encode())

    def __del__(self, _d=_destroy):
        if getattr(self, "h", None) and self.h.value:
            _d("pmx_pod_destroy", self.h)
            self.h = None


NSE_EXAMPLES = {"u2_p1": 9, "u0_p0": 13}
# pressure spaces: Q1 (continuous, velocity space 2) and discontinuous P1
# (D_P1_3D_H_A, velocity space 22), as NavierStokes.C:get_velocity_pressure_orders pairs them
NSE_PRESSURE = {"q1": 1, "p1disc": -110}
NSE_BLOCK = {"A": 0, "B": 1, "B1": 1, "B2": 2, "B3": 3, "combined": 4, "BT": 5, "B1T": 5, "B2T": 6, "B3T": 7}


class NSESystem:
    """One level of NavierStokes<3> with Q2 velocity and Q1 or discontinuous P1
    pressure on hexahedra and the NSTYPE-1 BlockFEMatrix (pmx_nse_*):
    assemble, apply, structures and values, interpolation of the example,
    errors, pressure projection.  `example` fixes the velocity boundary
    conditions ("u2_p1": Dirichlet everywhere; "u0_p0": Neumann on the
    interior of the face x = 1)."""

    def __init__(self, ctx, domain, level, pressure="q1", example="u2_p1"):
        self.ctx = ctx
        self.domain = domain
        self.example = example
        self.h = c_void_p()
        call("pmx_nse_create_pair", ctx.h, domain.h, level, NSE_PRESSURE[pressure], NSE_EXAMPLES[example],
             ctypes.byref(self.h))
        inf = NseInfo()
        call("pmx_nse_get_info", self.h, ctypes.byref(inf))
        self.info = inf
        self.n_u, self.n_u_active, self.n_p = inf.n_u, inf.n_u_active, inf.n_p
        self.n = inf.n_combined
        self.n_cells = inf.n_cells
        self.n_p_cell = inf.n_p_cell
        self.enclosed = bool(inf.enclosed)

    def _v(self, t, what):
        return _ptr(t, self.n, self.ctx.device, what)

    def zeros(self):
        import torch
        return torch.zeros(self.n, dtype=torch.float64, device=torch.device("cuda", self.ctx.device))

    def set_velocity_sell_rows(self, min_rows):
        """rows from which the velocity block runs on its SELL copy (default 131,072)"""
        call("pmx_nse_set_velocity_sell_rows", self.h, min_rows)

    def assemble(self, example=None, nu=1.0, w=None, rhs=None):
        import torch
        torch.cuda.synchronize()
        p = NseParams(NSE_EXAMPLES[example or self.example], nu)
        call("pmx_nse_assemble", self.ctx.h, self.h, ctypes.byref(p),
             _ptr(w, 3 * self.n_u, self.ctx.device, "w") if w is not None else None,
             self._v(rhs, "rhs") if rhs is not None else None)
        self.ctx.synchronize()

    def apply(self, x, y):
        import torch
        torch.cuda.synchronize()
        call("pmx_nse_apply", self.ctx.h, self.h, self._v(x, "x"), self._v(y, "y"))
        self.ctx.synchronize()

    def apply_scaled_add(self, x, y, a=1.0):
        import torch
        torch.cuda.synchronize()
        call("pmx_nse_apply_scaled_add", self.ctx.h, self.h, self._v(x, "x"), self._v(y, "y"), a)
        self.ctx.synchronize()

    def structure(self, which):
        nnz = c_int()
        call("pmx_nse_get_structure", self.h, NSE_BLOCK[which], ctypes.byref(nnz), None, None)
        n_rows = {"A": self.n_u, "B": self.n_p, "BT": self.n_u, "combined": self.n}[which]
        rp = np.zeros(n_rows + 1, np.int32)
        col = np.zeros(nnz.value, np.int32)
        P = ctypes.POINTER(c_int)
        call("pmx_nse_get_structure", self.h, NSE_BLOCK[which], ctypes.byref(nnz), rp.ctypes.data_as(P),
             col.ctypes.data_as(P))
        return rp, col

    def values(self, which):
        base = "B" if which in ("B1", "B2", "B3") else "BT" if which in ("B1T", "B2T", "B3T") else which
        rp, col = self.structure(base)
        out = np.zeros(len(col))
        call("pmx_nse_get_values", self.ctx.h, self.h, NSE_BLOCK[which], out.ctypes.data_as(ctypes.POINTER(c_double)))
        return out

    def global_numbers(self):
        gv = np.zeros(self.n_cells * 27, np.int32)
        gp = np.zeros(self.n_cells * self.n_p_cell, np.int32)
        P = ctypes.POINTER(c_int)
        call("pmx_nse_get_global_numbers", self.h, gv.ctypes.data_as(P), gp.ctypes.data_as(P))
        return gv.reshape(-1, 27), gp.reshape(-1, self.n_p_cell)

    def interpolate(self, x, example=None):
        import torch
        torch.cuda.synchronize()
        call("pmx_nse_interpolate", self.ctx.h, self.h, NSE_EXAMPLES[example or self.example], self._v(x, "x"))
        self.ctx.synchronize()

    def project_pressure(self, x):
        import torch
        torch.cuda.synchronize()
        call("pmx_nse_project_pressure", self.ctx.h, self.h, self._v(x, "x"))
        self.ctx.synchronize()

    def errors(self, x, example=None):
        import torch
        torch.cuda.synchronize()
        out = (c_double * 5)()
        call("pmx_nse_errors", self.ctx.h, self.h, NSE_EXAMPLES[example or self.example], self._v(x, "x"), out)
        return list(out)

    def restrict_velocity(self, fine, u_fine, u_coarse):
        """GridTransfer::RestrictFunction of [u1|u2|u3] from the next finer level."""
        import torch
        torch.cuda.synchronize()
        call("pmx_nse_restrict_velocity", self.ctx.h, self.h, fine.h, _ptr(u_fine, 3 * fine.n_u, self.ctx.device, "u_fine"),
             _ptr(u_coarse, 3 * self.n_u, self.ctx.device, "u_coarse"))
        self.ctx.synchronize()

    def __del__(self, _d=_destroy):
        if getattr(self, "h", None) and self.h.value:
            _d("pmx_nse_destroy", self.h)
            self.h = None


class NavierStokes:
    """NavierStokes<3> (src/System/NavierStokes.C) for NSTYPE 1, Q2/Q1 or
    Q2/P1disc (pressure="p1disc") on the unit cube: one NSESystem per
    multigrid level (finest last), the nonlinear loop of the reference's
    driver (test/nse3d_solvers_test.c++:107-126):

        ns.assemble_linear_terms(); ns.stop_it(0)
        for k in 1, 2, ...: ns.solve(); ns.assemble_nonlinear_term(); if ns.stop_it(k): break

    Parameters are the reference's database keys nonlinloop_epsilon,
    nonlinloop_maxit, nonlinloop_minit, nonlinloop_slowfactor,
    nonlinloop_damping_factor, nonlinloop_scale_epsilon_with_size,
    nonlinloop_residual_relative_to_rhs (defaults of
    src/General/ParameterDatabase.C:1294-1330)."""

    def __init__(self, ctx, domain, levels, solver_params, nu=1.0, example="u2_p1", pressure="q1", **db):
        import torch
        self.ctx = ctx
        self.example = example
        self.nu = nu
        self.systems = [NSESystem(ctx, domain, l, pressure, example) for l in levels]
        self.solution = [s.zeros() for s in self.systems]
        f = self.systems[-1]
        self.rhs = f.zeros()
        self.defect = f.zeros()
        self.solver = Solver(ctx, solver_params)
        self.db = {"nonlinloop_epsilon": 1e-10, "nonlinloop_maxit": 100, "nonlinloop_minit": 1,
                   "nonlinloop_slowfactor": 1e10, "nonlinloop_damping_factor": 1.0,
                   "nonlinloop_scale_epsilon_with_size": False, "nonlinloop_residual_relative_to_rhs": False}
        for k, v in db.items():
            if k not in self.db:
                raise KeyError(k)
            self.db[k] = v
        self.residuals = []           # (impuls^2, mass^2, full) per stop_it call
        self.linear_iterations = []
        self._torch = torch

    def _assemble_all(self):
        last = len(self.systems) - 1
        for k, (s, x) in enumerate(zip(self.systems, self.solution)):
            s.assemble(self.example, self.nu, x[:3 * s.n_u], self.rhs if k == last else None)

    def assemble_linear_terms(self):
        """All blocks with the current velocity as convection field (zero at
        the start); Dirichlet values copied from the rhs into the solution."""
        self._assemble_all()
        f, x = self.systems[-1], self.solution[-1]
        nu, na = f.n_u, f.n_u_active
        for k in range(3):
            x[k * nu + na:(k + 1) * nu] = self.rhs[k * nu + na:(k + 1) * nu]

    def assemble_nonlinear_term(self):
        """Velocity restricted to every coarser level (RestrictFunctionRepeatedly),
        then the blocks of every level with its own velocity as convection."""
        for l in range(len(self.systems) - 2, -1, -1):
            c, fsys = self.systems[l], self.systems[l + 1]
            c.restrict_velocity(fsys, self.solution[l + 1], self.solution[l])
        self._assemble_all()

    def compute_residuals(self):
        f = self.systems[-1]
        nu = f.n_u
        self.defect.copy_(self.rhs)
        f.apply_scaled_add(self.solution[-1], self.defect, -1.0)
        if f.enclosed:   # BlockFEMatrix::pressure_projection_enabled
            f.project_pressure(self.defect)
        imp = vector_dot(self.ctx, self.defect[:3 * nu], self.defect[:3 * nu])
        mass = vector_dot(self.ctx, self.defect[3 * nu:], self.defect[3 * nu:])
        self.residuals.append((imp, mass, (imp + mass) ** 0.5))

    def get_full_residual(self):
        return self.residuals[-1][2]

    def stop_it(self, k):
        """NavierStokes<d>::stop_it (src/System/NavierStokes.C:750-818)."""
        self.compute_residuals()
        res = self.get_full_residual()
        if k == 0:
            self.initial_residual = res
            self.initial_rhs_norm = vector_norm(self.ctx, self.rhs)
        if k < self.db["nonlinloop_minit"]:
            return False
        old = self.residuals[max(0, len(self.residuals) - 10)][2]
        slow = res >= self.db["nonlinloop_slowfactor"] * old
        limit = self.db["nonlinloop_epsilon"]
        if self.db["nonlinloop_scale_epsilon_with_size"]:
            limit *= self.systems[-1].n ** 0.5
        if self.db["nonlinloop_residual_relative_to_rhs"]:
            limit *= self.initial_rhs_norm
        return res <= limit or k == self.db["nonlinloop_maxit"] or slow

    def solve(self):
        """NavierStokes<d>::solve (src/System/NavierStokes.C:880-915): update the
        solver with the current matrices, solve from the current solution,
        damping, pressure projection."""
        f, x = self.systems[-1], self.solution[-1]
        damping = self.db["nonlinloop_damping_factor"]
        old = x.clone() if damping != 1.0 else None
        self.solver.update_saddle(self.systems)
        rep = self.solver.solve(self.rhs, x)
        self.linear_iterations.append(rep.iterations)
        if damping != 1.0:
            x.mul_(damping).add_(old, alpha=1.0 - damping)
        if f.enclosed:
            f.project_pressure(x)
        return rep
