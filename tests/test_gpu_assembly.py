"""The two Q2 hexahedron assembly routes of pmx_assemble_cd through the C ABI:
the element GEMM (per-cell coefficients times reference matrices on the
matrix cores) with the row gather, FP64 and FP32, against the quadrature
kernel and the oracle.  FP64 routes agree to 1e-13 of the largest entry (the
same sums in another order); the FP32 element matrices to 1e-6 of each row's
largest entry (FP32 operands, 32 terms accumulated in FP32 per element entry,
up to 8 element entries per matrix entry; measured below 1e-7)."""
import numpy as np
import pytest
import torch

from mymoonmd_amd import parmoon as pm
from mymoonmd_amd._lib import PmxError
from oracle import oracle as O

pytestmark = pytest.mark.gpu

B_CD3D = (1 / 14 ** 0.5, 2 / 14 ** 0.5, 3 / 14 ** 0.5)

CASES = [
    (2, dict(example="exp_sin_3d", disc="supg", eps=1e-3, b=B_CD3D, hexa_quad_rule=2)),
    (3, dict(example="exp_sin_3d", disc="supg", eps=1e-3, b=B_CD3D, hexa_quad_rule=2)),
    (3, dict(example="p2_3d", disc="galerkin", eps=1.0, b=(4, 3, 2), c=1.0, hexa_quad_rule=0)),
    (2, dict(example="exp_sin_3d", disc="supg", eps=1e-2, b=(0.3, -0.5, 0.8), c=1.5, hexa_quad_rule=0)),
    (4, dict(example="exp_sin_3d", disc="supg", eps=1e-3, b=B_CD3D, hexa_quad_rule=2)),
]


def assemble(ctx, s, **kw):
    A = pm.FEMatrix(ctx, s)
    f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
    pm.assemble_cd(ctx, s, pm.cd_params(**kw), A, f)
    torch.cuda.synchronize()
    return A.get_entries(), f.cpu().numpy()


def space(ctx, nref):
    return pm.TFESpace(pm.TDomain("Default_UnitCube_Hexa", nref), nref, 2, ctx)


@pytest.mark.parametrize("nref,kw", CASES)
def test_element_gemm_matches_quadrature_kernel(ctx, nref, kw):
    print("case", nref, kw, flush=True)
    s = space(ctx, nref)
    v0, f0 = assemble(ctx, s, **kw)
    v1, f1 = assemble(ctx, s, assembly_variant=1, **kw)
    rp, _ = pm.FEMatrix(ctx, s).get_structure()
    act = rp[s.active_bound]            # entries of the active rows (Dirichlet rows hold 1)
    scale = np.abs(v1[:act]).max()
    assert np.abs(v0 - v1).max() <= 1e-13 * scale
    assert np.abs(f0 - f1).max() <= 1e-13 * max(1.0, np.abs(f1).max())


@pytest.mark.parametrize("nref,kw", CASES[:3])
def test_element_gemm_matches_oracle(ctx, nref, kw):
    print("case", nref, kw, flush=True)
    s = space(ctx, nref)
    v0, f0 = assemble(ctx, s, **kw)
    sp, _ = O.build_space(2, nref, nref, 2)
    ex = {"exp_sin_3d": 5, "p2_3d": 4}[kw["example"]]
    p = O.cd_params(example=ex, disc={"galerkin": 0, "supg": 1}[kw["disc"]], eps=kw["eps"],
                    b=list(kw["b"]), c=kw.get("c", 0.0), hexa_rule=kw["hexa_quad_rule"])
    val, rhs = O.assemble(sp, p)
    assert np.abs(v0 - val).max() <= 1e-12 * np.abs(val).max()
    assert np.abs(f0 - rhs).max() <= 1e-12 * max(1.0, np.abs(rhs).max())


@pytest.mark.parametrize("nref,kw", [CASES[1], CASES[2], CASES[4]])
def test_fp32_element_matrices(ctx, nref, kw):
    """FP32 matrix cores: every entry within 1e-6 of its row's largest entry
    of the FP64 route and the median error at rounding level, the load vector
    (FP64 in both) equal, Dirichlet rows unit."""
    print("case", nref, kw, flush=True)
    s = space(ctx, nref)
    A = pm.FEMatrix(ctx, s)
    rp, _ = A.get_structure()
    v64, f64 = assemble(ctx, s, **kw)
    v32, f32 = assemble(ctx, s, fp32_assembly=True, **kw)
    rmax = np.repeat(np.maximum.reduceat(np.abs(v64), rp[:-1]), np.diff(rp))
    rel = np.abs(v32 - v64) / rmax
    print("fp32 entry error max %.2e median %.2e" % (rel.max(), np.median(rel)), flush=True)
    assert rel.max() <= 1e-6
    assert 0.0 < np.median(rel) <= 1e-7   # the FP32 route really ran
    assert np.array_equal(f32, f64)
    assert np.all(v32[rp[s.active_bound:-1]] == 1.0)


def test_assembly_adds_into_the_matrix(ctx):
    """pmx_assemble_cd adds (the caller zeroes first): two calls give twice
    one call's entries on the active rows."""
    kw = CASES[1][1]
    s = space(ctx, 3)
    A = pm.FEMatrix(ctx, s)
    f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
    prm = pm.cd_params(**kw)
    pm.assemble_cd(ctx, s, prm, A, f)
    once = A.get_entries().copy()
    f_once = f.cpu().numpy().copy()
    pm.assemble_cd(ctx, s, prm, A, f)
    twice = A.get_entries()
    rp, _ = A.get_structure()
    act = rp[s.active_bound]
    assert np.abs(twice[:act] - 2 * once[:act]).max() <= 1e-14 * np.abs(once).max()
    assert np.all(twice[act:] == 1.0)
    fa = f.cpu().numpy()[: s.active_bound]
    assert np.abs(fa - 2 * f_once[: s.active_bound]).max() <= 1e-14 * np.abs(f_once).max()


def test_fp32_assembly_needs_the_gemm_route(ctx):
    s = space(ctx, 2)
    A = pm.FEMatrix(ctx, s)
    f = torch.zeros(s.n_dof, dtype=torch.float64, device="cuda")
    for kw in (dict(disc="gls", assembly_variant=0), dict(disc="supg", assembly_variant=1)):
        with pytest.raises(PmxError):
            pm.assemble_cd(ctx, s, pm.cd_params(example="exp_sin_3d", eps=1e-3, b=B_CD3D, fp32_assembly=True,
                                                 **kw), A, f)
    with pytest.raises(PmxError):
        pm.assemble_cd(ctx, s, pm.cd_params(example="exp_sin_3d", assembly_variant=2), A, f)


@pytest.mark.parametrize("variant,fp32", [(0, False), (0, True), (1, False)])
def test_overwrite_sets_matrix_and_rhs(ctx, variant, fp32):
    """overwrite = 1 on a matrix and rhs full of other values gives exactly
    what zeroing them and adding gives, on both routes (FP32 element matrices
    exist on the GEMM route only)."""
    kw = dict(CASES[1][1], assembly_variant=variant, fp32_assembly=fp32)
    s = space(ctx, 3)
    ref_v, ref_f = assemble(ctx, s, **kw)
    A = pm.FEMatrix(ctx, s)
    rng = np.random.default_rng(3)
    A.set_entries(rng.standard_normal(len(ref_v)))
    f = torch.as_tensor(rng.standard_normal(s.n_dof)).cuda()
    pm.assemble_cd(ctx, s, pm.cd_params(overwrite=True, **kw), A, f)
    assert np.array_equal(A.get_entries(), ref_v)
    assert np.array_equal(f.cpu().numpy(), ref_f)
