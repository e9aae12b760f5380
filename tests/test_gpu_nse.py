"""Oseen / Navier-Stokes (Q2/Q1, NSTYPE 1) on the HIP path against the oracle:
numbering, patterns, assembled blocks, block operator, combined matrix,
interpolation, errors and the pressure projection."""
import numpy as np
import pytest
import torch

from mymoonmd_amd import parmoon as pm
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()


@pytest.fixture(scope="module", params=[1, 2])
def pair(ctx, request):
    nref = request.param
    d = pm.TDomain("Default_UnitCube_Hexa", nref)
    g = pm.NSESystem(ctx, d, nref)
    o = O.NseOracle(nref, nref)
    return nref, d, g, o


def test_nse_numbering_and_patterns(pair):
    nref, d, g, o = pair
    gv, gp = g.global_numbers()
    assert np.array_equal(gv, o.vs.gnum) and np.array_equal(gp, o.ps.gnum)
    rp, col = g.structure("A")
    assert np.array_equal(rp, o.vs.rp) and np.array_equal(col, o.vs.col)
    rp, col = g.structure("B")
    assert np.array_equal(rp, o.brp) and np.array_equal(col, o.bcol)


@pytest.mark.parametrize("nu,conv", [(1.0, False), (1e-2, True)])
def test_nse_assembly_apply_combined_match_oracle(ctx, pair, nu, conv):
    nref, d, g, o = pair
    print("case", nref, nu, conv, flush=True)
    rng = np.random.default_rng(7)
    w = rng.standard_normal(3 * o.n_u) if conv else None
    rhs = g.zeros()
    g.assemble("u2_p1", nu, dev(w) if conv else None, rhs)
    A, B, r = o.assemble(nu, w)
    ga = g.values("A")
    assert np.max(np.abs(ga - A)) <= 1e-12 * np.max(np.abs(A))
    for k in range(3):
        gb = g.values("B%d" % (k + 1))
        assert np.max(np.abs(gb - B[k])) <= 1e-12 * np.max(np.abs(B[k]))
    assert np.max(np.abs(rhs.cpu().numpy() - r)) <= 1e-12 * np.max(np.abs(r))
    x = rng.standard_normal(o.n)
    y = g.zeros()
    g.apply(dev(x), y)
    yo = o.apply(A, B, x)
    assert np.max(np.abs(y.cpu().numpy() - yo)) <= 1e-12 * np.max(np.abs(yo))
    K = o.combined(A, B)
    rp, col = g.structure("combined")
    import scipy.sparse as sp
    G = sp.csr_matrix((g.values("combined"), col, rp), shape=(o.n, o.n))
    assert abs(G - K).max() <= 1e-12 * abs(K).max()


def test_nse_interpolant_residual_errors_and_projection(ctx, pair):
    """The interpolant of test_u_2_p_1 solves the Oseen system with w = u; its
    errors vanish; the pressure projection removes a constant shift."""
    nref, d, g, o = pair
    x = g.zeros()
    g.interpolate(x)
    assert np.max(np.abs(x.cpu().numpy() - O.nse_interpolate(o))) < 1e-14
    rhs = g.zeros()
    g.assemble("u2_p1", 1e-2, x[:3 * g.n_u].clone(), rhs)
    r = rhs.clone()
    g.apply_scaled_add(x, r, -1.0)
    assert torch.max(torch.abs(r)).item() < 1e-13
    e = g.errors(x)
    assert e[0] < 1e-13 and e[1] < 1e-12 and e[3] < 1e-13 and e[4] < 1e-12, e
    x[3 * g.n_u:] += 0.37
    assert g.errors(x)[3] > 0.3
    g.project_pressure(x)
    assert g.errors(x)[3] < 1e-13
