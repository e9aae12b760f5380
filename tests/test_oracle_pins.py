"""Pins of the oracle (and the product's host set-up) by facts that the
reference's source tables or the mathematics fix, independently of the
oracle's own restatement:

* the DOF numbering of the refined unit cube (Q1, Q2) against a fixture
  derived from the reference's refinement, boundary and nodal-point tables
  (tests/golden/make_hexa_numbering_fixture.py);
* the assembled Galerkin and SUPG operators against closed forms: for p in the
  FE space and an interior DOF i, (A p)_i = int (-eps lap p + b.grad p + c p)
  phi_i and (A_SUPG - A_Gal) p)_i = -delta int b.grad(-eps lap p + b.grad p)
  phi_i (integration by parts, phi_i in H^1_0); with the integrands linear
  and phi_i symmetric about its node these are a value at the node times
  int phi_i, so delta of Compute_SDFEM_delta (ConvDiff.C:273-345) is read off
  the matrix;
* the quadrature rule, through the load vector of a non-polynomial f against
  tensor Gauss-Legendre rules built here;
* the Jacobi damping of the benchmark: lambda_max(D^-1 A) of the Q2 SUPG
  operator, and what damping 0.8 and 0.7 do to the solve;
* one V(2,2) cycle against the dense composition of its factors
  S = I - w D^-1 A, R = P^T, P and A_0^-1 built here in numpy.
All of it runs on the CPU."""
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
B_CD3D = (1 / 14 ** 0.5, 2 / 14 ** 0.5, 3 / 14 ** 0.5)


# ------------------------------------------------------------ numbering
@pytest.fixture(scope="module")
def numbering():
    with open(os.path.join(ROOT, "tests", "golden", "hexa_numbering.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("level", [1, 2])
@pytest.mark.parametrize("order", [1, 2])
def test_hexa_numbering_matches_reference_tables(numbering, level, order):
    """GlobalNumbers of the oracle and of the product's host space (no
    device needed) equal the fixture; so do the mesh vertices."""
    from mymoonmd_amd import parmoon as pm
    fx = numbering["level%d_Q%d" % (level, order)]
    g_fix = np.array(fx["global_numbers"]).reshape(-1, 27 if order == 2 else 8)
    sp, info = O.build_space(2, 2, level, order)
    assert np.array_equal(sp.gnum, g_fix)
    assert sp.n_dof == fx["n_dof"] and sp.active == fx["n_inner"]
    dom = pm.TDomain("Default_UnitCube_Hexa", 2)
    s = pm.TFESpace(dom, level, order, None)
    assert np.array_equal(s.global_numbers(), g_fix)
    assert np.array_equal(dom.vertices(level).reshape(-1), np.array(numbering["level%d_vertices" % level]))


# ------------------------------------------------- closed-form operators
def phi_integral(x, h, order):
    """int phi_i for the interior node x of a uniform hexahedral mesh with cell
    size h: per direction h for Q1; h/3 at a vertex coordinate and 2h/3 at a
    midpoint coordinate for Q2 (1D P2: 1/6, 2/3, 1/6 of the interval)"""
    if order == 1:
        return h ** 3
    t = np.round(x / (h / 2)).astype(int)
    per = np.where(t % 2 == 0, h / 3, 2 * h / 3)
    return np.prod(per, axis=1)


def sdfem_delta(h_diam, eps, b, degree=1.0):
    """SDFEM_TYPE 2 with INTERNAL_POLYNOMIAL_DEGREE `degree` and DELTA0 = 1"""
    nb = math.sqrt(sum(v * v for v in b))
    nu = 1.0 / degree
    alpha = nu * nb * h_diam / (2 * eps)
    return nu * h_diam * (1 / math.tanh(alpha) - 1 / alpha) / (2 * nb)


def operator_setup(nref, order, disc, eps, b, c=0.0, rule=2, poly_degree=1.0):
    sp, _ = O.build_space(2, nref, nref, order)
    val, _ = O.assemble(sp, O.cd_params(example=0, disc=disc, eps=eps, b=b, c=c, hexa_rule=rule,
                                        poly_degree=poly_degree))
    x = O.nodal_coords(sp)
    return sp, val, x


def apply(sp, val, v):
    y = np.zeros(sp.n_dof)
    O.apply_scaled_add(sp.n_dof, sp.active, sp.rp, sp.col, val, v, y, 1.0)
    return y[:sp.active]


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("rule", [0, 2])
def test_galerkin_operator_closed_form(order, rule):
    """Galerkin rows: p = x^2 (c = 0) gives (-2 eps + 2 b1 x_i) int phi_i;
    p = x + 2y - z with c = 0.7 gives (b1 + 2 b2 - b3 + c p(x_i)) int phi_i.
    A convection term with phi_i and phi_j exchanged flips the sign of the
    b terms; a dropped diffusion, convection or reaction term, or a wrong
    Jacobian, changes the values."""
    nref, h = 2, 0.25
    eps, b = 0.3, (0.9, -0.4, 0.7)
    if order == 2:
        sp, val, x = operator_setup(nref, order, 0, eps, b, rule=rule)
        r = apply(sp, val, x[:, 0] ** 2)
        xa = x[:sp.active]
        ref = (-2 * eps + 2 * b[0] * xa[:, 0]) * phi_integral(xa, h, order)
        assert np.max(np.abs(r - ref)) <= 1e-13 * np.max(np.abs(ref))
    c = 0.7
    sp, val, x = operator_setup(nref, order, 0, eps, b, c=c, rule=rule)
    p = x[:, 0] + 2 * x[:, 1] - x[:, 2]
    r = apply(sp, val, p)
    xa = x[:sp.active]
    ref = (b[0] + 2 * b[1] - b[2] + c * p[:sp.active]) * phi_integral(xa, h, order)
    assert np.max(np.abs(r - ref)) <= 1e-13 * np.max(np.abs(ref))


@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("eps,b", [(1e-3, B_CD3D), (0.05, (1.0, 2.0, 3.0)), (0.2, (-0.5, 0.3, 0.8))])
def test_supg_delta_from_the_operator(order, eps, b):
    """(A_SUPG - A_Gal) p on the active rows for p = x^2 y (Q2) or p = xyz
    (Q1; the Laplacian of a trilinear function vanishes) equals
    -delta (b.grad g)(x_i) int phi_i with g = -eps lap p + b.grad p:
        Q2: b.grad g = 2 b1^2 y + 4 b1 b2 x - 2 eps b2,
        Q1: b.grad g = 2 (b2 b3 x + b1 b3 y + b1 b2 z).
    The delta read off every row equals SDFEM_TYPE 2 with h_K the cell
    diameter (THexahedron::GetDiameter, src/Geometry/Hexahedron.C:78-104; the
    default CELL_MEASURE 0, Database.C:2671) and polynomial degree 1 (the
    default INTERNAL_POLYNOMIAL_DEGREE, Database.C:2984; the consistency check
    that would reset both, check_parameters_consistency_CD, has no caller).
    The edge length for h_K, degree 2, or |b|_1 for |b| give deltas that this
    test tells apart; a SUPG term with test and trial function exchanged, or
    without its -eps lap phi_j part, changes the Q2 rows."""
    nref, h = 2, 0.25
    spg, vg, x = operator_setup(nref, order, 0, eps, b)
    sps, vs, _ = operator_setup(nref, order, 1, eps, b)
    assert np.array_equal(spg.col, sps.col)
    X, Y, Z = x[:, 0], x[:, 1], x[:, 2]
    p = X * X * Y if order == 2 else X * Y * Z
    d = apply(sps, vs, p) - apply(spg, vg, p)
    xa = x[:spg.active]
    if order == 2:
        bg = 2 * b[0] ** 2 * xa[:, 1] + 4 * b[0] * b[1] * xa[:, 0] - 2 * eps * b[1]
    else:
        bg = 2 * (b[1] * b[2] * xa[:, 0] + b[0] * b[2] * xa[:, 1] + b[0] * b[1] * xa[:, 2])
    w = -bg * phi_integral(xa, h, order)
    ok = np.abs(w) > 1e-3 * np.max(np.abs(w))
    delta_rows = d[ok] / w[ok]
    delta = sdfem_delta(h * math.sqrt(3), eps, b)
    assert np.max(np.abs(delta_rows - delta)) <= 1e-10 * delta
    # the alternatives a misreading would give are distinguishable
    for alt in (sdfem_delta(h, eps, b), sdfem_delta(h * math.sqrt(3), eps, b, degree=2.0),
                sdfem_delta(h * math.sqrt(3), eps, [sum(abs(v) for v in b), 0, 0])):
        assert abs(alt - delta) > 1e-3 * delta
    # planted mistake: the oracle run with polynomial degree 2 fails the check
    _, vw, _ = operator_setup(nref, order, 1, eps, b, poly_degree=2.0)
    dw = apply(sps, vw, p) - apply(spg, vg, p)
    assert np.max(np.abs(dw[ok] / w[ok] - delta)) > 1e-3 * delta


def gauss_1d(n):
    xg, wg = np.polynomial.legendre.leggauss(n)
    return (xg + 1) / 2, wg / 2      # on [0, 1]


def basis_1d_moment(node, h, g, n):
    """int over the support of the 1D P2 basis of nodal index `node` (node
    spacing h/2) of g(x) * basis(x), each interval by an n-point Gauss rule"""
    t, w = gauss_1d(n)
    total = 0.0
    for k0 in ((node // 2 - 1, node // 2) if node % 2 == 0 else (node // 2,)):
        a = k0 * h
        if a < -1e-14 or a > 1 - h + 1e-14:
            continue
        xs = a + h * t
        local = (node - 2 * k0) / 2.0          # 0, 0.5 or 1 within the interval
        if local == 0.0:
            phi = (1 - t) * (1 - 2 * t)
        elif local == 1.0:
            phi = t * (2 * t - 1)
        else:
            phi = 4 * t * (1 - t)
        total += h * np.sum(w * g(xs) * phi)
    return total


@pytest.mark.parametrize("rule,n_gauss", [(2, 3), (0, 4)])
def test_load_vector_quadrature(rule, n_gauss):
    """The Galerkin load vector of u = exp(-x) sin(pi y) sin(pi z)
    (f = -eps lap u + b.grad u) on Q2 equals tensor Gauss-Legendre sums with
    n_gauss points per direction (INTERNAL_QUAD_RULE 2 -> Gauss3; the Q2
    default -> Gauss4), built here from numpy's Gauss-Legendre nodes and the
    1D P2 basis; f is a sum of separable terms, so the 3D rule factorises.
    The other rule differs by far more than the tolerance."""
    nref, h = 2, 0.25
    eps, b = 1e-3, B_CD3D
    sp, _ = O.build_space(2, nref, nref, 2)
    _, rhs = O.assemble(sp, O.cd_params(example=5, disc=0, eps=eps, b=b, hexa_rule=rule))
    x = O.nodal_coords(sp)
    pi = math.pi
    ex = lambda s: np.exp(-s)                    # noqa: E731
    sn = lambda s: np.sin(pi * s)                # noqa: E731
    cs = lambda s: pi * np.cos(pi * s)           # noqa: E731
    # f = (-eps (1 - 2 pi^2) - b1) e^-x s(y) s(z) + b2 e^-x c(y) s(z) + b3 e^-x s(y) c(z)
    terms = [(-eps * (1 - 2 * pi * pi) - b[0], ex, sn, sn), (b[1], ex, cs, sn), (b[2], ex, sn, cs)]

    def load(i, n):
        nodes = np.round(x[i] / (h / 2)).astype(int)
        return sum(cf * basis_1d_moment(nodes[0], h, gx, n) * basis_1d_moment(nodes[1], h, gy, n) *
                   basis_1d_moment(nodes[2], h, gz, n) for cf, gx, gy, gz in terms)

    rows = range(0, sp.active, 7)
    mine = np.array([load(i, n_gauss) for i in rows])
    other = np.array([load(i, 7 - n_gauss) for i in rows])
    got = rhs[list(rows)]
    scale = np.max(np.abs(mine))
    assert np.max(np.abs(got - mine)) <= 1e-13 * scale
    assert np.max(np.abs(got - other)) > 1e-8 * scale


# ----------------------------------------------- the Jacobi damping claim
def dense_levels(nref, n_levels, eps=1e-3, b=B_CD3D, omega_example=5):
    spaces = [O.build_space(2, nref, l, 2)[0] for l in range(nref - n_levels + 1, nref + 1)]
    p = O.cd_params(example=omega_example, disc=1, eps=eps, b=b, hexa_rule=2)
    mats, rhss = [], []
    for sp in spaces:
        v, r = O.assemble(sp, p)
        A = np.zeros((sp.n_dof, sp.n_dof))
        for i in range(sp.n_dof):
            A[i, sp.col[sp.rp[i]:sp.rp[i + 1]]] = v[sp.rp[i]:sp.rp[i + 1]]
        mats.append((v, A))
        rhss.append(r)
    return spaces, mats, rhss


def test_jacobi_damping_08_amplifies_on_the_supg_operator():
    """The benchmark's SUPG Q2 operator (eps 1e-3, b = (1,2,3)/sqrt(14),
    Gauss3) at 8^3 cells: lambda_max(D^-1 A) on the active block is above
    2.5, so |1 - 0.8 lambda| > 1 and damped Jacobi with 0.8 amplifies that
    mode, while 0.7 keeps |1 - 0.7 lambda| < 1.15.  FGMRES(30) + MG V(2,2)
    on 3 levels: with 0.8 the residual is still above 1e-7 after 150
    iterations; with 0.7 it reaches 1e-8 in fewer than 100.  (At 64^3 the
    stall with 0.8 sits near 1e-5.)"""
    spaces, mats, rhss = dense_levels(3, 3)
    sp = spaces[-1]
    A = mats[-1][1][:sp.active, :sp.active]
    ev = np.linalg.eigvals(A / np.diag(A)[:, None])
    lam = np.max(ev.real)
    assert lam > 2.5, lam
    assert abs(1 - 0.8 * lam) > 1.0 and abs(1 - 0.7 * lam) < 1.15
    b = rhss[-1]
    x0 = np.zeros(sp.n_dof)
    x0[sp.active:] = b[sp.active:]
    res = {}
    for omega in (0.8, 0.7):
        op = O.solver_params(iterative_solver_type=0, preconditioner=2, gmres_restart=30, residual_tolerance=0.0,
                             residual_reduction=1e-8, max_n_iterations=150, mg_cycle=0, mg_smoother_coarse=1,
                             mg_n_pre_smooth=2, mg_n_post_smooth=2, jacobi_damp=omega, mg_coarse_residual=0.1,
                             mg_coarse_max_n_iterations=10)
        x, it, hist, st = O.solve(op, spaces, [m[0] for m in mats], b, x0)
        res[omega] = (it, hist[-1] / hist[0])
    assert res[0.8][1] > 1e-7 and res[0.8][0] == 150, res
    assert res[0.7][1] <= 1e-8 and res[0.7][0] < 100, res


def test_v_cycle_is_the_composition_of_its_factors():
    """One V(2,2) cycle (zero start, damped Jacobi 0.7, exact coarse solve,
    correction damping 1) on 2 levels (2^3 -> 4^3 cells) equals
        z = S^2 (I + P A0^-1 R (I - A S^2)...) applied to v,
    written here as the dense sequence x1 = S-steps from 0, d = v - A x1
    restricted with R = P^T (Dirichlet rows of the coarse defect zeroed),
    e = A0^-1 d_c, x2 = x1 + P e on the active rows, then 2 more steps;
    S(x) = x + w D^-1 (v - A x) on active rows with the Dirichlet values of v
    kept.  P is the interpolation matrix of the nested Q2 spaces, checked to
    reproduce quadratics exactly."""
    nref = 2
    spaces, mats, rhss = dense_levels(nref, 2)
    cs, fs = spaces
    Ac, Af = mats[0][1], mats[1][1]
    nc, nf = cs.n_dof, fs.n_dof
    P = np.zeros((nf, nc))
    for j in range(nc):
        e = np.zeros(nc)
        e[j] = 1.0
        P[:, j] = O.prolongate(cs, fs, e)
    xc, xf = O.nodal_coords(cs), O.nodal_coords(fs)
    q = lambda x: x[:, 0] ** 2 + x[:, 1] * x[:, 2] - 2 * x[:, 2] ** 2 + 1   # noqa: E731
    assert np.max(np.abs(P @ q(xc) - q(xf))) < 1e-13
    R = P.T
    w = 0.7
    rng = np.random.default_rng(4)
    v = rng.standard_normal(nf)

    def smooth(x):
        xn = x.copy()
        r = v - Af @ x
        xn[:fs.active] = x[:fs.active] + w * r[:fs.active] / np.diag(Af)[:fs.active]
        xn[fs.active:] = v[fs.active:]
        return xn

    x = np.zeros(nf)
    x[fs.active:] = v[fs.active:]
    x = smooth(smooth(x))
    d = v - Af @ x
    d[fs.active:] = 0.0
    dc = R @ d
    dc[cs.active:] = 0.0
    e = np.linalg.solve(Ac, dc)
    x[:fs.active] += (P @ e)[:fs.active]
    z = smooth(smooth(x))
    op = O.solver_params(preconditioner=2, mg_cycle=0, mg_smoother_coarse=1, mg_n_pre_smooth=2,
                         mg_n_post_smooth=2, jacobi_damp=w, mg_correction_damp=1.0)
    zo = O.mg_apply(op, spaces, [m[0] for m in mats], v)
    assert np.max(np.abs(zo - z)) <= 1e-11 * np.max(np.abs(z))
