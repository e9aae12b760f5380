"""Pins the oracle's restatement of SnapshotsCollector and POD (oracle/pod.py):
the reference's snaps_pod_rom_test golden values (mean of the snapshots and
POD eigenvalues of a TCD2D run with the L2 inner product), planted mistakes
that those values catch, and what the mathematics fixes (eigenvalues are the
squared singular values of L^T S for W = L L^T, the basis is W-orthonormal,
the discarded eigenvalues are the W-norm projection error)."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import pod as OP


def dense(space, vals):
    A = np.zeros((space.n_dof, space.n_dof))
    for i in range(space.n_dof):
        for k in range(space.rp[i], space.rp[i + 1]):
            A[i, space.col[k]] = vals[k]
    return A


def golden_run(g):
    """snaps_pod_rom_test.c++:148-186 on the oracle: TCD2D example 0 (eps 1,
    b = (1, -1), c = 1), UnitSquare refined twice, Q1, Galerkin, backward Euler;
    snapshots of the initial condition and of every steps_per_snap-th step."""
    sp, _ = O.build_space(0, g["n_refine"], g["n_refine"], 1)
    tau, k = g["time_step_length"], g["steps_per_snap"]
    p0 = O.cd_params(example=7, disc=0, eps=1.0, b=(1.0, -1.0, 0.0), c=1.0, time=0.0, theta1=1.0, tau=tau)
    snaps = [O.interpolate(sp, p0)]

    def on_step(step, t, u):
        if step % k == 0:
            snaps.append(u.copy())

    O.tcd_time_integration(sp, 7, 0, O.BACKWARD_EULER, tau, g["n_steps"], eps=1.0, b=(1.0, -1.0), c=1.0,
                           on_step=on_step)
    _, m, _ = O.assemble_tcd(sp, p0)
    return sp, np.array(snaps).T, dense(sp, m)


def test_pod_golden_snaps_pod_rom(golden):
    g = golden["snaps_pod_rom_test"]
    sp, S, M = golden_run(g)
    assert S.shape == (25, 3)
    res = OP.pod(S, M, fluctuations_only=True)
    assert np.max(np.abs(res["mean"] - np.array(g["mean"]))) < g["tolerance"]
    assert np.max(np.abs(res["eigs"][:2] - np.array(g["eigs"]))) < g["tolerance"]
    assert res["valid"] == 2 and res["rank"] == 2


def test_pod_golden_detects_planted_errors(golden):
    """The golden eigenvalues reject the euclidean inner product, a missing
    mean subtraction and a mass matrix without its Dirichlet rows."""
    g = golden["snaps_pod_rom_test"]
    sp, S, M = golden_run(g)
    tol = g["tolerance"]
    for res in (OP.pod(S, None), OP.pod(S, M, fluctuations_only=False)):
        assert np.max(np.abs(res["eigs"][:2] - np.array(g["eigs"]))) > 100 * tol
    M2 = M.copy()
    M2[sp.active:, :] = 0.0
    res = OP.pod(S, M2)
    # the Dirichlet fluctuations vanish here: this one must still pass, the
    # boundary rows do not carry the result
    assert np.max(np.abs(res["eigs"][:2] - np.array(g["eigs"]))) < tol
    # a wrong time-step count changes the snapshots
    g2 = dict(g, n_steps=9, steps_per_snap=3)
    _, S3, M3 = golden_run(g2)
    res = OP.pod(S3, M3)
    assert np.max(np.abs(res["eigs"][:2] - np.array(g["eigs"]))) > 100 * tol


@pytest.mark.parametrize("fluct", [True, False])
@pytest.mark.parametrize("m", [3, 7, 20])
def test_pod_closed_forms(fluct, m):
    rng = np.random.default_rng(10 + m)
    n = 60
    S = rng.standard_normal((n, m)) @ np.diag(np.logspace(0, -3, m))
    B = rng.standard_normal((n, n))
    W = B @ B.T / n + np.eye(n)
    res = OP.pod(S, W, fluctuations_only=fluct)
    Sc = S - res["mean"][:, None]
    if fluct:
        assert np.allclose(res["mean"], S.mean(axis=1), rtol=1e-14, atol=1e-15)
    else:
        assert np.all(res["mean"] == 0)
    L = np.linalg.cholesky(W)
    sv = np.linalg.svd(L.T @ Sc, compute_uv=False)
    k = min(len(sv), m)
    assert np.allclose(res["eigs"][:k], sv[:k] ** 2, rtol=1e-10, atol=1e-12 * sv[0] ** 2)
    Phi = res["basis"]
    r = res["rank"]
    assert np.allclose(Phi.T @ W @ Phi, np.eye(r), atol=1e-9)
    for q in (1, r // 2, r):
        P = Phi[:, :q]
        E = Sc - P @ (P.T @ W @ Sc)
        err = np.trace(E.T @ W @ E)
        assert abs(err - np.sum(res["eigs"][q:][res["eigs"][q:] > 0])) <= 1e-9 * np.sum(np.abs(res["eigs"])) + 1e-12
    # pod_rank truncates to the leading modes
    res2 = OP.pod(S, W, fluctuations_only=fluct, pod_rank=2)
    assert res2["rank"] == 2 and np.allclose(res2["basis"], Phi[:, :2])


def test_snapshot_text_format():
    v = np.array([1.0, -0.0, 1.0 / 3.0, 2.5e-17, -123456.789012345, 0.0])
    line = OP.snapshot_line(v)
    assert line == "1 -0 0.333333333333 2.5e-17 -123456.789012 0 \n"
    back = OP.read_data(line + line)
    assert len(back) == 2 and np.allclose(back[0], v, rtol=1e-11, atol=0)
    res = dict(basis=np.eye(3)[:, :2], rank=2, mean=np.array([0.5, 0.0, -1.0]), eigs=np.array([3.0, 1.0, 0.0]))
    pod_text, mean_text, eig_text = OP.pod_files(res)
    assert pod_text == "POD basis: fluctuating field\n1 0 0 \n0 1 0 \n"
    assert mean_text == "0.5 0 -1 "
    assert eig_text == "1 \t3 \t0.75\n2 \t1 \t1\n"
