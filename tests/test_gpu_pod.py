"""Snapshots and POD on the HIP path (through the C ABI) against the oracle
(oracle/pod.py) and the reference's snaps_pod_rom_test golden values: the
mean, the eigenvalues (to 1e-12 of the largest) and the modes (to 1e-9) of
random snapshot sets whose Gram matrices cover one and several 16 x 16 MFMA
tiles and the 64 x 64 blocking; the golden TCD2D run with the device
collector and with the snapshot file; the files of write_pod; the batched
POD of a TCD sweep against the oracle per problem."""
import os

import numpy as np
import pytest
import torch

from mymoonmd_amd import parmoon as pm
from oracle import oracle as O
from oracle import pod as OP

pytestmark = pytest.mark.gpu


def dense_csr(rp, col, vals, n):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, col[rp[i]:rp[i + 1]]] = vals[rp[i]:rp[i + 1]]
    return A


def mass_matrix(ctx, nref):
    d = pm.TDomain("UnitSquare", nref)
    s = pm.TFESpace(d, nref, 1, ctx)
    M = pm.FEMatrix(ctx, s)
    pm.assemble_tcd(ctx, s, pm.cd_params("tcd_exp"), None, M, None)
    return s, M


def check_against_oracle(pod, S, W, fluct, pod_rank=0):
    res = OP.pod(S, W, fluctuations_only=fluct, pod_rank=pod_rank)
    n, m, r, valid = pod.info()
    assert (n, m) == S.shape and r == res["rank"] and valid == res["valid"]
    eigs = pod.get_pod_eigs()
    scale = abs(res["eigs"][0])
    assert np.max(np.abs(eigs - res["eigs"])) <= 1e-12 * scale
    assert np.max(np.abs(pod.get_snaps_avr().cpu().numpy() - res["mean"])) <= 1e-14 * max(1.0, np.max(np.abs(S)))
    B = pod.get_basis().cpu().numpy().T
    assert B.shape == res["basis"].shape
    # modes of well separated eigenvalues agree to the eigenvector conditioning
    ev = res["eigs"]
    for k in range(r):
        gap = min(abs(ev[k] - ev[j]) for j in range(len(ev)) if j != k)
        if gap > 1e-3 * scale:
            assert np.max(np.abs(B[:, k] - res["basis"][:, k])) <= 1e-9 * np.max(np.abs(res["basis"][:, k]))


@pytest.mark.parametrize("m", [3, 17, 70])
@pytest.mark.parametrize("fluct", [True, False])
def test_pod_euclidean_matches_oracle(ctx, m, fluct):
    print("case euclidean m", m, "fluct", fluct, flush=True)
    rng = np.random.default_rng(m)
    n = 9000                       # three row chunks
    S = rng.standard_normal((n, m)) @ np.diag(np.logspace(0, -4, m))
    pod = pm.POD(ctx, pod_fluctuations_only=fluct)
    pod.set_snapshots(torch.as_tensor(np.ascontiguousarray(S.T)).cuda())
    pod.compute_basis()
    check_against_oracle(pod, S, None, fluct)


@pytest.mark.parametrize("m", [4, 21])
def test_pod_l2_matches_oracle(ctx, m):
    print("case L2 m", m, flush=True)
    s, M = mass_matrix(ctx, 5)
    rp, col = M.get_structure()
    W = dense_csr(rp, col, M.get_entries(), s.n_dof)
    rng = np.random.default_rng(100 + m)
    S = rng.standard_normal((s.n_dof, m))
    for rank in (0, 2):
        pod = pm.POD(ctx, pod_rank=rank, pod_inner_product="L2")
        pod.set_gramian(M)
        pod.set_snapshots(torch.as_tensor(np.ascontiguousarray(S.T)).cuda())
        pod.compute_basis()
        check_against_oracle(pod, S, W, True, pod_rank=rank)
    # with the Dirichlet entries of the snapshots zero (homogeneous boundary
    # fluctuations) S^T W S is symmetric and the modes are W-orthonormal
    S[s.active_bound:] = 0.0
    pod = pm.POD(ctx, pod_inner_product="L2")
    pod.set_gramian(M)
    pod.set_snapshots(torch.as_tensor(np.ascontiguousarray(S.T)).cuda())
    pod.compute_basis()
    check_against_oracle(pod, S, W, True)
    B = pod.get_basis().cpu().numpy().T
    assert np.allclose(B.T @ W @ B, np.eye(B.shape[1]), atol=1e-10)


def test_pod_golden_snaps_pod_rom_on_gpu(ctx, golden, tmp_path):
    """test/snaps_pod_rom_test.c++ through the device path: TCD2D run with the
    collector (device copy and snapshot file), POD with the L2 Gramian from
    the device collector and from the file, files of write_pod."""
    g = golden["snaps_pod_rom_test"]
    tau, n_steps = g["time_step_length"], g["n_steps"]
    d = pm.TDomain("UnitSquare", g["n_refine"])
    tss = pm.TimeDiscretization(ctx, "backward_euler", tau, 0.0, tau * n_steps)
    sp = pm.solver_params(iterative_solver_type="fgmres", preconditioner="jacobi", gmres_restart=50,
                          max_n_iterations=3000, residual_tolerance=1e-14, residual_reduction=0.0)
    tcd = pm.TimeConvectionDiffusion(ctx, d, 1, "tcd_exp", tss, sp, eps=1.0, b=(1.0, -1.0, 0.0), c=1.0)
    n = tcd.spaces[-1].n_dof
    snap_dir = str(tmp_path / "snapshots_tmp")
    snaps = pm.SnapshotsCollector(ctx, n, 8, snaps_directory=snap_dir, snaps_basename="snaps_tmp",
                                  steps_per_snap=g["steps_per_snap"])
    tss.current_step_ = 0
    tcd.assemble_initial_time()
    assert snaps.write_data(tcd.solution)
    while not tss.reached_final_time_step():
        tss.current_step_ += 1
        tss.current_time_ += tss.tau
        tcd.assemble()
        rep = tcd.solve()
        assert rep.converged
        snaps.write_data(tcd.solution, tss.current_step_)
    assert snaps.count() == 3
    pod_dir = str(tmp_path / "pod_basis_tmp")
    for source in ("device", "file"):
        tp = pm.TimeConvectionDiffusionPOD(ctx, d, 1, pod_directory=pod_dir, pod_basename="pod_tmp",
                                           pod_inner_product="L2")
        tp.compute_pod_basis(snaps if source == "device" else snaps.filename)
        mean = tp.pod.get_snaps_avr().cpu().numpy()
        eigs = tp.pod.get_pod_eigs()
        assert np.max(np.abs(mean - np.array(g["mean"]))) < g["tolerance"], source
        assert np.max(np.abs(eigs[:2] - np.array(g["eigs"]))) < g["tolerance"], source
    # the snapshot file: three lines of 25 values, 12 significant digits
    text = open(snaps.filename).read()
    rows = OP.read_data(text)
    assert len(rows) == 3 and all(len(r) == n for r in rows) and text.endswith(" \n")
    assert text.splitlines()[0] == OP.snapshot_line(np.array(rows[0]))[:-1]
    # the files of write_pod: same structure as the oracle's, same numbers
    base = os.path.join(pod_dir, "pod_tmp.L2.")
    pod_text = open(base + "pod").read().splitlines()
    assert pod_text[0] == "POD basis: fluctuating field" and len(pod_text) == 3
    B = tp.pod.get_basis().cpu().numpy()
    for k in range(2):
        assert np.allclose(np.array(pod_text[1 + k].split(), float), B[k], rtol=1e-11, atol=1e-300)
    assert np.allclose(np.array(open(base + "mean").read().split(), float), mean, rtol=1e-11, atol=1e-300)
    eig_lines = open(base + "eigs").read().splitlines()
    assert len(eig_lines) == 2 and all(len(line.split(" \t")) == 3 for line in eig_lines)
    assert eig_lines[0].startswith("1 \t")
    assert abs(float(eig_lines[0].split(" \t")[1]) - g["eigs"][0]) < g["tolerance"]


@pytest.mark.parametrize("inner", ["euclidean", "L2"])
def test_batch_pod_matches_oracle(ctx, inner):
    """Snapshots of every problem of a batched backward-Euler sweep, POD of all
    problems at once, against the oracle's POD of each problem's snapshots
    (with its own mass matrix for L2)."""
    probs = [(8, 1.0, (1.0, -1.0), 1.0), (16, 0.05, (0.6, 0.8), 0.0), (12, 0.3, (-1.0, 0.2), 0.5)]
    b = pm.BatchTCD(ctx, probs, tau=0.01, residual_reduction=1e-12, inner_reduction=1e-5)
    b.initialize()
    b.reserve_snapshots(6)
    host = [[b.solution(p)] for p in range(len(probs))]
    b.write_snapshot()
    for _ in range(4):
        b.run(2)
        b.write_snapshot()
        for p in range(len(probs)):
            host[p].append(b.solution(p))
    b.compute_pod(pod_inner_product=inner)
    for p in range(len(probs)):
        S = np.array(host[p]).T
        W = None
        if inner == "L2":
            rp, col = b.structure(p)
            _, mvals, _ = b.matrices(p)
            W = dense_csr(rp, col, mvals, S.shape[0])
        res = OP.pod(S, W)
        eigs, basis, mean = b.pod(p)
        scale = res["eigs"][0]
        assert np.max(np.abs(eigs - res["eigs"])) <= 1e-11 * scale, p
        assert np.max(np.abs(mean - res["mean"])) <= 1e-13, p
        assert basis.shape == res["basis"].shape, p
        ev = res["eigs"]
        for k in range(res["rank"]):
            gap = min(abs(ev[k] - ev[j]) for j in range(len(ev)) if j != k)
            if gap > 1e-3 * scale:
                assert np.max(np.abs(basis[:, k] - res["basis"][:, k])) <= 1e-8 * np.max(np.abs(res["basis"][:, k]))
