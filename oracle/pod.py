"""ORACLE — test infrastructure only (tests/ and their fixtures).

numpy restatement of the reference's snapshot / POD path:

* snapshot_line(): one line of SnapshotsCollector::write_data
  (src/ReducedOrderModels/SnapshotsCollector.C:76-95): every entry printed by
  an ostream with setprecision(12) (printf %.12g) followed by a blank, then a
  newline.
* read_data(): POD::read_data (src/ReducedOrderModels/POD.C:564-589): one
  vector per line, values separated by blanks.
* pod(): POD::compute_basis (POD.C:121-235) with decompose_snaps (:534-561)
  and compute_autocorr_mat (:564-597): mean over the snapshots (subtracted with
  pod_fluctuations_only), autocorrelation S^T S or S^T M S, the eigenproblem
  of its upper triangle (dsyev 'U' there, numpy.linalg.eigh here),
  eigenvalues above 1e-10, basis
  phi_k = S v_i / sqrt(lambda_i) for the largest eigenvalues, eigenvalues
  returned largest first.  Eigenvector signs are normalised so that the
  largest-magnitude component is positive (the reference leaves the sign to
  LAPACK).
* pod_files(): the text of POD::write_pod / write_averages /
  write_eigenvalues (POD.C:299-330, 424-487).

This module never runs in the product path.
"""
import numpy as np

EIGEN_THRESHOLD = 1.0e-10   # POD::eigen_threshold (POD.C:66)


def fmt(v):
    """std::ostream << double after setprecision(12), default float field"""
    return "%.12g" % v


def snapshot_line(vec):
    return "".join(fmt(v) + " " for v in vec) + "\n"


def read_data(text):
    out = []
    for line in text.splitlines():
        vals = []
        for tok in line.split():
            try:
                vals.append(float(tok))
            except ValueError:
                break
        out.append(vals)
    return out


def _normalise_sign(v):
    k = int(np.argmax(np.abs(v)))
    return v if v[k] >= 0 else -v


def pod(S, W=None, fluctuations_only=True, pod_rank=0):
    """S: snapshots as columns (length x m); W: inner-product matrix (dense or
    anything with @), None for euclidean.  Returns dict(mean, eigs (largest
    first), basis (length x rank), rank, valid, G)."""
    S = np.array(S, dtype=np.float64, copy=True)
    n, m = S.shape
    mean = np.zeros(n)
    if fluctuations_only:
        for i in range(n):
            s = 0.0
            for j in range(m):
                s += S[i, j]
            mean[i] = (1.0 / m) * s
        S -= mean[:, None]
    WS = S if W is None else W @ S
    G = S.T @ WS
    # dsyev reads the upper triangle only (uplo 'U', POD.C:157-178); S^T W S
    # is not symmetric for a Gramian with unit Dirichlet rows
    G = np.triu(G) + np.triu(G, 1).T
    w, V = np.linalg.eigh(G)          # ascending, like dsyev
    valid = int(np.sum(w > EIGEN_THRESHOLD))
    rank = valid if pod_rank <= 0 else min(pod_rank, valid)
    basis = np.zeros((n, rank))
    for k in range(rank):
        i = m - 1 - k
        v = _normalise_sign(V[:, i])
        basis[:, k] = S @ (1.0 / np.sqrt(w[i]) * v)
    return dict(mean=mean, eigs=w[::-1].copy(), basis=basis, rank=rank, valid=valid, G=G)


def pod_files(result, fluctuations_only=True):
    """(pod, mean or None, eigs) file texts of POD::write_pod"""
    head = "POD basis: fluctuating field\n" if fluctuations_only else "POD basis: full field\n"
    pod_text = head + "".join(snapshot_line(result["basis"][:, k]) for k in range(result["rank"]))
    mean_text = "".join(fmt(v) + " " for v in result["mean"]) if fluctuations_only else None
    eigs = result["eigs"]
    valid = int(np.sum(eigs >= EIGEN_THRESHOLD))
    total = 0.0
    for e in eigs:
        total += e
    cum, lines = 0.0, []
    for i in range(min(len(eigs), valid)):
        cum += eigs[i] / total
        lines.append("%d \t%s \t%s\n" % (i + 1, fmt(eigs[i]), fmt(cum)))
    return pod_text, mean_text, "".join(lines)
