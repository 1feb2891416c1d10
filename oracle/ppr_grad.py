"""FP64 CPU restatement of the differentiable weighted personalized PageRank.

Test infrastructure only (the checker of cytom_ppr_weighted /
cytom_ppr_adjoint).  No reference counterpart: the definition is

    P_e = w_e / W_u   (W_u = sum of the weights leaving u; rows without
                       out-edges drop their mass)
    r   = (1 - a) s + a P^T r
    lam = g + a P lam                 (g = dL/dr)
    dL/ds   = (1 - a) lam
    dL/dw_e = a r_u (lam_v - sum_{e' leaving u} P_e' lam_dst(e')) / W_u

Pinned by: a dense direct solve on small graphs (``dense=True``), central
finite differences of ``forward`` (see tests/test_oracle.py), and the
identity with global PageRank when s is uniform and every weight is 1.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _matrix(V, src, dst, w):
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    w = np.asarray(w, np.float64)
    W = np.bincount(src, weights=w, minlength=V)
    p = w / W[src]
    # P[u, v] summed over parallel edges
    P = sp.csr_matrix((p, (src, dst)), shape=(V, V))
    return P, W, p


def forward(V, src, dst, w, s, alpha=0.85, tol=1e-15, max_iter=100000, dense=False):
    """r = (1-a) s + a P^T r, FP64 (Jacobi to tol, or a dense direct solve)."""
    P, _, _ = _matrix(V, src, dst, w)
    s = np.asarray(s, np.float64)
    if dense:
        A = np.eye(V) - alpha * P.T.toarray()
        return np.linalg.solve(A, (1 - alpha) * s)
    PT = P.T.tocsr()
    r = (1 - alpha) * s
    for _ in range(max_iter):
        nr = (1 - alpha) * s + alpha * (PT @ r)
        if np.max(np.abs(nr - r)) <= tol:
            return nr
        r = nr
    return r


def adjoint(V, src, dst, w, r, g, alpha=0.85, tol=1e-15, max_iter=100000, dense=False):
    """Returns (dL/dw per edge in input order, dL/ds, lambda), FP64."""
    P, W, p = _matrix(V, src, dst, w)
    g = np.asarray(g, np.float64)
    if dense:
        lam = np.linalg.solve(np.eye(V) - alpha * P.toarray(), g)
    else:
        lam = g.copy()
        for _ in range(max_iter):
            nl = g + alpha * (P @ lam)
            if np.max(np.abs(nl - lam)) <= tol:
                lam = nl
                break
            lam = nl
    mu = P @ lam
    src = np.asarray(src, np.int64)
    dst = np.asarray(dst, np.int64)
    r = np.asarray(r, np.float64)
    dw = alpha * r[src] * (lam[dst] - mu[src]) / W[src]
    return dw, (1 - alpha) * lam, lam


def loss_sq(r, targets):
    """L = sum_{i in targets} r_i^2 and its gradient."""
    g = np.zeros_like(np.asarray(r, np.float64))
    t = np.asarray(targets, np.int64)
    np.add.at(g, t, 2.0 * np.asarray(r, np.float64)[t])
    return float(np.sum(np.asarray(r, np.float64)[t] ** 2)), g
