"""Error measures shared by the GPU tests (test infrastructure)."""
import numpy as np


def norm_inf(A):
    """||A||_inf of a CSR operator (rp / val): the largest row sum of |a_ij|."""
    rs = np.zeros(len(A.rp) - 1)
    rows = np.repeat(np.arange(len(A.rp) - 1), np.diff(A.rp))
    np.add.at(rs, rows, np.abs(A.val))
    return float(rs.max()) if len(rs) else 0.0


def berr(A, X, B):
    """Largest per-column normwise backward error in the infinity norm,
    ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf): the refinement's stopping quantity."""
    X = np.asarray(X).reshape(A.n, -1)
    B = np.asarray(B).reshape(A.n, -1)
    na = norm_inf(A)
    out = 0.0
    for j in range(X.shape[1]):
        r = B[:, j] - A.matvec(X[:, j])
        den = na * np.abs(X[:, j]).max() + np.abs(B[:, j]).max()
        out = max(out, float(np.abs(r).max() / den) if den > 0 else 0.0)
    return out


def relerr(a, b):
    """||a - b|| / ||b|| (2-norm over all entries)."""
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def inv_norm(solve, n, iters=30, seed=0):
    """||A^-1||_2 of an SPD operator from its solve: the Rayleigh quotient of A^-1 after
    `iters` inverse iterations (it approaches the norm from below)."""
    x = np.random.default_rng(seed).standard_normal(n)
    x /= np.linalg.norm(x)
    q = 0.0
    for _ in range(iters):
        y = solve(x)
        q = float(np.dot(x, y))
        x = y / np.linalg.norm(y)
    return q


def forward_within_bound(A, X, Xref, B, invnorm, slack=1.5):
    """Per column: ||x - x_ref|| <= slack * ||A^-1|| (||b - A x|| + ||b - A x_ref||), the forward
    error a solution's own residual allows (x - x* = A^-1 r), for both x and the reference.
    Returns (ok, worst ratio of the error to its bound)."""
    X = np.asarray(X).reshape(A.n, -1)
    Xref = np.asarray(Xref).reshape(A.n, -1)
    B = np.asarray(B).reshape(A.n, -1)
    worst = 0.0
    for j in range(X.shape[1]):
        r = np.linalg.norm(B[:, j] - A.matvec(X[:, j]))
        r0 = np.linalg.norm(B[:, j] - A.matvec(Xref[:, j]))
        err = np.linalg.norm(X[:, j] - Xref[:, j])
        bound = invnorm * (r + r0)
        worst = max(worst, err / bound if bound > 0 else (0.0 if err == 0 else np.inf))
    return worst <= slack, worst
