"""FP64 personalized PageRank and its adjoint by dense linear solves.

TEST INFRASTRUCTURE ONLY (small graphs).  The forward is the fixed point of
the PageRank update of algorithm/pr.cu:449-516 generalised to a
personalization vector s:  p = alpha * P^T p + (1 - alpha) * s with
P = D^-1 A (rows of vertices without out-edges are zero, as in pr.cu).
Pinned in tests/test_oracle.py by central finite differences.
"""
import numpy as np


def transition(n, rowptr, colind):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, colind[rowptr[i]:rowptr[i + 1]]] = 1.0
    deg = A.sum(1)
    inv = np.where(deg > 0, 1.0 / np.maximum(deg, 1), 0.0)
    return A * inv[:, None]  # P = D^-1 A


def forward(P, S, alpha):
    n = P.shape[0]
    return (1.0 - alpha) * np.linalg.solve(np.eye(n) - alpha * P.T, S)


def adjoint(P, S, W, alpha):
    """(dL/dS, dL/dalpha) for L = sum(W * forward(P, S, alpha))."""
    n = P.shape[0]
    lam = np.linalg.solve(np.eye(n) - alpha * P, W)
    p = forward(P, S, alpha)
    return (1.0 - alpha) * lam, float((lam * (P.T @ p - S)).sum())
