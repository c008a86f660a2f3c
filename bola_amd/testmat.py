"""Seeded synthetic matrices used by the benchmark and the tests.

`conditioned(n, kappa, seed)` builds A = U diag(sigma) V^T with sigma_i = kappa^(-i/(n-1))
(geometric from 1 to 1/kappa) and U = H(u1) H(u2), V = H(v1) H(v2) products of two
Householder reflectors H(v) = I - 2 v v^T / (v^T v), v ~ U[-1, 1) from numpy's PCG64 with
the given seed.  Built in FP64 with torch on the given device (or numpy on the host).
"""
from __future__ import annotations

import numpy as np


def _reflectors(n, seed):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-1.0, 1.0, n) for _ in range(4)]  # u1, u2, v1, v2


def conditioned(n: int, kappa: float, seed: int = 7, device=None):
    """Column-major n x n matrix as a flat FP64 array (numpy) or CUDA tensor (device)."""
    u1, u2, v1, v2 = _reflectors(n, seed)
    sig = kappa ** (-np.arange(n, dtype=np.float64) / max(n - 1, 1))
    if device is None:
        M = np.diag(sig)
        for v in (v2, v1):          # M <- M H(v)
            M = M - np.outer(M @ v, v) * (2.0 / (v @ v))
        for v in (u2, u1):          # M <- H(v) M
            M = M - np.outer(v, v @ M) * (2.0 / (v @ v))
        return np.ascontiguousarray(M.T).reshape(-1)  # row-major of M^T = column-major of M
    import torch
    M = torch.diag(torch.from_numpy(sig).to(device))
    for v in (v2, v1):
        t = torch.from_numpy(v).to(device)
        M = M - torch.outer(M @ t, t) * (2.0 / float(v @ v))
    for v in (u2, u1):
        t = torch.from_numpy(v).to(device)
        M = M - torch.outer(t, t @ M) * (2.0 / float(v @ v))
    return M.t().contiguous().reshape(-1)
