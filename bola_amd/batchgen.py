"""Synthetic uneven batches for the batched solver: orders uniform in [nmin, nmax], half
general systems (entries uniform [-1, 1), LU) and half SPD (G G^T + n I with G uniform
[-1, 1), Cholesky), right-hand sides uniform [-1, 1) with nrhs uniform in [1, nrhs_max].

The descriptors are drawn with numpy's PCG64 from `seed`; the arena is filled on the
device with torch's generator seeded from `seed`, and the SPD matrices are formed there by
one batched product per distinct order.
"""
from __future__ import annotations

import numpy as np

from . import BATCH_CHOL, BATCH_DESC_DTYPE, BATCH_LU


def make_descs(count, nmin=8, nmax=512, nrhs_max=16, seed=1, spd_fraction=0.5):
    rng = np.random.default_rng(seed)
    n = rng.integers(nmin, nmax + 1, count)
    nrhs = rng.integers(1, nrhs_max + 1, count)
    kind = np.where(rng.random(count) < spd_fraction, BATCH_CHOL, BATCH_LU)
    d = np.zeros(count, BATCH_DESC_DTYPE)
    d["n"], d["nrhs"], d["kind"] = n, nrhs, kind
    sizes = n.astype(np.int64) * n + n.astype(np.int64) * nrhs
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    d["a"] = starts
    d["b"] = starts + n.astype(np.int64) * n
    pstarts = np.concatenate([[0], np.cumsum(n.astype(np.int64))[:-1]])
    d["piv"] = pstarts
    return d, int(sizes.sum()), int(n.sum())


def fill_arena(descs, total, seed=1, device="cuda"):
    """The arena (float64, `total` elements) for `descs`, on `device`."""
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    arena = torch.rand(total, dtype=torch.float64, device=device, generator=g)
    arena.mul_(2.0).sub_(1.0)
    spd = np.nonzero(descs["kind"] == BATCH_CHOL)[0]
    for n in np.unique(descs["n"][spd]):
        idx = spd[descs["n"][spd] == n]
        n = int(n)
        offs = torch.as_tensor(descs["a"][idx], device=device)
        ar = torch.arange(n * n, device=device)
        gather = (offs[:, None] + ar[None, :])
        G = arena[gather].view(len(idx), n, n)
        A = torch.bmm(G, G.transpose(1, 2).contiguous())
        A.diagonal(dim1=1, dim2=2).add_(float(n))
        arena[gather] = A.reshape(len(idx), n * n)
    return arena


def make_batch(count, nmin=8, nmax=512, nrhs_max=16, seed=1, device="cuda"):
    d, total, ptotal = make_descs(count, nmin, nmax, nrhs_max, seed)
    return d, fill_arena(d, total, seed, device), ptotal


def cost_flops(descs):
    """FP64-equivalent work: 2n^3/3 (LU) or n^3/3 (Cholesky) plus 2 n^2 nrhs per system."""
    n = descs["n"].astype(np.float64)
    f = np.where(descs["kind"] == BATCH_LU, 2.0 * n ** 3 / 3.0, n ** 3 / 3.0) + 2.0 * n * n * descs["nrhs"]
    return float(f.sum())
