"""Batched solves (no reference counterpart): every system of an uneven batch must equal
its standalone solve within tolerance.  The standalone solve is the oracle's FP64 GETRF +
GETRS (general systems) or POTRF + POTRS (SPD systems) on the same matrix; the batch and
the oracle order their arithmetic differently (blocked right-looking vs the crate's
panel order), so the check is: the same pivots, the solutions within the conditioning
bound of each other, and the batch's own backward error at FP64 level.  The refined
FP32 variant must reach backward error <= n * 2^-52 in every column, leave A unchanged,
and report its iteration count.  A captured hipGraph must replay to the eager result bit
for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def B():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import bola_amd
    bola_amd.lib()
    return bola_amd


@pytest.fixture(scope="module")
def O():
    import oracle
    oracle.lib()
    return oracle


def _batch(B, count, nmin, nmax, seed):
    from bola_amd import batchgen
    d, arena, ptot = batchgen.make_batch(count, nmin, nmax, 16, seed)
    return d, arena, ptot


def _sys(d, arena_h, i):
    n, r = int(d["n"][i]), int(d["nrhs"][i])
    A = arena_h[d["a"][i]:d["a"][i] + n * n].reshape(n, n).T
    Bm = arena_h[d["b"][i]:d["b"][i] + n * r].reshape(r, n).T
    return n, r, A, Bm


def _berr(A, X, Bm):
    R = Bm - A @ X
    an = np.max(np.sum(np.abs(A), axis=1))
    return float(np.max(np.max(np.abs(R), axis=0) / (an * np.max(np.abs(X), axis=0) + np.max(np.abs(Bm), axis=0))))


def _oracle_solve(O, B, kind, n, r, A, Bm):
    a = np.asfortranarray(A).ravel(order="F").copy()
    x = np.asfortranarray(Bm).ravel(order="F").copy()
    if kind == B.BATCH_LU:
        p = np.zeros(n, np.int32)
        O.getrf(n, n, a, n, p)
        O.getrs(n, r, a, n, p, x, n)
        return x.reshape(r, n).T, p
    O.potrf("L", n, a, n)
    O.potrs("L", n, r, a, n, x, n)
    return x.reshape(r, n).T, None


def _run(B, variant, d, arena, ptot, config):
    import torch
    count = len(d)
    piv = torch.zeros(max(ptot, 1), dtype=torch.int32, device="cuda")
    info = torch.full((count,), -7, dtype=torch.int32, device="cuda")
    iters = torch.full((count,), -7, dtype=torch.int32, device="cuda")
    work = arena.clone()
    B.batch_solve(variant, d, work, piv, info, iters, config)
    torch.cuda.synchronize()
    return work.cpu().numpy(), piv.cpu().numpy(), info.cpu().numpy(), iters.cpu().numpy()


@pytest.mark.parametrize("config", [(16, 2), (8, 1), None])
def test_batch_fp64_equals_standalone(B, O, config):
    d, arena, ptot = _batch(B, 48, 8, 150, seed=3)
    orig = arena.cpu().numpy()
    out, piv, info, iters = _run(B, B.BATCH_FP64, d, arena, ptot, config)
    assert np.all(info == 0) and np.all(iters == 0)
    for i in range(len(d)):
        n, r, A, Bm = _sys(d, orig, i)
        X = out[d["b"][i]:d["b"][i] + n * r].reshape(r, n).T
        kind = int(d["kind"][i])
        xo, po = _oracle_solve(O, B, kind, n, r, A if kind == B.BATCH_LU else np.tril(A) + np.tril(A, -1).T, Bm)
        if kind == B.BATCH_LU:
            assert np.array_equal(piv[d["piv"][i]:d["piv"][i] + n], po), i
        assert _berr(A, X, Bm) <= 10 * n * EPS, i
        ferr = np.max(np.abs(X - xo)) / np.max(np.abs(xo))
        assert ferr <= 1e-9, (i, n, ferr)


def test_batch_fp32_ir_reaches_fp64_accuracy(B, O):
    d, arena, ptot = _batch(B, 40, 8, 200, seed=5)
    orig = arena.cpu().numpy()
    out, piv, info, iters = _run(B, B.BATCH_FP32_IR, d, arena, ptot, (16, 2))
    assert np.all(info == 0)
    assert np.all((iters >= 1) | (iters == -1)) and np.any(iters >= 1)
    for i in range(len(d)):
        n, r, A, Bm = _sys(d, orig, i)
        # A is left as it was
        assert np.array_equal(out[d["a"][i]:d["a"][i] + n * n], orig[d["a"][i]:d["a"][i] + n * n])
        X = out[d["b"][i]:d["b"][i] + n * r].reshape(r, n).T
        Af = A if int(d["kind"][i]) == B.BATCH_LU else np.tril(A) + np.tril(A, -1).T
        assert _berr(Af, X, Bm) <= 2 * n * EPS, (i, n, iters[i])


@pytest.mark.parametrize("variant", ["fp64", "fp32_ir"])
def test_batch_full_order_range_against_oracle(B, O, variant):
    # the batched configuration's whole order range, 8 .. 512, on a sample of 240 systems
    d, arena, ptot = _batch(B, 240, 8, 512, seed=9)
    assert int(d["n"].max()) > 480 and int(d["n"].min()) < 40
    orig = arena.cpu().numpy()
    v = B.BATCH_FP64 if variant == "fp64" else B.BATCH_FP32_IR
    out, piv, info, iters = _run(B, v, d, arena, ptot, None)
    assert np.all(info == 0)
    for i in range(len(d)):
        n, r, A, Bm = _sys(d, orig, i)
        X = out[d["b"][i]:d["b"][i] + n * r].reshape(r, n).T
        kind = int(d["kind"][i])
        Af = A if kind == B.BATCH_LU else np.tril(A) + np.tril(A, -1).T
        xo, po = _oracle_solve(O, B, kind, n, r, Af, Bm)
        if variant == "fp64":
            assert iters[i] == 0
            if kind == B.BATCH_LU:
                assert np.array_equal(piv[d["piv"][i]:d["piv"][i] + n], po), i
            assert _berr(Af, X, Bm) <= 10 * n * EPS, i
        else:
            assert iters[i] >= 1 or iters[i] == -1, (i, iters[i])
            assert _berr(Af, X, Bm) <= 2 * n * EPS, (i, n, iters[i])
        # forward error against the oracle's solution, within the conditioning bound
        ferr = np.max(np.abs(X - xo)) / np.max(np.abs(xo))
        assert ferr <= 10 * np.linalg.cond(Af) * n * EPS, (i, n, ferr)


def test_batch_graph_replays_equal_eager(B):
    import torch
    d, arena, ptot = _batch(B, 24, 8, 96, seed=7)
    ref, _, _, _ = _run(B, B.BATCH_FP64, d, arena, ptot, (16, 2))
    work = arena.clone()
    piv = torch.zeros(ptot, dtype=torch.int32, device="cuda")
    info = torch.zeros(len(d), dtype=torch.int32, device="cuda")
    g = B.BatchGraph(B.BATCH_FP64, d, work, piv, info, None, (16, 2))
    try:
        for rep in range(3):
            work.copy_(arena)  # eager work between the replays
            torch.cuda.synchronize()
            g.launch()
            torch.cuda.synchronize()
            assert np.array_equal(work.cpu().numpy(), ref), rep
            assert int(info.abs().sum()) == 0
    finally:
        g.close()


def test_batch_tune_keeps_the_arena(B):
    import torch
    d, arena, ptot = _batch(B, 64, 8, 128, seed=9)
    before = arena.clone()
    nb, wpc = B.batch_tune(B.BATCH_FP64, d, arena)
    torch.cuda.synchronize()
    assert nb in (8, 16) and 1 <= wpc <= 4
    assert torch.equal(before, arena)
    # the tuned configuration is then the default for such a batch
    out, _, info, _ = _run(B, B.BATCH_FP64, d, arena, ptot, None)
    assert np.all(info == 0)


def test_batch_reports_singular_and_not_spd_per_system(B):
    import torch
    d, arena, ptot = _batch(B, 6, 20, 40, seed=11)
    h = arena.cpu().numpy()
    lu = np.nonzero(d["kind"] == B.BATCH_LU)[0]
    ch = np.nonzero(d["kind"] == B.BATCH_CHOL)[0]
    assert len(lu) and len(ch)
    i, j = int(lu[0]), int(ch[0])
    n = int(d["n"][i])
    h[d["a"][i] + 4 * n:d["a"][i] + 5 * n] = 0.0        # column 5 of a general system is zero
    nj = int(d["n"][j])
    h[d["a"][j] + 3 * nj + 3] = -1.0                      # A(4, 4) of an SPD system < 0
    arena = torch.from_numpy(h).cuda()
    _, _, info, _ = _run(B, B.BATCH_FP64, d, arena, ptot, (8, 1))
    assert info[i] >= 1 and info[i] <= 5
    assert info[j] == 4
    others = [k for k in range(len(d)) if k not in (i, j)]
    assert np.all(info[others] == 0)


def test_batch_rejects_bad_descriptors(B):
    import torch
    d, arena, ptot = _batch(B, 4, 8, 16, seed=13)
    d2 = d.copy()
    d2["n"][1] = 513
    with pytest.raises(B.BolaError) as e:
        B.batch_solve(B.BATCH_FP64, d2, arena, torch.zeros(ptot, dtype=torch.int32, device="cuda"),
                      torch.zeros(4, dtype=torch.int32, device="cuda"))
    assert e.value.code == -4
    with pytest.raises(B.BolaError):
        B.batch_solve(B.BATCH_FP64, d, arena, torch.zeros(ptot, dtype=torch.int32, device="cuda"),
                      torch.zeros(4, dtype=torch.int32, device="cuda"), config=(12, 1))


def test_batch_rejects_a_missing_pivot_arena_for_lu_systems(B):
    # an LU descriptor with no pivot arena is an argument error (no launch), not a device fault
    import torch
    d, arena, ptot = _batch(B, 16, 8, 40, 5)
    assert (d["kind"] == B.BATCH_LU).any()
    info = torch.zeros(len(d), dtype=torch.int32, device="cuda")
    with pytest.raises(B.BolaError) as e:
        B.batch_solve(B.BATCH_FP64, d, arena, None, info)
    assert e.value.code == -6 and "pivot arena" in str(e.value)
    with pytest.raises(B.BolaError) as e:
        B.BatchGraph(B.BATCH_FP64, d, arena, None, info)
    assert e.value.code == -6
    # Cholesky-only batches need none
    dc = d[d["kind"] == B.BATCH_CHOL]
    if len(dc):
        infc = torch.zeros(len(dc), dtype=torch.int32, device="cuda")
        B.batch_solve(B.BATCH_FP64, dc, arena, None, infc)
        torch.cuda.synchronize()
        assert int((infc != 0).sum()) == 0
