"""Factorization objects (bola_hip_factor_*; no reference counterpart -- the reference's
callers keep an (a, ipiv) pair): a factorization made once serves later right-hand sides.
The FP64 kinds must give exactly the oracle's getrf + getrs / potrf + potrs (the
reference's algorithm) for every right-hand-side set; the mixed kind must give exactly
what the one-shot mixed solver gives for the same matrix and right-hand sides, reach
backward error <= n * 2^-52 in every column, and keep its FP64 fallback factorization
once it has needed it."""
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


def _cuda(x):
    import torch
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _berr(A, X, Bm):
    R = Bm - A @ X
    an = np.max(np.sum(np.abs(A), axis=1))
    return np.max(np.abs(R), axis=0) / (an * np.max(np.abs(X), axis=0) + np.max(np.abs(Bm), axis=0))


def test_fp64_lu_factor_serves_repeated_right_hand_sides(B, O):
    import torch
    n = 1000
    rng = np.random.default_rng(1)
    a = rng.uniform(-1, 1, n * n)
    ref, p = a.copy(), np.zeros(n, np.int32)
    assert O.getrf(n, n, ref, n, p)[0] == 0
    with B.Factor(B.FACTOR_LU, n, a) as f:           # from host memory
        d = f.describe()
        assert d["kind"] == B.FACTOR_LU and d["n"] == n and d["ld"] == n
        assert d["bytes"] == B.factor_bytes(B.FACTOR_LU, n)
        for nrhs, seed in [(5, 2), (70, 3), (1, 4)]:
            b = np.random.default_rng(seed).uniform(-1, 1, n * nrhs)
            want = b.copy()
            O.getrs(n, nrhs, ref, n, p, want, n)
            x = _cuda(b)
            f.solve(nrhs, x, x)                      # in place
            torch.cuda.synchronize()
            assert np.array_equal(x.cpu().numpy(), want), nrhs


@pytest.mark.parametrize("kind,uplo", [("CHOL_L", "L"), ("CHOL_U", "U")])
def test_fp64_cholesky_factor(B, O, kind, uplo):
    import torch
    n = 777
    rng = np.random.default_rng(5)
    G = rng.uniform(-1, 1, (n, n))
    a = (G @ G.T + n * np.eye(n)).ravel(order="F").copy()
    ref = a.copy()
    assert O.potrf(uplo, n, ref, n)[0] == 0
    with B.Factor(getattr(B, "FACTOR_" + kind), n, _cuda(a)) as f:  # from device memory
        for nrhs in (3, 40):
            b = rng.uniform(-1, 1, n * nrhs)
            want = b.copy()
            O.potrs(uplo, n, nrhs, ref, n, want, n)
            bx, x = _cuda(b), torch.zeros(n * nrhs, dtype=torch.float64, device="cuda")
            f.solve(nrhs, bx, x)
            torch.cuda.synchronize()
            assert np.array_equal(x.cpu().numpy(), want)


def test_factor_errors(B):
    n = 64
    a = np.random.default_rng(6).uniform(-1, 1, (n, n))
    a[:, 40] = 0.0
    with pytest.raises(B.BolaError) as e:
        B.Factor(B.FACTOR_LU, n, a.ravel(order="F").copy())
    assert e.value.code == 41 and "singular" in e.value.message
    with pytest.raises(B.BolaError) as e:
        B.Factor(B.FACTOR_CHOL_L, n, -np.eye(n).ravel())
    assert e.value.code == 1
    with pytest.raises(B.BolaError) as e:
        B.Factor(9, n, np.eye(n).ravel())
    assert e.value.code == -2


@pytest.mark.parametrize("precision,rbt", [("fp32", False), ("fp32", True), ("fp16", True), ("bf16", False)])
def test_mixed_factor_equals_the_one_shot_solver(B, precision, rbt):
    import torch
    from bola_amd import testmat
    n = 1024
    A = testmat.conditioned(n, 1e3, 11, device="cuda")
    An = A.cpu().numpy().reshape(n, n).T
    with B.Factor(B.FACTOR_MIXED, n, A, precision=precision, rbt=rbt) as f:
        d = f.describe()
        assert d["butterfly"] == (1 if rbt else 0) and d["bytes"] == B.factor_bytes(B.FACTOR_MIXED, n, precision, rbt)
        for seed, nrhs in [(1, 1), (2, 16), (3, 20)]:
            b = torch.from_numpy(np.random.default_rng(seed).uniform(-1, 1, n * nrhs)).cuda()
            x = torch.zeros_like(b)
            st = f.solve(nrhs, b, x)
            torch.cuda.synchronize()
            assert st["converged"] == 1 and st["final_berr"] <= n * EPS
            X = x.cpu().numpy().reshape(nrhs, n).T
            assert np.all(_berr(An, X, b.cpu().numpy().reshape(nrhs, n).T) <= n * EPS)
            if nrhs <= 16:  # the one-shot solver on the same system: the same bits
                x1 = torch.zeros_like(b)
                B.dgesv_mixed(n, nrhs, A, n, b, n, x1, n, precision, rbt=rbt)
                torch.cuda.synchronize()
                assert torch.equal(x, x1), (seed, nrhs)


def test_mixed_factor_keeps_its_fp64_fallback(B):
    # kappa = 1e12 in fp16 without pivoting: refinement and GMRES-IR give up, the FP64 LU
    # answers (bit-identical to getrf + getrs), and the next solve goes straight to it
    import torch
    from bola_amd import testmat
    n = 512
    A = testmat.conditioned(n, 1e12, 12, device="cuda")
    a = A.clone()
    p = torch.zeros(n, dtype=torch.int32, device="cuda")
    B.getrf(n, n, a, n, p)
    with B.Factor(B.FACTOR_MIXED, n, A, precision="fp16", rbt=True) as f:
        for seed in (1, 2):
            b = torch.from_numpy(np.random.default_rng(seed).uniform(-1, 1, n)).cuda()
            x = torch.zeros_like(b)
            st = f.solve(1, b, x)
            want = b.clone()
            B.getrs(n, 1, a, n, p, want, n)
            torch.cuda.synchronize()
            assert st["fallback"] == 1, st
            assert torch.equal(x, want)
            if seed == 2:
                assert st["ir_iterations"] == 0 and st["gmres_iterations"] == 0
        assert f.describe()["bytes"] > B.factor_bytes(B.FACTOR_MIXED, n, "fp16", True)


def test_server_mixed_precision_requests(B):
    # mixed-precision LU requests beside FP64 ones; resident mixed factorizations serve
    # repeated right-hand sides, every solution at FP64 backward error
    from bola_amd import testmat
    rng = np.random.default_rng(9)
    reqs = []
    for i, n in enumerate([700, 1500, 300, 1500, 2048]):
        A = testmat.conditioned(n, 1e4, 40 + i)
        nrhs = int(rng.integers(1, 20))
        prec = B.serve_mixed("fp32", rbt=(i % 2 == 0)) if i != 2 else 0
        reqs.append((n, A, nrhs, rng.uniform(-1, 1, n * nrhs), prec, 100 + i if i in (1, 4) else 0))
    reqs.append((1500, reqs[1][1], 7, rng.uniform(-1, 1, 1500 * 7), reqs[1][4], 101))
    with B.Server(8 << 30, 4) as s:
        live = []
        for k, (n, A, nrhs, b, prec, key) in enumerate(reqs):
            if k == len(reqs) - 1:
                s.wait(live[1][0])  # the repeat comes once its factorization is resident
            x = np.zeros(n * nrhs)
            live.append((s.submit(B.SERVE_LU, n, nrhs, A, b, x, key, precision=prec), x))
        for k, (t, x) in enumerate(live):
            if k == 1:
                continue
            res = s.wait(t)
            assert res["status"] == 0 and res["info"] == 0
            if k == len(live) - 1:
                assert res["cache_hit"] == 1
        for (n, A, nrhs, b, prec, key), (t, x) in zip(reqs, live):
            An = A.reshape(n, n).T
            assert np.all(_berr(An, x.reshape(nrhs, n).T, b.reshape(nrhs, n).T) <= n * EPS), n
    with B.Server(1 << 30, 1) as s:
        with pytest.raises(B.BolaError) as e:
            s.submit(B.SERVE_CHOL, 100, 1, np.eye(100).ravel(), np.ones(100), np.zeros(100), 0,
                     precision=B.serve_mixed("fp32"))
        assert e.value.code == -2
