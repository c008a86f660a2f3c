"""The solve server (no reference counterpart): each result must be the same alone and in
the mix.  The server's slots run the one-shot FP64 entry points on their own streams, so
"the same" is bit for bit, and also bit for bit the result of calling bola_amd.getrf +
getrs (or potrf + potrs) directly.  Admission must respect the HBM budget, resident
factorizations must serve later right-hand sides and be evicted least recently used
first, and bad requests must be refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import bola_amd
    bola_amd.lib()
    return bola_amd


def _matrix(n, kind, seed, B):
    rng = np.random.default_rng(seed)
    G = rng.uniform(-1, 1, (n, n))
    A = G @ G.T + n * np.eye(n) if kind == B.SERVE_CHOL else G
    return np.asfortranarray(A).ravel(order="F").copy()


def _requests(B, seed=0):
    sizes = [33, 64, 100, 150, 257, 300, 450, 600, 129, 77, 513, 200, 350, 48]
    reqs = []
    rng = np.random.default_rng(seed)
    for i, n in enumerate(sizes):
        kind = B.SERVE_CHOL if i % 3 == 2 else B.SERVE_LU
        nrhs = int(rng.integers(1, 20))
        a = _matrix(n, kind, 100 + i, B)
        b = rng.uniform(-1, 1, n * nrhs)
        key = 1000 + i if i in (1, 4, 8) else 0
        reqs.append(dict(kind=kind, n=n, nrhs=nrhs, a=a, b=b, key=key, dev=(i % 2 == 0)))
    # repeated right-hand sides on three of them
    for j in (1, 4, 8):
        r = dict(reqs[j])
        r["b"] = rng.uniform(-1, 1, r["n"] * r["nrhs"])
        r["repeat"] = True
        reqs.append(r)
    return reqs


def _operands(r):
    import torch
    x = np.zeros(r["n"] * r["nrhs"])
    if r["dev"]:
        return torch.from_numpy(r["a"]).cuda(), torch.from_numpy(r["b"]).cuda(), torch.from_numpy(x).cuda()
    return r["a"].copy(), r["b"].copy(), x


def _host(x):
    return x.cpu().numpy() if hasattr(x, "cpu") else x


def _standalone(B, r):
    import torch
    n, nrhs = r["n"], r["nrhs"]
    a = torch.from_numpy(r["a"].copy()).cuda()
    x = torch.from_numpy(r["b"].copy()).cuda()
    if r["kind"] == B.SERVE_LU:
        p = torch.zeros(n, dtype=torch.int32, device="cuda")
        B.getrf(n, n, a, n, p)
        B.getrs(n, nrhs, a, n, p, x, n)
    else:
        B.potrf("L", n, a, n)
        B.potrs("L", n, nrhs, a, n, x, n)
    torch.cuda.synchronize()
    return x.cpu().numpy()


def test_serve_each_result_is_the_same_alone_and_in_the_mix(B):
    reqs = _requests(B)
    # alone: one request at a time
    alone = []
    with B.Server(8 << 30, 4) as s:
        for r in reqs:
            a, b, x = _operands(r)
            t = s.submit(r["kind"], r["n"], r["nrhs"], a, b, x, r["key"])
            res = s.wait(t)
            assert res["status"] == 0 and res["info"] == 0
            if r.get("repeat"):
                assert res["cache_hit"] == 1
            alone.append(_host(x).copy())
    # in the mix: the first requests all at once; the repeats once their factorizations
    # exist, while the rest are still running
    first = [r for r in reqs if not r.get("repeat")]
    rest = [r for r in reqs if r.get("repeat")]
    mix = [None] * len(reqs)
    with B.Server(8 << 30, 4) as s:
        live = []
        for i, r in enumerate(first):
            ops = _operands(r)
            live.append((i, s.submit(r["kind"], r["n"], r["nrhs"], *ops, r["key"]), ops))
        keyed = {r["key"]: i for i, r in enumerate(first) if r["key"]}
        waited = set()
        for k, i in keyed.items():
            tk = next(t for (j, t, _) in live if j == i)
            assert s.wait(tk)["status"] == 0
            waited.add(tk)
        for j, r in enumerate(rest):
            ops = _operands(r)
            live.append((len(first) + j, s.submit(r["kind"], r["n"], r["nrhs"], *ops, r["key"]), ops))
        slots = set()
        for i, t, ops in live:
            if t not in waited:
                res = s.wait(t)
                assert res["status"] == 0 and res["info"] == 0
                slots.add(res["slot"])
                if i >= len(first):
                    assert res["cache_hit"] == 1
            mix[i] = _host(ops[2]).copy()
        st = s.stats()
        assert st["completed"] == len(reqs) and st["cache_hits"] == len(rest)
        assert st["p50_us"] > 0 and st["p99_us"] >= st["p50_us"]
        assert len(slots) > 1  # the work was spread over the slots
    for i, r in enumerate(reqs):
        assert np.array_equal(alone[i], mix[i]), i
        assert np.array_equal(alone[i], _standalone(B, r)), i


def test_serve_admission_respects_the_budget(B):
    sizes = [400, 380, 360, 340, 320, 300]
    budget = 400 * 400 * 8 + 4096 + 400 * 8 * 4 + (1 << 20)  # room for one large request at a time
    reqs = [dict(kind=B.SERVE_LU, n=n, nrhs=3, a=_matrix(n, B.SERVE_LU, n, B),
                 b=np.random.default_rng(n).uniform(-1, 1, n * 3), key=0, dev=True) for n in sizes]
    with B.Server(budget, 4) as s:
        live = []
        for r in reqs:
            ops = _operands(r)
            live.append((r, s.submit(r["kind"], r["n"], r["nrhs"], *ops, 0), ops))
        for r, t, ops in live:
            assert s.wait(t)["status"] == 0
            assert np.array_equal(_host(ops[2]), _standalone(B, r))
        st = s.stats()
        assert st["hbm_peak"] <= budget
        assert st["hbm_in_use"] == 0


def test_serve_evicts_the_least_recently_used_factorization(B):
    n = 256
    f = n * n * 8 + 1024           # one resident factorization (with its pivots)
    budget = 2 * f + f // 2         # two resident, and room for one right-hand side
    with B.Server(budget, 2) as s:
        def run(key, seed):
            r = dict(kind=B.SERVE_LU, n=n, nrhs=1, a=_matrix(n, B.SERVE_LU, key, B),
                     b=np.random.default_rng(seed).uniform(-1, 1, n), key=key, dev=False)
            ops = _operands(r)
            return s.wait(s.submit(B.SERVE_LU, n, 1, *ops, key))
        assert run(1, 0)["cache_hit"] == 0
        assert run(2, 1)["cache_hit"] == 0
        assert run(1, 2)["cache_hit"] == 1   # 1 is now the most recent
        assert run(3, 3)["cache_hit"] == 0   # evicts 2
        assert run(1, 4)["cache_hit"] == 1
        assert run(2, 5)["cache_hit"] == 0
        assert s.stats()["evictions"] >= 2


def test_serve_reports_a_singular_matrix_and_refuses_bad_requests(B):
    n = 40
    a = _matrix(n, B.SERVE_LU, 1, B)
    a[5 * n:6 * n] = 0.0
    b = np.ones(n)
    x = np.full(n, 7.0)
    with B.Server(1 << 28, 2) as s:
        res = s.wait(s.submit(B.SERVE_LU, n, 1, a, b, x, 0))
        assert res["status"] == 0 and 1 <= res["info"] <= 6
        assert np.all(x == 7.0)
        with pytest.raises(B.BolaError) as e:
            s.submit(B.SERVE_LU, 20000, 1, np.zeros(4), np.zeros(4), np.zeros(4), 0)
        assert e.value.code == -3
        with pytest.raises(B.BolaError) as e:
            s.submit(7, n, 1, a, b, x, 0)
        assert e.value.code == -2
        with pytest.raises(B.BolaError) as e:
            s.submit(B.SERVE_LU, n, 1, None, b, x, 99)  # no matrix, nothing resident
        assert e.value.code == -3


def test_serve_large_factorizations_in_the_mix(B):
    # LU requests above 2048 take turns (their persistent panels must all be resident);
    # the small ones and the Cholesky ones run beside them; every result is the standalone one
    sizes = [(2600, B.SERVE_LU), (300, B.SERVE_LU), (2100, B.SERVE_LU), (2300, B.SERVE_CHOL), (3000, B.SERVE_LU),
             (150, B.SERVE_CHOL)]
    reqs = []
    rng = np.random.default_rng(77)
    for i, (n, kind) in enumerate(sizes):
        nrhs = 2
        reqs.append(dict(kind=kind, n=n, nrhs=nrhs, a=_matrix(n, kind, 500 + i, B),
                         b=rng.uniform(-1, 1, n * nrhs), key=0, dev=(i % 2 == 1)))
    with B.Server(8 << 30, 4) as s:
        live = []
        for r in reqs:
            ops = _operands(r)
            live.append((r, s.submit(r["kind"], r["n"], r["nrhs"], *ops, 0), ops))
        for r, t, ops in live:
            res = s.wait(t)
            assert res["status"] == 0 and res["info"] == 0, r["n"]
            assert np.array_equal(_host(ops[2]), _standalone(B, r)), r["n"]


def test_serve_hit_on_the_least_recently_used_factorization_under_budget_pressure(B):
    # Two resident factorizations fill the budget.  A hit on the least recently used one
    # whose right-hand sides need room must evict the other one, never its own.
    n, nrhs = 256, 40
    f = n * n * 8 + ((n * 4 + 255) & ~255)
    budget = 2 * f + n * nrhs * 8 // 2
    with B.Server(budget, 2) as s:
        reqs = {}
        for key in (1, 2):
            r = dict(kind=B.SERVE_LU, n=n, nrhs=1, a=_matrix(n, B.SERVE_LU, key, B),
                     b=np.random.default_rng(key).uniform(-1, 1, n), key=key, dev=False)
            reqs[key] = r
            assert s.wait(s.submit(B.SERVE_LU, n, 1, *_operands(r), key))["cache_hit"] == 0
        r = dict(reqs[1])
        r["nrhs"] = nrhs
        r["b"] = np.random.default_rng(9).uniform(-1, 1, n * nrhs)
        ops = _operands(r)
        res = s.wait(s.submit(B.SERVE_LU, n, nrhs, *ops, 1))
        assert res["status"] == 0 and res["info"] == 0 and res["cache_hit"] == 1
        assert np.array_equal(_host(ops[2]), _standalone(B, r))
        st = s.stats()
        assert st["evictions"] == 1 and st["hbm_peak"] <= budget
        # the evicted key is factored again; the kept one still serves
        assert s.wait(s.submit(B.SERVE_LU, n, 1, *_operands(reqs[2]), 2))["cache_hit"] == 0


def test_serve_config_sizes_match_the_oracle(B):
    # Orders of the serving configuration up to 8192, none of them a multiple of 64 except
    # the largest, in exactly sized stream-ordered buffers, LU and Cholesky side by side on
    # four slots, one repeated right-hand side from a resident factorization: every
    # solution is the oracle's (the reference's algorithm), bit for bit.
    import oracle as O
    O.lib()
    rng = np.random.default_rng(2026)
    spec = [(8192, B.SERVE_LU, 3, 11), (5661, B.SERVE_LU, 64, 0), (4097, B.SERVE_CHOL, 17, 12),
            (3913, B.SERVE_LU, 1, 0), (2705, B.SERVE_CHOL, 40, 0), (1000, B.SERVE_LU, 5, 0),
            (607, B.SERVE_CHOL, 2, 0), (8192, B.SERVE_LU, 9, 11)]
    mats = {}
    reqs = []
    for n, kind, nrhs, key in spec:
        if (n, kind) not in mats:
            mats[(n, kind)] = _matrix(n, kind, n, B)
        reqs.append(dict(kind=kind, n=n, nrhs=nrhs, a=mats[(n, kind)], b=rng.uniform(-1, 1, n * nrhs), key=key))
    with B.Server(16 << 30, 4) as s:
        live = []
        for i, r in enumerate(reqs):
            if i == len(reqs) - 1:  # the repeat comes once its factorization is resident
                s.wait(live[0][1])
            x = np.zeros(r["n"] * r["nrhs"])
            live.append((r, s.submit(r["kind"], r["n"], r["nrhs"], r["a"], r["b"], x, r["key"]), x))
        results = []
        for r, t, x in live[1:]:
            results.append((r, s.wait(t), x))
        st = s.stats()
    assert results[-1][1]["cache_hit"] == 1
    factors = {}
    for r, res, x in [(live[0][0], {"status": 0, "info": 0}, live[0][2])] + results:
        assert res["status"] == 0 and res["info"] == 0, r["n"]
        n, kind = r["n"], r["kind"]
        if (n, kind) not in factors:
            a = r["a"].copy()
            if kind == B.SERVE_LU:
                p = np.zeros(n, np.int32)
                assert O.getrf(n, n, a, n, p)[0] == 0
            else:
                p = None
                assert O.potrf("L", n, a, n)[0] == 0
            factors[(n, kind)] = (a, p)
        a, p = factors[(n, kind)]
        ref = r["b"].copy()
        if kind == B.SERVE_LU:
            O.getrs(n, r["nrhs"], a, n, p, ref, n)
        else:
            O.potrs("L", n, r["nrhs"], a, n, ref, n)
        assert np.array_equal(x, ref), (n, kind)
    assert st["completed"] == len(reqs) and st["hbm_in_use"] == st["hbm_cached"]
