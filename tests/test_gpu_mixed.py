"""Mixed-precision solver (no reference counterpart): the low-precision trailing-update
GEMM against an FP64 product of the same rounded operands, and the refined solve against
its own contract (backward error <= n * eps64 in every column) and against the FP64
oracle's solution (forward error within the conditioning bound).  The FP64 fallback is
the reference-exact path, so when it is taken the answer must equal the oracle's GETRF +
GETRS bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS64 = np.finfo(np.float64).eps
EPS32 = float(np.finfo(np.float32).eps)


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


def colmajor(M):
    return np.asfortranarray(M).ravel(order="F")


@pytest.mark.parametrize("prec", ["fp32", "fp16", "bf16"])
@pytest.mark.parametrize("m,n,k", [(128, 128, 32), (300, 200, 96), (1000, 513, 256), (64, 700, 33), (257, 129, 1)])
def test_sgemm_lp_matches_fp64_product_of_rounded_operands(B, prec, m, n, k):
    import torch
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    Bm = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    Cm = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    dA = torch.from_numpy(colmajor(A)).cuda()
    dB = torch.from_numpy(colmajor(Bm)).cuda()
    dC = torch.from_numpy(colmajor(Cm)).cuda()
    B.sgemm_lp(m, n, k, dA, m, dB, k, dC, m, prec)
    torch.cuda.synchronize()
    got = dC.cpu().numpy().reshape(n, m).T.astype(np.float64)
    rnd = {"fp32": torch.float32, "fp16": torch.float16, "bf16": torch.bfloat16}[prec]
    Ar = torch.from_numpy(A).to(rnd).double().numpy()
    Br = torch.from_numpy(Bm).to(rnd).double().numpy()
    want = Cm.astype(np.float64) - Ar @ Br
    # fp32 accumulation over k products plus the final subtraction
    bound = (k + 2) * EPS32 * (np.abs(Ar) @ np.abs(Br) + np.abs(Cm))
    assert np.all(np.abs(got - want) <= bound + 1e-30), float(np.max(np.abs(got - want) - bound))


@pytest.mark.parametrize("m,n,k", [(128, 128, 32), (300, 200, 96), (1000, 513, 256), (64, 700, 33), (257, 129, 1),
                                   (2048, 1024, 512)])
def test_sgemm_split_fp16_is_as_accurate_as_fp32(B, m, n, k):
    # each fp32 operand as h + 2^-11 l (h, l fp16): the product, with the 2^-22 l l term
    # dropped, against the FP64 product of the unrounded fp32 operands.  Per product the
    # representation errors are near 2^-22 relative; fp32 accumulation adds k eps32.
    import torch
    rng = np.random.default_rng(m * 5 + n * 11 + k)
    A = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    Bm = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    Cm = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    dA = torch.from_numpy(colmajor(A)).cuda()
    dB = torch.from_numpy(colmajor(Bm)).cuda()
    dC = torch.from_numpy(colmajor(Cm)).cuda()
    B.sgemm_lp(m, n, k, dA, m, dB, k, dC, m, "fp16x3")
    torch.cuda.synchronize()
    got = dC.cpu().numpy().reshape(n, m).T.astype(np.float64)
    A64, B64 = A.astype(np.float64), Bm.astype(np.float64)
    want = Cm.astype(np.float64) - A64 @ B64
    bound = ((k + 2) * EPS32 + 8 * 2.0 ** -22) * (np.abs(A64) @ np.abs(B64) + np.abs(Cm))
    assert np.all(np.abs(got - want) <= bound + 1e-30), float(np.max(np.abs(got - want) - bound))
    # and clearly better than plain fp16 products of the same operands
    dC2 = torch.from_numpy(colmajor(Cm)).cuda()
    B.sgemm_lp(m, n, k, dA, m, dB, k, dC2, m, "fp16")
    torch.cuda.synchronize()
    got16 = dC2.cpu().numpy().reshape(n, m).T.astype(np.float64)
    if k >= 32:
        assert np.max(np.abs(got - want)) * 64 < np.max(np.abs(got16 - want))


def _solve(B, n, nrhs, kappa, prec, seed=7, rbt=False):
    import torch
    from bola_amd import testmat
    A = testmat.conditioned(n, kappa, seed, device="cuda")
    rng = np.random.default_rng(seed + 1)
    b = rng.uniform(-1, 1, n * nrhs)
    db = torch.from_numpy(b).cuda()
    dx = torch.zeros(n * nrhs, dtype=torch.float64, device="cuda")
    st = B.dgesv_mixed(n, nrhs, A, n, db, n, dx, n, prec, rbt=rbt)
    torch.cuda.synchronize()
    return A.cpu().numpy(), b, dx.cpu().numpy(), st


def _berr(Af, b, x, n, nrhs):
    Am = Af.reshape(n, n).T
    X = x.reshape(nrhs, n).T
    Bm = b.reshape(nrhs, n).T
    R = Bm - Am @ X
    an = np.max(np.sum(np.abs(Am), axis=1))
    return np.max(np.max(np.abs(R), axis=0) / (an * np.max(np.abs(X), axis=0) + np.max(np.abs(Bm), axis=0)))


@pytest.mark.parametrize("rbt", [False, True], ids=["tournament", "butterfly"])
@pytest.mark.parametrize("prec,kappa", [("fp32", 1e2), ("fp32", 1e5), ("fp32", 1e8), ("fp16", 1e2),
                                        ("fp16", 1e5), ("bf16", 1e2), ("bf16", 1e5), ("fp16x3", 1e2),
                                        ("fp16x3", 1e5), ("fp16x3", 1e8)])
@pytest.mark.parametrize("n,nrhs", [(512, 1), (1024, 16), (1000, 3)])
def test_mixed_solve_reaches_fp64_backward_error(B, O, prec, kappa, n, nrhs, rbt):
    Af, b, x, st = _solve(B, n, nrhs, kappa, prec, rbt=rbt)
    print(f"n={n} nrhs={nrhs} {prec} kappa={kappa:g} rbt={rbt}: {st}")
    assert st["converged"] == 1, st
    assert st["final_berr"] <= n * EPS64
    # the same criterion recomputed on the host
    assert _berr(Af, b, x, n, nrhs) <= 2 * n * EPS64
    # bf16 products keep 8 significand bits: at these orders the refinement's contraction is
    # close to 1 and, seed by seed, GMRES-IR may stall into the FP64 fallback (which still
    # meets the contract above), whichever panel factorization produced the factors
    if (kappa <= 1e2 and prec != "bf16") or (prec == "fp16x3" and kappa <= 1e5):
        assert st["fallback"] == 0, st
    # forward error against the FP64 oracle's solution, within the conditioning bound
    a = Af.copy()
    p = np.zeros(n, np.int32)
    code, _ = O.getrf(n, n, a, n, p)
    assert code == 0
    xo = b.copy()
    O.getrs(n, nrhs, a, n, p, xo, n)
    X, XO = x.reshape(nrhs, n), xo.reshape(nrhs, n)
    ferr = np.max(np.abs(X - XO), axis=1) / np.max(np.abs(XO), axis=1)
    assert np.all(ferr <= 10 * kappa * n * EPS64), ferr


def test_mixed_host_flavour_equals_device(B):
    import torch
    n, nrhs = 700, 2
    Af, b, x_dev, st_dev = _solve(B, n, nrhs, 1e3, "fp32", seed=11)
    x = np.zeros(n * nrhs)
    st = B.dgesv_mixed(n, nrhs, Af.copy(), n, b.copy(), n, x, n, "fp32")
    assert np.array_equal(x, x_dev)
    assert st["ir_iterations"] == st_dev["ir_iterations"]
    del torch


def test_mixed_fallback_is_the_exact_fp64_path(B, O):
    # bf16 factors of a kappa = 1e13 matrix cannot refine to n * eps64: the FP64 path answers
    n, nrhs = 384, 2
    Af, b, x, st = _solve(B, n, nrhs, 1e13, "bf16", seed=5)
    assert st["fallback"] == 1, st
    a = Af.copy()
    p = np.zeros(n, np.int32)
    O.getrf(n, n, a, n, p)
    xo = b.copy()
    O.getrs(n, nrhs, a, n, p, xo, n)
    assert np.array_equal(x, xo)


def test_mixed_rejects_bad_arguments(B):
    import torch
    a = torch.zeros(16, dtype=torch.float64, device="cuda")
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    with pytest.raises(B.BolaError) as e:
        B.dgesv_mixed(4, 17, a, 4, x, 4, x, 4, "fp32")
    assert e.value.code == -4
    with pytest.raises(B.BolaError) as e:
        B.dgesv_mixed(4, 1, a, 3, x, 4, x, 4, "fp32")
    assert e.value.code == -6


def test_mixed_butterfly_needs_no_pivoting(B, O):
    # a matrix whose leading entry is zero cannot be factored without interchanges, but
    # its butterfly-randomised image can
    import torch
    n, nrhs = 256, 1
    rng = np.random.default_rng(3)
    M = rng.uniform(-1, 1, (n, n))
    M[0, 0] = 0.0
    M[:, 0] = 0.0
    M[1, 0] = 1.0
    a = torch.from_numpy(colmajor(M)).cuda()
    b = rng.uniform(-1, 1, n)
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    st = B.dgesv_mixed(n, nrhs, a, n, torch.from_numpy(b).cuda(), n, x, n, "fp32", rbt=True)
    assert st["converged"] == 1 and st["fallback"] == 0, st
    assert _berr(colmajor(M), b, x.cpu().numpy(), n, nrhs) <= 2 * n * EPS64


@pytest.mark.parametrize("n,nrhs", [(514, 16), (1001, 2)])
def test_mixed_butterfly_request_falls_back_to_tournament_when_n_is_not_a_multiple_of_4(B, O, n, nrhs):
    # two butterfly levels need n divisible by 4: other orders are factored with
    # tournament pivoting and must meet the same contract
    Af, b, x, st = _solve(B, n, nrhs, 1e3, "fp32", seed=21, rbt=True)
    assert st["converged"] == 1 and st["fallback"] == 0, st
    assert _berr(Af, b, x, n, nrhs) <= 2 * n * EPS64


# ---- the mixed-precision configuration at full size: n = 8192, kappa in {1e2, 1e5, 1e8},
# nrhs in {1, 16}, every product precision with and without butterflies

_BIG = {}


def _big_case(O, kappa):
    """A (device and host copies) and the oracle's LU of it, built once per kappa."""
    if kappa not in _BIG:
        import torch
        from bola_amd import testmat
        n = 8192
        A = testmat.conditioned(n, kappa, 7, device="cuda")
        Af = A.cpu().numpy()
        lu = Af.copy()
        p = np.zeros(n, np.int32)
        code, _ = O.getrf(n, n, lu, n, p)
        assert code == 0
        _BIG.clear()  # one kappa at a time (half a GiB each)
        torch.cuda.empty_cache()
        _BIG[kappa] = (A, Af, lu, p)
    return _BIG[kappa]


@pytest.mark.parametrize("prec,rbt", [("fp32", True), ("fp32", False), ("fp16", True), ("fp16", False),
                                      ("bf16", True), ("bf16", False), ("fp16x3", True), ("fp16x3", False)],
                         ids=["fp32-butterfly", "fp32-tournament", "fp16-butterfly", "fp16-tournament",
                              "bf16-butterfly", "bf16-tournament", "fp16x3-butterfly", "fp16x3-tournament"])
@pytest.mark.parametrize("nrhs", [1, 16])
@pytest.mark.parametrize("kappa", [1e2, 1e5, 1e8])  # outermost: the matrix is built once per kappa
def test_mixed_config_n8192(B, O, kappa, nrhs, prec, rbt):
    import torch
    n = 8192
    A, Af, lu, p = _big_case(O, kappa)
    b = np.random.default_rng(8).uniform(-1, 1, n * nrhs)
    dx = torch.zeros(n * nrhs, dtype=torch.float64, device="cuda")
    st = B.dgesv_mixed(n, nrhs, A, n, torch.from_numpy(b).cuda(), n, dx, n, prec, rbt=rbt)
    torch.cuda.synchronize()
    x = dx.cpu().numpy()
    print(f"kappa={kappa:g} nrhs={nrhs} {prec} rbt={rbt}: ir {st['ir_iterations']} gmres "
          f"{st['gmres_iterations']} fallback {st['fallback']} berr {st['final_berr']:.2e}")
    assert st["converged"] == 1, st
    assert st["final_berr"] <= n * EPS64
    assert _berr(Af, b, x, n, nrhs) <= 2 * n * EPS64
    if prec in ("fp32", "fp16x3") and kappa <= 1e5:
        assert st["fallback"] == 0, st
    xo = b.copy()
    O.getrs(n, nrhs, lu, n, p, xo, n)
    if st["fallback"]:  # the FP64 path is the oracle's
        assert np.array_equal(x, xo)
    X, XO = x.reshape(nrhs, n), xo.reshape(nrhs, n)
    ferr = np.max(np.abs(X - XO), axis=1) / np.max(np.abs(XO), axis=1)
    assert np.all(ferr <= 10 * kappa * n * EPS64), ferr


@pytest.mark.parametrize("prec,rbt", [("fp32", True), ("fp16", False)], ids=["fp32-butterfly", "fp16-tournament"])
def test_mixed_gmres_ir_converges_without_fallback(B, O, prec, rbt):
    # at kappa = 1e8 plain refinement from these factors stalls and GMRES-IR finishes the
    # solve, with no FP64 factorization
    import torch
    n, nrhs, kappa = 8192, 1, 1e8
    A, Af, lu, p = _big_case(O, kappa)
    b = np.random.default_rng(8).uniform(-1, 1, n)
    dx = torch.zeros(n, dtype=torch.float64, device="cuda")
    st = B.dgesv_mixed(n, nrhs, A, n, torch.from_numpy(b).cuda(), n, dx, n, prec, rbt=rbt)
    torch.cuda.synchronize()
    assert st["converged"] == 1 and st["fallback"] == 0, st
    assert st["gmres_iterations"] > 0, st
    assert st["final_berr"] <= n * EPS64
    assert _berr(Af, b, dx.cpu().numpy(), n, nrhs) <= 2 * n * EPS64
