"""Parity of the gfx950 path with the oracle (the CPU restatement of the reference's
algorithm), through the C ABI.  The GPU kernels keep the reference's rounding structure,
so the bar here is bit-for-bit equality; where a routine is a tree reduction (dot) or
takes the BLAS meaning the reference does not compute (syrk (U,N)/(L,T) beyond 32), the
tolerance is stated in the test."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps
HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "kats.json")))


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


def same(got, want, what=""):
    got = np.asarray(got)
    want = np.asarray(want)
    if np.array_equal(got, want, equal_nan=True):
        return
    bad = np.flatnonzero(~((got == want) | (np.isnan(got) & np.isnan(want))))
    i = bad[0]
    raise AssertionError(f"{what}: {len(bad)} of {got.size} differ; first at {i}: {got.flat[i]!r} vs {want.flat[i]!r}")


def arr(x):
    return np.array(x, dtype=np.float64)


# ------------------------------------------------------------------ reference KATs on the GPU

@pytest.mark.parametrize("name", ["gemm_nn", "gemm_nn_dup", "gemm_nt"])
def test_kat_gemm(B, name):
    t = KATS[name]
    c = arr(t["c"])
    B.gemm(t["transa"], t["transb"], t["m"], t["n"], t["k"], t["alpha"], arr(t["a"]), t["lda"], arr(t["b"]),
           t["ldb"], t["beta"], c, t["ldc"])
    assert np.all(np.abs(c - arr(t["expected"])) < t["tol"])


def test_kat_laswp_scal(B):
    t = KATS["laswp"]
    a = arr(t["a"])
    B.laswp(t["n"], a, t["lda"], t["k1"], t["k2"], np.array(t["ipiv"], np.int32), t["incx"])
    same(a, arr(t["expected"]), "laswp KAT")
    t = KATS["scal"]
    x = arr(t["x"])
    B.scal(len(x), t["alpha"], x, 1)
    same(x, arr(t["expected"]), "scal KAT")


@pytest.mark.parametrize("name", ["syrk_ln", "syrk_ut"])
def test_kat_syrk(B, name):
    t = KATS[name]
    c = arr(t["c"])
    n = t["n"]
    B.syrk(t["uplo"], t["trans"], n, t["k"], t["alpha"], arr(t["a"]), t["lda"], t["beta"], c, t["ldc"])
    exp = arr(t["expected_triangle"])
    for j in range(n):
        for i in (range(j, n) if t["uplo"] == "L" else range(0, j + 1)):
            assert abs(c[i + j * t["ldc"]] - exp[i + j * t["ldc"]]) < t["tol"]


@pytest.mark.parametrize("name", ["trsm_llnn", "trsm_rutn", "trsm_rltn"])
def test_kat_trsm(B, name):
    t = KATS[name]
    b = arr(t["b"])
    B.trsm(t["side"], t["uplo"], t["transa"], t["diag"], t["m"], t["n"], t["alpha"], arr(t["a"]), t["lda"], b,
           t["ldb"])
    assert np.all(np.abs(b - arr(t["expected"])) < t["tol"])


@pytest.mark.parametrize("name", ["potrf_u", "potrf_l"])
def test_kat_potrf(B, name):
    t = KATS[name]
    a = arr(t["a"])
    B.potrf(t["uplo"], t["n"], a, t["lda"])
    exp = arr(t["expected_triangle"])
    n = t["n"]
    for j in range(n):
        for i in (range(j, n) if t["uplo"] == "L" else range(0, j + 1)):
            assert abs(a[i + j * n] - exp[i + j * n]) < t["tol"]


def test_kat_getrf(B):
    t = KATS["getrf"]
    a = arr(t["a"])
    ipiv = np.zeros(3, np.int32)
    B.getrf(3, 3, a, 3, ipiv)
    assert np.all(np.abs(a - arr(t["expected"])) < t["tol"])
    assert list(ipiv) == t["ipiv_expected"]


# ------------------------------------------------------------------ IDAMAX / LASWP (bit-exact)

def test_idamax_random_and_ties(B, O):
    rng = np.random.default_rng(1)
    for n in [0, 1, 5, 15, 16, 17, 31, 32, 33, 100, 1000, 4099, 100000, 1 << 20]:
        x = rng.uniform(-1, 1, max(n, 1))
        assert B.idamax(n, x) == O.idamax(n, x), n
    for trial in range(400):
        n = int(rng.integers(1, 300))
        x = rng.integers(0, 4, n).astype(np.float64) * rng.choice([-1.0, 1.0], n)
        if trial % 3 == 0:
            x[rng.random(n) < 0.2] = np.nan
        if trial % 7 == 0:
            x[:] = 0.0
        if trial % 11 == 0:
            x[:] = np.nan
        assert B.idamax(n, x) == O.idamax(n, x), (trial, n)


@pytest.mark.parametrize("ncols", [1, 5, 64, 300])
@pytest.mark.parametrize("incx", [1, -1, 2])
def test_laswp_bitexact(B, O, ncols, incx):
    rng = np.random.default_rng(ncols * 7 + incx)
    m, lda = 200, 203
    k1, k2 = 5, 150
    span = (k1 - 1) + (k2 - k1) * abs(incx) + 1
    ipiv = rng.integers(1, m + 1, span).astype(np.int32)
    a = rng.uniform(-1, 1, lda * ncols)
    ref = a.copy()
    O.laswp(ncols, ref, lda, k1, k2, ipiv, incx)
    B.laswp(ncols, a, lda, k1, k2, ipiv, incx)
    same(a, ref, "laswp")


# ------------------------------------------------------------------ BLAS-1/2

def test_ger_scal_daxpy_bitexact(B, O):
    rng = np.random.default_rng(3)
    for m, n in [(1, 1), (7, 5), (130, 33), (1000, 64)]:
        x = rng.uniform(-1, 1, m)
        y = rng.uniform(-1, 1, 3 * n)
        y[::7] = 0.0
        a = rng.uniform(-1, 1, (m + 2) * n)
        ref = a.copy()
        O.ger(m, n, -1.3, x, y, 3, ref, m + 2)
        B.ger(m, n, -1.3, x, y, 3, a, m + 2)
        same(a, ref, f"ger {m}x{n}")
    for n in [1, 3, 4, 1001]:
        for inc in [1, 3]:
            x = rng.uniform(-1, 1, (n - 1) * inc + 1)
            y = rng.uniform(-1, 1, (n - 1) * inc + 1)
            rx, ry = x.copy(), y.copy()
            O.scal(n, 0.37, rx, inc)
            B.scal(n, 0.37, x, inc)
            same(x, rx, "scal")
            O.daxpy(n, 1.7, x, inc, ry, inc)
            B.daxpy(n, 1.7, x, inc, y, inc)
            same(y, ry, "daxpy")


def test_dot_tolerance(B, O):
    rng = np.random.default_rng(4)
    for n in [0, 1, 5, 1000, 1 << 20]:
        x = rng.uniform(-1, 1, max(n, 1))
        y = rng.uniform(-1, 1, max(n, 1))
        got = B.dot_product(n, x, 1, y, 1)
        want = O.dot_product(n, x, 1, y, 1)
        bound = n * EPS * float(np.sum(np.abs(x[:n] * y[:n]))) + 1e-300
        assert abs(got - want) <= 2 * bound, (n, got, want)


# ------------------------------------------------------------------ GEMM (bit-exact)

def _rand_op(rng, trans, rows, cols, ld_extra=0):
    """storage for op(X) of shape rows x cols, column-major with padding"""
    sr, sc = (rows, cols) if trans == "N" else (cols, rows)
    ld = sr + ld_extra
    x = rng.uniform(-1, 1, ld * sc)
    return x, ld


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (7, 5, 3), (64, 64, 64), (100, 37, 130), (257, 129, 200),
                                   (300, 600, 65)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 1.0), (-1.0, 1.0), (0.7, 0.0), (2.5, -0.3)])
def test_gemm_bitexact(B, O, ta, tb, m, n, k, alpha, beta):
    rng = np.random.default_rng(m * 131 + n * 17 + k)
    a, lda = _rand_op(rng, ta, m, k, 3)
    b, ldb = _rand_op(rng, tb, k, n, 1)
    ldc = m + 2
    c = rng.uniform(-1, 1, ldc * n)
    ref = c.copy()
    O.gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, ref, ldc)
    B.gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    same(c, ref, f"gemm {ta}{tb} {m}x{n}x{k}")


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("alpha,beta", [(1.0, 1.0), (2.5, -0.3), (0.7, 0.0)])
def test_gemm_large_tiles_bitexact(B, O, ta, tb, alpha, beta):
    # products large enough for the 128 x 64 tiles at two workgroups per CU; ragged edges
    # and k = 150 (two full 64-deep slabs and a partial one)
    m, n, k = 2048 + 37, 2048 + 5, 150
    rng = np.random.default_rng(7 + (ta == "T") * 2 + (tb == "T"))
    a, lda = _rand_op(rng, ta, m, k, 3)
    b, ldb = _rand_op(rng, tb, k, n, 1)
    ldc = m + 2
    c = rng.uniform(-1, 1, ldc * n)
    ref = c.copy()
    O.gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, ref, ldc)
    B.gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    same(c, ref, f"gemm {ta}{tb} {m}x{n}x{k}")


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("alpha,beta", [(1.0, 1.0), (2.5, -0.3)])
def test_gemm_128x128_tiles_bitexact(B, O, ta, tb, alpha, beta):
    # products with at least four 128 x 128 tiles per CU take the 8-wavefront tile; ragged
    # edges (m % 8, n % 4 nonzero) and k = 150
    m, n, k = 4096 + 37, 4096 + 9, 150
    rng = np.random.default_rng(11 + (ta == "T") * 2 + (tb == "T"))
    a, lda = _rand_op(rng, ta, m, k, 3)
    b, ldb = _rand_op(rng, tb, k, n, 1)
    ldc = m + 2
    c = rng.uniform(-1, 1, ldc * n)
    ref = c.copy()
    O.gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, ref, ldc)
    B.gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    same(c, ref, f"gemm {ta}{tb} {m}x{n}x{k}")


@pytest.mark.parametrize("alpha,beta", [(1.0, 1.0), (2.5, -0.3), (0.7, 0.0)])
@pytest.mark.parametrize("m,n,k", [(2048, 2048, 208), (2048, 2048, 1040), (4096, 4096, 256)])
def test_gemm_direct_load_tiles_bitexact(B, O, alpha, beta, m, n, k):
    # whole tiles of a plain product with even leading dimensions and k % 16 == 0 take the
    # direct-to-LDS kernels (three LDS stages, swizzled B columns): 128 x 64 tiles for the
    # 2048^2 products (k = 208 ends in a partial 64-deep slab, k = 1040 runs many stages
    # round the three buffers), 128 x 128 tiles on 8 wavefronts for the short-k 4096^2
    # product (the shape of the LU's trailing updates)
    lda, ldb, ldc = m + 2, k + 4, m + 6
    rng = np.random.default_rng(k + int(10 * alpha))
    a = rng.uniform(-1, 1, lda * k)
    b = rng.uniform(-1, 1, ldb * n)
    c = rng.uniform(-1, 1, ldc * n)
    ref = c.copy()
    O.gemm("N", "N", m, n, k, alpha, a, lda, b, ldb, beta, ref, ldc)
    B.gemm("N", "N", m, n, k, alpha, a, lda, b, ldb, beta, c, ldc)
    same(c, ref, f"gemm NN {m}x{n}x{k} direct loads")


def _fma_exact(a, b, c):
    """fma(a, b, c) rounded once: exact rational arithmetic, then float() (round to nearest even)"""
    from fractions import Fraction
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def test_mfma_f64_rounds_as_four_chained_fmas(B, O):
    # The bit-exactness of every product rests on v_mfma_f64_16x16x4f64 rounding like
    # fma(a3, b3, fma(a2, b2, fma(a1, b1, fma(a0, b0, acc)))), the reference's _mm256_fmadd_pd
    # chain.  A 16 x 16 x 4 product with alpha = 1, beta = 0 is one MFMA per element (16 x 16: no
    # edge elements), C = fma(1, acc, 0) = acc.  Terms spread over 2^+-40 with mixed signs make
    # the chained result differ from a once-rounded sum of the four products in many elements,
    # so the comparison separates the two.
    rng = np.random.default_rng(2024)
    m = n = 16
    k = 4
    mag = lambda size: rng.choice([-1.0, 1.0], size) * np.ldexp(rng.uniform(1, 2, size), rng.integers(-20, 21, size))
    a = mag(m * k)
    b = mag(k * n)
    c = np.full(m * n, np.nan)  # beta = 0: C is not read
    B.gemm("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, c, m)
    from fractions import Fraction
    chained = np.empty(m * n)
    backward = np.empty(m * n)
    once = np.empty(m * n)
    for j in range(n):
        for i in range(m):
            terms = [(a[i + q * m], b[q + j * k]) for q in range(k)]
            acc = 0.0
            for x, y in terms:
                acc = _fma_exact(x, y, acc)
            chained[i + j * m] = acc
            acc = 0.0
            for x, y in reversed(terms):
                acc = _fma_exact(x, y, acc)
            backward[i + j * m] = acc
            once[i + j * m] = float(sum(Fraction(x) * Fraction(y) for x, y in terms))
    same(c, chained, "MFMA vs four chained FMAs in k order")
    # the data separates the chain from a once-rounded sum and from the chain in reverse k order
    assert np.count_nonzero(chained != once) >= 20
    assert np.count_nonzero(chained != backward) >= 20
    # and the oracle (the reference's chain in C) agrees
    ref = np.full(m * n, np.nan)
    O.gemm("N", "N", m, n, k, 1.0, a, m, b, k, 0.0, ref, m)
    same(c, ref, "oracle")


def test_gemm_alpha_zero_beta_rules(B, O):
    c = np.array([np.nan, 1.0, 2.0, 3.0])
    B.gemm("N", "N", 2, 2, 1, 0.0, np.ones(2), 2, np.ones(2), 1, 0.0, c, 2)
    same(c, np.zeros(4), "beta=0")
    c = np.array([1.0, 2.0, 3.0, 4.0])
    B.gemm("N", "N", 2, 2, 1, 0.0, np.ones(2), 2, np.ones(2), 1, 3.0, c, 2)
    same(c, arr([3, 6, 9, 12]), "alpha=0")


def test_gemm_config1_1024(B, O):
    """DGEMM 'N','N' 1024^3, U[0,1), alpha = beta = 1: bit-exact with the oracle, and
    elementwise relative error <= k*eps against a float128-free float64 reference."""
    n = 1024
    rng = np.random.default_rng(1)
    a = rng.random(n * n)
    b = rng.random(n * n)
    c = rng.random(n * n)
    ref = c.copy()
    O.gemm("N", "N", n, n, n, 1.0, a, n, b, n, 1.0, ref, n)
    got = c.copy()
    B.gemm("N", "N", n, n, n, 1.0, a, n, b, n, 1.0, got, n)
    same(got, ref, "gemm 1024")
    exact = (a.reshape(n, n).T @ b.reshape(n, n).T + c.reshape(n, n).T).T.ravel()
    assert np.max(np.abs(got - exact) / exact) <= n * EPS


# ------------------------------------------------------------------ TRSM (bit-exact)

@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("size", [5, 64, 65, 200])
def test_trsm_bitexact(B, O, side, uplo, trans, diag, size):
    rng = np.random.default_rng(size * 16 + "LR".index(side) * 8 + "LU".index(uplo) * 4 + "NT".index(trans) * 2
                                + "NU".index(diag))
    m, n = (size, 37) if side == "L" else (37, size)
    na = m if side == "L" else n
    lda = na + 1
    A = rng.uniform(-1, 1, (na, na)) / na
    np.fill_diagonal(A, rng.uniform(1, 2, na))
    a = np.zeros(lda * na)
    a.reshape(na, lda).T[:na] = A
    b = rng.uniform(-1, 1, (m + 3) * n)
    b[5] = 0.0
    ref = b.copy()
    O.trsm(side, uplo, trans, diag, m, n, 1.5, a, lda, ref, m + 3)
    B.trsm(side, uplo, trans, diag, m, n, 1.5, a, lda, b, m + 3)
    same(b, ref, f"trsm {side}{uplo}{trans}{diag} {m}x{n}")


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("diag", ["N", "U"])
def test_trsm_sparse_rhs_bitexact(B, O, uplo, diag):
    # right-hand sides with most entries +0.0 or -0.0: the reference skips every step whose
    # entry is zero (a -0.0 entry stays -0.0); the lane solves turn those skips into selects
    m, n = 200, 41
    rng = np.random.default_rng(5 + "LU".index(uplo) * 2 + "NU".index(diag))
    lda = m + 1
    A = rng.uniform(-1, 1, (m, m)) / m
    np.fill_diagonal(A, rng.uniform(1, 2, m))
    a = np.zeros(lda * m)
    a.reshape(m, lda).T[:m] = A
    b = rng.uniform(-1, 1, m * n)
    pick = rng.random(m * n)
    b[pick < 0.6] = 0.0
    b[pick < 0.2] = -0.0
    ref = b.copy()
    O.trsm("L", uplo, "N", diag, m, n, 1.0, a, lda, ref, m)
    B.trsm("L", uplo, "N", diag, m, n, 1.0, a, lda, b, m)
    same(b, ref, f"trsm L{uplo}N{diag} sparse")
    assert np.array_equal(np.signbit(b), np.signbit(ref))


# ------------------------------------------------------------------ SYRK

@pytest.mark.parametrize("uplo,trans", [("L", "N"), ("U", "T")])
@pytest.mark.parametrize("n,k", [(3, 2), (32, 32), (33, 70), (100, 64), (130, 129)])
@pytest.mark.parametrize("beta", [1.0, 0.0, 0.5])
def test_syrk_bitexact(B, O, uplo, trans, n, k, beta):
    rng = np.random.default_rng(n * 1000 + k)
    ar, ac = (n, k) if trans == "N" else (k, n)
    a = rng.uniform(-1, 1, ar * ac)
    c = rng.uniform(-1, 1, (n + 1) * n)
    ref = c.copy()
    O.syrk(uplo, trans, n, k, -1.0, a, ar, beta, ref, n + 1)
    B.syrk(uplo, trans, n, k, -1.0, a, ar, beta, c, n + 1)
    same(c, ref, f"syrk {uplo}{trans} {n}x{k}")


@pytest.mark.parametrize("uplo,trans", [("L", "N"), ("U", "T")])
def test_syrk_large_bitexact(B, O, uplo, trans):
    # off-diagonal blocks on the masked MFMA GEMM at an order with at least four 128 x 128
    # tiles per CU, ragged (n % 32 = 8, k % 32 = 0 and k % 64 = 32)
    n, k = 4200, 96
    rng = np.random.default_rng(n + k + (uplo == "U"))
    rows = n if trans == "N" else k
    cols = k if trans == "N" else n
    a = rng.uniform(-1, 1, rows * cols)
    c = rng.uniform(-1, 1, n * n)
    ref = c.copy()
    O.syrk(uplo, trans, n, k, -1.0, a, rows, 1.0, ref, n)
    B.syrk(uplo, trans, n, k, -1.0, a, rows, 1.0, c, n)
    same(c, ref, f"syrk {uplo}{trans} {n}x{k}")


@pytest.mark.parametrize("uplo,trans", [("U", "N"), ("L", "T")])
def test_syrk_blas_meaning_other_combos(B, O, uplo, trans):
    rng = np.random.default_rng(9)
    # n, k <= 32: the reference computes the BLAS meaning; equal bit for bit
    n, k = 20, 30
    ar, ac = (n, k) if trans == "N" else (k, n)
    a = rng.uniform(-1, 1, ar * ac)
    c = rng.uniform(-1, 1, n * n)
    ref = c.copy()
    O.syrk(uplo, trans, n, k, 1.0, a, ar, 1.0, ref, n)
    B.syrk(uplo, trans, n, k, 1.0, a, ar, 1.0, c, n)
    same(c, ref, "syrk small")
    # larger: BLAS meaning, tolerance k*eps*|A||A|
    n, k = 90, 70
    ar, ac = (n, k) if trans == "N" else (k, n)
    A = rng.uniform(-1, 1, (ar, ac))
    c0 = rng.uniform(-1, 1, (n, n))
    c = c0.ravel(order="F").copy()
    B.syrk(uplo, trans, n, k, 1.0, A.ravel(order="F").copy(), ar, 1.0, c, n)
    full = (A @ A.T if trans == "N" else A.T @ A) + c0
    got = c.reshape(n, n).T
    mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    assert np.max(np.abs(got[mask] - full[mask])) <= 4 * k * EPS * k
    assert np.array_equal(got[~mask], c0[~mask])


# ------------------------------------------------------------------ GETRF / GETRS (bit-exact)

@pytest.mark.parametrize("m,n", [(1, 1), (3, 3), (31, 31), (32, 32), (33, 33), (100, 100), (257, 257),
                                 (1000, 1000), (300, 100), (100, 300), (20, 90), (90, 20), (513, 64)])
def test_getrf_bitexact(B, O, m, n):
    rng = np.random.default_rng(m * 7919 + n)
    lda = m + 1
    a = rng.uniform(-1, 1, lda * n)
    ipiv = np.zeros(min(m, n), np.int32)
    ref, rp = a.copy(), ipiv.copy()
    code, msg = O.getrf(m, n, ref, lda, rp)
    assert code == 0, msg
    B.getrf(m, n, a, lda, ipiv)
    same(ipiv, rp, "ipiv")
    same(a, ref, f"getrf {m}x{n}")


@pytest.mark.parametrize("m,n", [(80, 80), (70, 100), (100, 70), (65, 65)])
def test_getrf2_recursive_bitexact(B, O, m, n):
    rng = np.random.default_rng(m * 31 + n)
    a = rng.uniform(-1, 1, m * n)
    ipiv = np.zeros(min(m, n), np.int32)
    ref, rp = a.copy(), ipiv.copy()
    assert O.getrf2(m, n, ref, m, rp)[0] == 0
    B.getrf2(m, n, a, m, ipiv)
    same(ipiv, rp, "ipiv")
    same(a, ref, "getrf2")


def test_getrf_ties_and_zero_rows(B, O):
    # integer-valued matrix with many equal-magnitude candidates: pivot choice follows
    # the reference's IDAMAX tie rule, not "first index"
    rng = np.random.default_rng(12)
    n = 96
    a = rng.integers(-2, 3, n * n).astype(np.float64)
    a[: n * 4] = np.where(a[: n * 4] == 0, 1.0, a[: n * 4])
    ref, rp = a.copy(), np.zeros(n, np.int32)
    code, msg = O.getrf(n, n, ref, n, rp)
    ipiv = np.zeros(n, np.int32)
    if code == 0:
        B.getrf(n, n, a, n, ipiv)
        same(ipiv, rp, "ipiv ties")
        same(a, ref, "getrf ties")
    else:
        with pytest.raises(B.BolaError) as e:
            B.getrf(n, n, a, n, ipiv)
        assert e.value.message == msg


@pytest.mark.parametrize("m,n,ints", [(4500, 64, False), (8192, 40, False), (8192, 36, True), (9000, 40, False),
                                      (9000, 33, True)])
def test_getrf_tall_panels_bitexact(B, O, m, n, ints):
    # tall panels spread over 71..141 wavefronts (one row per lane up to 8192 rows, two
    # above), with and without equal-magnitude candidates in many wavefronts at once
    rng = np.random.default_rng(m + n)
    a = rng.integers(-3, 4, m * n).astype(np.float64) if ints else rng.uniform(-1, 1, m * n)
    ipiv = np.zeros(n, np.int32)
    ref, rp = a.copy(), ipiv.copy()
    code, msg = O.getrf(m, n, ref, m, rp)
    if code != 0:
        with pytest.raises(B.BolaError) as e:
            B.getrf(m, n, a, m, ipiv)
        assert e.value.message == msg
        return
    B.getrf(m, n, a, m, ipiv)
    same(ipiv, rp, "ipiv")
    same(a, ref, f"getrf {m}x{n}")


def test_getrf_errors(B, O):
    with pytest.raises(B.BolaError) as e:
        B.getrf(4, 4, np.zeros(16), 3, np.zeros(4, np.int32))
    assert str(e.value) == "Argument 4 to getrf had an illegal value of 3" and e.value.code == -4
    n = 96
    rng = np.random.default_rng(13)
    A = rng.uniform(-1, 1, (n, n))
    A[:, 70] = 0.0
    a = A.ravel(order="F").copy()
    ref = a.copy()
    code, msg = O.getrf(n, n, ref, n, np.zeros(n, np.int32))
    assert code == 7
    with pytest.raises(B.BolaError) as e:
        B.getrf(n, n, a, n, np.zeros(n, np.int32))
    assert e.value.message == msg and e.value.code == 71
    same(a, ref, "partially factored state")


@pytest.mark.parametrize("m,n,zc", [
    (600, 600, 293),      # 128-column super-panels: the second panel of the third one
    (2600, 2600, 620),    # 256-column super-panels: the fourth panel of the third one
    (2600, 2600, 1300),   # the first panel of a super-panel, while the previous one's bulk update runs
    (2600, 2600, 2590),   # the partial last super-panel
    (1200, 800, 290),     # rectangular, m > n
    (2048, 2600, 1800),   # rectangular, m < n, columns beyond the last super-panel
])
def test_getrf_zero_pivot_state_bitexact(B, O, m, n, zc):
    # A zero column stays exactly zero through every update, so the factorization stops at
    # its pivot.  The reference returns there (getrf2.rs:166-168, getrf.rs:53): the panels
    # before the failing one are applied to every column, the failing panel is left as far
    # as it got, ipiv of the failing panel stays panel-local.  The whole matrix and ipiv
    # must be the reference's, every time.
    rng = np.random.default_rng(m + n + zc)
    A = rng.uniform(-1, 1, (m, n))
    A[:, zc] = 0.0
    a0 = A.ravel(order="F").copy()
    ref = a0.copy()
    rp = np.full(min(m, n), -7, np.int32)
    code, msg = O.getrf(m, n, ref, m, rp)
    assert code == zc % 32 + 1
    for rep in range(2):
        a = a0.copy()
        ipiv = np.full(min(m, n), -7, np.int32)
        with pytest.raises(B.BolaError) as e:
            B.getrf(m, n, a, m, ipiv)
        assert e.value.message == msg and e.value.code == zc + 1
        same(ipiv, rp, f"ipiv after the zero pivot ({rep})")
        same(a, ref, f"partially factored state ({rep})")


def test_getrs_bitexact(B, O):
    rng = np.random.default_rng(14)
    n, nrhs = 300, 64
    a = rng.uniform(-1, 1, n * n)
    ipiv = np.zeros(n, np.int32)
    assert O.getrf(n, n, a, n, ipiv)[0] == 0
    b = rng.uniform(-1, 1, n * nrhs)
    ref = b.copy()
    O.getrs(n, nrhs, a, n, ipiv, ref, n)
    B.getrs(n, nrhs, a, n, ipiv, b, n)
    same(b, ref, "getrs")


@pytest.mark.parametrize("n", [2600, 2601])
def test_getrf_large_square_bitexact(B, O, n):
    # 256-column super-panels with a partial last one (2600), and the reference's own
    # panel-by-panel order for an order that is not a multiple of 8 (2601)
    rng = np.random.default_rng(n)
    a = rng.uniform(-1, 1, n * n)
    ref, rp = a.copy(), np.zeros(n, np.int32)
    assert O.getrf(n, n, ref, n, rp)[0] == 0
    ipiv = np.zeros(n, np.int32)
    B.getrf(n, n, a, n, ipiv)
    same(ipiv, rp, f"ipiv {n}")
    same(a, ref, f"lu {n}")


def test_getrf_getrs_2048_bitexact(B, O):
    n, nrhs = 2048, 16
    rng = np.random.default_rng(15)
    a = rng.uniform(-1, 1, n * n)
    b = rng.uniform(-1, 1, n * nrhs)
    ra, rp, rb = a.copy(), np.zeros(n, np.int32), b.copy()
    assert O.getrf(n, n, ra, n, rp)[0] == 0
    O.getrs(n, nrhs, ra, n, rp, rb, n)
    ipiv = np.zeros(n, np.int32)
    B.getrf(n, n, a, n, ipiv)
    B.getrs(n, nrhs, a, n, ipiv, b, n)
    same(ipiv, rp, "ipiv 2048")
    same(a, ra, "lu 2048")
    same(b, rb, "x 2048")


# ------------------------------------------------------------------ POTRF / POTRS (bit-exact)

def _spd(rng, n):
    G = rng.uniform(-1, 1, (n, n))
    return (G @ G.T + n * np.eye(n)).ravel(order="F").copy()


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n", [1, 5, 32, 33, 100, 300])
def test_potrf_bitexact(B, O, uplo, n):
    rng = np.random.default_rng(n * 3 + (uplo == "U"))
    a = _spd(rng, n)
    ref = a.copy()
    assert O.potrf(uplo, n, ref, n)[0] == 0
    B.potrf(uplo, n, a, n)
    same(a, ref, f"potrf {uplo} {n}")


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n", [1000, 2085])
def test_potrf_large_bitexact(B, O, uplo, n):
    # the blocked loop over many 32-column blocks: SYRK against thousands of previous
    # columns, GEMMs with k up to n - 32, a ragged last block
    rng = np.random.default_rng(n + (uplo == "U"))
    a = _spd(rng, n)
    ref = a.copy()
    assert O.potrf(uplo, n, ref, n)[0] == 0
    B.potrf(uplo, n, a, n)
    same(a, ref, f"potrf {uplo} {n}")


@pytest.mark.parametrize("uplo,n", [("L", 4096), ("U", 4096), ("U", 8192), ("L", 4133)])
def test_potrf_config_sizes_bitexact(B, O, uplo, n):
    # the plain FP64 configuration's Cholesky orders (SPD A = G G^T + n I, G uniform [-1, 1)),
    # and an order with ragged edges (n % 8 = 5, n % 32 = 5) on the device flavour
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(n)
    G = torch.rand(n, n, dtype=torch.float64, device="cuda", generator=g) * 2.0 - 1.0
    A = (G @ G.t() + n * torch.eye(n, dtype=torch.float64, device="cuda")).t().contiguous().reshape(-1)
    ref = A.cpu().numpy().copy()
    assert O.potrf(uplo, n, ref, n)[0] == 0
    B.potrf(uplo, n, A, n)
    torch.cuda.synchronize()
    same(A.cpu().numpy(), ref, f"potrf {uplo} {n}")


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n,k", [(600, 305), (600, 288), (2085, 1000), (2085, 2070)])
def test_potrf_failure_state_bitexact(B, O, uplo, n, k):
    # a non-positive leading minor at column k + 1: the reference returns at that block
    # (potrf.rs:124/144) with the earlier block columns (rows) factored, the failing
    # diagonal block as far as POTRF2 got, and everything later untouched
    rng = np.random.default_rng(n + k)
    A = _spd(rng, n).reshape(n, n, order="F")
    A[k, k] = -float(n) * n
    a0 = A.ravel(order="F").copy()
    ref = a0.copy()
    code, msg = O.potrf(uplo, n, ref, n)
    assert code == k % 32 + 1
    for rep in range(2):
        a = a0.copy()
        with pytest.raises(B.BolaError) as e:
            B.potrf(uplo, n, a, n)
        assert e.value.message == msg and e.value.code == k + 1
        same(a, ref, f"potrf {uplo} state after the failure ({rep})")


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_potrs_config_bitexact(B, O, uplo):
    n, nrhs = 4096, 64
    rng = np.random.default_rng(4096)
    f = _spd(rng, n)
    assert O.potrf(uplo, n, f, n)[0] == 0
    b = rng.uniform(-1, 1, n * nrhs)
    ref = b.copy()
    O.potrs(uplo, n, nrhs, f, n, ref, n)
    B.potrs(uplo, n, nrhs, f, n, b, n)
    same(b, ref, f"potrs {uplo} {n}x{nrhs}")


@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n", [20, 50, 77])
def test_potrf2_bitexact(B, O, uplo, n):
    rng = np.random.default_rng(n)
    a = _spd(rng, n)
    ref = a.copy()
    assert O.potrf2(uplo, n, ref, n)[0] == 0
    B.potrf2(uplo, n, a, n)
    same(a, ref, f"potrf2 {uplo} {n}")


def test_potrf_errors(B, O):
    with pytest.raises(B.BolaError) as e:
        B.potrf("X", 3, np.eye(3).ravel(), 3)
    assert str(e.value) == "Argument 1 to potrf had an illegal value"
    with pytest.raises(B.BolaError) as e:
        B.potrf("L", 3, np.eye(3).ravel(), 2)
    assert str(e.value) == "Argument 4 to potrf had an illegal value"
    n = 64
    for uplo in "LU":
        A = np.eye(n)
        A[36, 36] = -1.0
        ref = A.ravel(order="F").copy()
        code, msg = O.potrf(uplo, n, ref, n)
        with pytest.raises(B.BolaError) as e:
            B.potrf(uplo, n, A.ravel(order="F").copy(), n)
        assert e.value.message == msg and e.value.code == 37


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_potrs_bitexact(B, O, uplo):
    rng = np.random.default_rng(16)
    n, nrhs = 200, 9
    f = _spd(rng, n)
    assert O.potrf(uplo, n, f, n)[0] == 0
    b = rng.uniform(-1, 1, n * nrhs)
    ref = b.copy()
    O.potrs(uplo, n, nrhs, f, n, ref, n)
    B.potrs(uplo, n, nrhs, f, n, b, n)
    same(b, ref, "potrs")


@pytest.mark.parametrize("uplo", ["L", "U"])
def test_potrs_large_bitexact(B, O, uplo):
    # the recursive triangular solves (halving down to 64-row base cases, transposed GEMMs in
    # between) over many columns
    rng = np.random.default_rng(18 + (uplo == "U"))
    n, nrhs = 2085, 64
    f = _spd(rng, n)
    assert O.potrf(uplo, n, f, n)[0] == 0
    b = rng.uniform(-1, 1, n * nrhs)
    ref = b.copy()
    O.potrs(uplo, n, nrhs, f, n, ref, n)
    B.potrs(uplo, n, nrhs, f, n, b, n)
    same(b, ref, f"potrs {uplo} {n}")


@pytest.mark.parametrize("nrhs", [64, 37])
def test_getrs_large_bitexact(B, O, nrhs):
    # nrhs = 37: the last nrhs % 4 columns of every update GEMM take the reference's unfused
    # edge path, beside the 64-deep MFMA tiles of the recursive solve
    rng = np.random.default_rng(19)
    n = 2600
    a = rng.uniform(-1, 1, n * n)
    ipiv = np.zeros(n, np.int32)
    assert O.getrf(n, n, a, n, ipiv)[0] == 0
    b = rng.uniform(-1, 1, n * nrhs)
    ref = b.copy()
    O.getrs(n, nrhs, a, n, ipiv, ref, n)
    B.getrs(n, nrhs, a, n, ipiv, b, n)
    same(b, ref, f"getrs {n} nrhs {nrhs}")


# ------------------------------------------------------------------ device-pointer flavour

def test_device_flavour_matches_host(B, O):
    import torch
    rng = np.random.default_rng(17)
    n = 500
    a = rng.uniform(-1, 1, n * n)
    ref, rp = a.copy(), np.zeros(n, np.int32)
    O.getrf(n, n, ref, n, rp)
    da = torch.from_numpy(a.copy()).cuda()
    dp = torch.zeros(n, dtype=torch.int32, device="cuda")
    B.getrf(n, n, da, n, dp)
    torch.cuda.synchronize()
    same(dp.cpu().numpy(), rp, "dev ipiv")
    same(da.cpu().numpy(), ref, "dev lu")
    x = torch.from_numpy(rng.uniform(-1, 1, 1000)).cuda()
    assert B.idamax(1000, x) == O.idamax(1000, x.cpu().numpy())


@pytest.mark.parametrize("n,nrhs", [(4096, 1), (4096, 3), (4096, 64), (8192, 64)])
def test_getrf_getrs_config_bitexact(B, O, n, nrhs):
    # the plain FP64 configuration (n = 4096 and 8192, nrhs = 64) and the mixed-precision
    # configuration's few right-hand sides, whose updates all take the reference's unfused
    # edge path (nrhs < 4): LU factors, pivots and solutions bit for bit, device flavour
    import torch
    rng = np.random.default_rng(n + nrhs)
    a0 = rng.uniform(-1, 1, n * n)
    b0 = rng.uniform(-1, 1, n * nrhs)
    ref, rp, rb = a0.copy(), np.zeros(n, np.int32), b0.copy()
    assert O.getrf(n, n, ref, n, rp)[0] == 0
    O.getrs(n, nrhs, ref, n, rp, rb, n)
    a = torch.from_numpy(a0).cuda()
    p = torch.zeros(n, dtype=torch.int32, device="cuda")
    b = torch.from_numpy(b0).cuda()
    B.getrf(n, n, a, n, p)
    B.getrs(n, nrhs, a, n, p, b, n)
    torch.cuda.synchronize()
    same(p.cpu().numpy(), rp, f"ipiv {n}")
    same(a.cpu().numpy(), ref, f"lu {n}")
    same(b.cpu().numpy(), rb, f"x {n}x{nrhs}")


@pytest.mark.parametrize("n,nrhs,ldb,special", [(128, 16, 128, False), (256, 48, 260, False), (1024, 64, 1024, True),
                                                (2048, 128, 2051, True), (16384, 16, 16384, False)])
def test_getrs_one_launch_solves_bitexact(B, O, n, nrhs, ldb, special):
    # device GETRS on shapes the one-launch triangular solves take (n = 64 * 2^k, nrhs a
    # multiple of 16; lda and ldb padded): against the oracle's recursive solves.  The
    # factor is arbitrary data (the solves read its unit-lower and upper triangles), with a
    # dominant diagonal; 'special' plants exact zeros in B (the base case's skipped steps),
    # a zero on U's diagonal (division by zero, infinities and NaN through the rest) and
    # columns of L that are zero
    import torch
    rng = np.random.default_rng(n + nrhs)
    lda = n + 3
    lu = rng.uniform(-1, 1, lda * n)
    for j in range(n):
        lu[j + j * lda] += 4.0 if j % 3 else -4.0
    piv = np.array([rng.integers(j, n) + 1 for j in range(n)], np.int32)
    b0 = rng.uniform(-1, 1, ldb * nrhs)
    if special:
        b0[rng.integers(0, ldb * nrhs, ldb * nrhs // 8)] = 0.0
        lu[(n // 2 + 5) * (lda + 1)] = 0.0
        for j in (3, 70, n - 40):
            lu[j * lda + j + 1: j * lda + n] = 0.0
    rb = b0.copy()
    O.getrs(n, nrhs, lu, lda, piv, rb, ldb)
    a = torch.from_numpy(lu).cuda()
    p = torch.from_numpy(piv).cuda()
    b = torch.from_numpy(b0.copy()).cuda()
    B.getrs(n, nrhs, a, lda, p, b, ldb)
    torch.cuda.synchronize()
    got = b.cpu().numpy().reshape(nrhs, ldb)[:, :n]
    same(got, rb.reshape(nrhs, ldb)[:, :n], f"getrs {n}x{nrhs}")
    if special:
        assert np.isnan(got).any() or np.isinf(got).any()


def test_shared_device_handle_same_bits(B, O):
    # a handle told that its work shares the GPU takes the undivided LU streams and the
    # recursive triangular solves: the same bits as the oracle (and as the default handle)
    import torch
    n, nrhs = 2048, 64
    rng = np.random.default_rng(77)
    a0 = rng.uniform(-1, 1, n * n)
    b0 = rng.uniform(-1, 1, n * nrhs)
    ref, rp, rb = a0.copy(), np.zeros(n, np.int32), b0.copy()
    assert O.getrf(n, n, ref, n, rp)[0] == 0
    O.getrs(n, nrhs, ref, n, rp, rb, n)
    h = B.Handle()
    h.set_exclusive(False)
    a = torch.from_numpy(a0).cuda()
    p = torch.zeros(n, dtype=torch.int32, device="cuda")
    b = torch.from_numpy(b0).cuda()
    assert B.lib().bola_hip_dgetrf_dev(h.ptr, n, n, B._ptr(a, np.float64), n, B._ptr(p, np.int32), None) == 0
    assert B.lib().bola_hip_dgetrs_dev(h.ptr, n, nrhs, B._ptr(a, np.float64), n, B._ptr(p, np.int32),
                                       B._ptr(b, np.float64), n) == 0
    h.synchronize()
    same(p.cpu().numpy(), rp, "ipiv")
    same(a.cpu().numpy(), ref, "lu")
    same(b.cpu().numpy(), rb, "x")


@pytest.mark.parametrize("m,n,ld", [(8192, 8192, 8192), (3000, 2500, 3011), (2600, 2601, 2600), (700, 700, 705)])
def test_host_flavour_getrf_matches_the_device_flavour(B, m, n, ld):
    # the host-pointer entry point stages through pinned memory, uploads columns while the
    # factorization runs and downloads row strips as they become final: the caller's
    # matrix (including the padding rows of a larger leading dimension) and ipiv must be
    # exactly what the device flavour leaves
    import torch
    rng = np.random.default_rng(m + n + ld)
    a = rng.uniform(-1, 1, ld * n)
    pad = a.reshape(n, ld)[:, m:].copy()
    dev = torch.from_numpy(a.reshape(n, ld)[:, :m].copy().ravel()).cuda()
    dp = torch.zeros(min(m, n), dtype=torch.int32, device="cuda")
    B.getrf(m, n, dev, m, dp)
    hp = np.zeros(min(m, n), np.int32)
    B.getrf(m, n, a, ld, hp)
    same(hp, dp.cpu().numpy(), "ipiv")
    same(a.reshape(n, ld)[:, :m].ravel(), dev.cpu().numpy(), "factors")
    same(a.reshape(n, ld)[:, m:], pad, "rows beyond m untouched")


def test_host_flavour_large_operands(B, O):
    # pinned, chunked transfers for every operand of a host-pointer GEMM and GETRS
    rng = np.random.default_rng(77)
    m, n, k = 1536, 1100, 1300
    a = rng.uniform(-1, 1, m * k)
    b = rng.uniform(-1, 1, k * n)
    c = rng.uniform(-1, 1, m * n)
    ref = c.copy()
    O.gemm("N", "N", m, n, k, 1.0, a, m, b, k, 1.0, ref, m)
    B.gemm("N", "N", m, n, k, 1.0, a, m, b, k, 1.0, c, m)
    same(c, ref, "host gemm")
    n2, nrhs = 1500, 200
    lu = rng.uniform(-1, 1, n2 * n2)
    p = np.zeros(n2, np.int32)
    assert O.getrf(n2, n2, lu, n2, p)[0] == 0
    x = rng.uniform(-1, 1, n2 * nrhs)
    rx = x.copy()
    O.getrs(n2, nrhs, lu, n2, p, rx, n2)
    B.getrs(n2, nrhs, lu, n2, p, x, n2)
    same(x, rx, "host getrs")
