"""Pins the CPU oracle (oracle/bola_oracle.c) before anything is checked against it.

Each routine is checked against (a) the known-answer tests the reference crate's own
tests hold (tests/golden/kats.json), and (b) checks the mathematics fixes independently
of the oracle's code: exact integer arithmetic, LAPACK through scipy, brute force on
tiny inputs, correctly rounded fused multiply-adds computed with exact rationals.
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest
import scipy.linalg as sla

EPS = np.finfo(np.float64).eps
HERE = os.path.dirname(os.path.abspath(__file__))
KATS = json.load(open(os.path.join(HERE, "golden", "kats.json")))


def fma(a, b, c):
    """Correctly rounded a*b+c via exact rationals (independent of any FMA hardware)."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def arr(x):
    return np.array(x, dtype=np.float64)


# ----------------------------------------------------------------- reference KATs

@pytest.mark.parametrize("name", ["gemm_nn", "gemm_nn_dup", "gemm_nt"])
def test_kat_gemm(oracle, name):
    t = KATS[name]
    c = arr(t["c"])
    oracle.gemm(t["transa"], t["transb"], t["m"], t["n"], t["k"], t["alpha"], arr(t["a"]), t["lda"], arr(t["b"]),
                t["ldb"], t["beta"], c, t["ldc"])
    assert np.all(np.abs(c - arr(t["expected"])) < t["tol"])


def test_kat_laswp(oracle):
    t = KATS["laswp"]
    a = arr(t["a"])
    oracle.laswp(t["n"], a, t["lda"], t["k1"], t["k2"], np.array(t["ipiv"], np.int32), t["incx"])
    assert np.all(np.abs(a - arr(t["expected"])) < t["tol"])


def test_kat_scal(oracle):
    t = KATS["scal"]
    x = arr(t["x"])
    done = oracle.scal_kernel(len(x), t["alpha"], x)
    for i in range(done, len(x)):
        x[i] *= t["alpha"]
    assert done == 4
    assert list(x) == t["expected"]


@pytest.mark.parametrize("name", ["syrk_ln", "syrk_ut"])
def test_kat_syrk(oracle, name):
    t = KATS[name]
    c = arr(t["c"])
    n = t["n"]
    oracle.syrk(t["uplo"], t["trans"], n, t["k"], t["alpha"], arr(t["a"]), t["lda"], t["beta"], c, t["ldc"])
    tri = np.zeros_like(c)
    for j in range(n):
        rows = range(j, n) if t["uplo"] == "L" else range(0, j + 1)
        for i in rows:
            tri[i + j * t["ldc"]] = c[i + j * t["ldc"]]
    assert np.all(np.abs(tri - arr(t["expected_triangle"])) < t["tol"])


@pytest.mark.parametrize("name", ["trsm_llnn", "trsm_rutn", "trsm_rltn"])
def test_kat_trsm(oracle, name):
    t = KATS[name]
    b = arr(t["b"])
    oracle.trsm(t["side"], t["uplo"], t["transa"], t["diag"], t["m"], t["n"], t["alpha"], arr(t["a"]), t["lda"], b,
                t["ldb"])
    assert np.all(np.abs(b - arr(t["expected"])) < t["tol"])


@pytest.mark.parametrize("name", ["potrf_u", "potrf_l"])
def test_kat_potrf(oracle, name):
    t = KATS[name]
    a = arr(t["a"])
    n = t["n"]
    code, msg = oracle.potrf(t["uplo"], n, a, t["lda"])
    assert code == 0, msg
    tri = np.zeros_like(a)
    for j in range(n):
        rows = range(j, n) if t["uplo"] == "L" else range(0, j + 1)
        for i in rows:
            tri[i + j * n] = a[i + j * n]
    assert np.all(np.abs(tri - arr(t["expected_triangle"])) < t["tol"])


def test_kat_getrf(oracle):
    t = KATS["getrf"]
    a = arr(t["a"])
    ipiv = np.zeros(3, np.int32)
    code, msg = oracle.getrf(t["m"], t["n"], a, t["lda"], ipiv)
    assert code == 0, msg
    assert np.all(np.abs(a - arr(t["expected"])) < t["tol"])
    assert list(ipiv) == t["ipiv_expected"]


# ----------------------------------------------------------------- IDAMAX rule

def idamax_rule(x):
    """The x86_64 idamax result as a closed-form selection rule (not a simulation of its
    accumulators): max |x| over non-NaN entries; ties among the first T = 16*floor(n/16)
    positions go to the smallest key (i mod 4, (i mod 16) div 4, i); a tail position wins
    only if strictly greater than the vector-phase maximum, first occurrence first; a
    zero vector-phase maximum reports index 0.  Below 16 elements: first strict maximum."""
    n = len(x)
    ax = np.abs(x)
    if n < 16:
        best, idx = -1.0, 0
        for i in range(n):
            if ax[i] > best:
                best, idx = ax[i], i
        return idx
    T = 16 * (n // 16)
    vec = [(i % 4, (i % 16) // 4, i) for i in range(T) if not np.isnan(ax[i])]
    vmax = max((ax[k[2]] for k in vec), default=0.0)
    if vmax > 0:
        vidx = min(k for k in vec if ax[k[2]] == vmax)[2]
    else:
        vmax, vidx = 0.0, 0
    tail = [i for i in range(T, n) if not np.isnan(ax[i])]
    tmax = max((ax[i] for i in tail), default=-1.0)
    if tmax > vmax:
        return min(i for i in tail if ax[i] == tmax)
    return vidx


def test_idamax_unique_max_matches_argmax(oracle):
    rng = np.random.default_rng(1)
    for n in list(range(0, 70)) + [127, 128, 129, 1000, 4099]:
        x = rng.uniform(-1, 1, n)
        got = oracle.idamax(n, x) if n else oracle.idamax(0, np.zeros(1))
        want = int(np.argmax(np.abs(x))) if n else 0
        assert got == want, (n, got, want)


def test_idamax_ties_nan_rule(oracle):
    rng = np.random.default_rng(2)
    for trial in range(3000):
        n = int(rng.integers(1, 80))
        # few distinct magnitudes -> many ties; sprinkle NaN, zeros, signs
        x = rng.integers(0, 4, n).astype(np.float64) * rng.choice([-1.0, 1.0], n)
        if trial % 3 == 0:
            x[rng.random(n) < 0.2] = np.nan
        if trial % 7 == 0:
            x[:] = 0.0
        if trial % 11 == 0:
            x[:] = np.nan
        assert oracle.idamax(n, x) == idamax_rule(x), (trial, n, x)


def test_idamax_tie_key_examples(oracle):
    # Equal maxima at 1 and 4: key (1,0,1) vs (0,1,4) -> position 4 wins (not first index).
    x = np.zeros(16)
    x[1] = 5.0
    x[4] = -5.0
    assert oracle.idamax(16, x) == 4
    # Tail never beats an equal vector-phase maximum.
    x = np.zeros(17)
    x[3] = 2.0
    x[16] = 2.0
    assert oracle.idamax(17, x) == 3
    # All-NaN vector phase with a finite tail element: tail wins (> 0).
    x = np.full(18, np.nan)
    x[17] = 0.5
    assert oracle.idamax(18, x) == 17


# ----------------------------------------------------------------- LASWP

def laswp_perm(nrows, k1, k2, ipiv, incx):
    order = list(range(k1, k2 + 1))
    pivs = list(ipiv)
    if incx < 0:
        order = order[::-1]
        pivs_used = [pivs[(k2 - k1) - t] for t in range(k2 - k1 + 1)]
    else:
        pivs_used = [pivs[t] for t in range(k2 - k1 + 1)]
    rows = list(range(nrows))
    for i, p in zip(order, pivs_used):
        rows[i - 1], rows[p - 1] = rows[p - 1], rows[i - 1]
    return rows


@pytest.mark.parametrize("ncols", [1, 3, 4, 7, 255, 256, 300])
@pytest.mark.parametrize("incx", [1, -1])
def test_laswp_is_composition_of_transpositions(oracle, ncols, incx):
    rng = np.random.default_rng(ncols * 10 + (incx > 0))
    m = 50
    lda = 53
    k1, k2 = 3, 20
    ipiv = np.zeros(k2, np.int32)
    for i in range(k2):
        ipiv[i] = rng.integers(i + 1, m + 1)
    a = rng.uniform(-1, 1, lda * ncols)
    ref = a.copy().reshape(ncols, lda).T
    if incx > 0:
        piv = ipiv[k1 - 1:k2]
    else:
        piv = ipiv[k1 - 1:k2]
    rows = laswp_perm(m, k1, k2, piv if incx > 0 else piv, incx)
    expect = ref.copy()
    expect[:m, :] = ref[rows, :]
    oracle.laswp(ncols, a, lda, k1, k2, ipiv, incx)
    assert np.array_equal(a.reshape(ncols, lda).T, expect)


# ----------------------------------------------------------------- BLAS-1/2 numerics

def test_daxpy_fma_placement(oracle):
    rng = np.random.default_rng(3)
    n = 11
    x = rng.uniform(-1, 1, n)
    y = rng.uniform(-1, 1, n)
    alpha = 0.1 + 1e-9
    got = y.copy()
    oracle.daxpy(n, alpha, x, 1, got, 1)
    want = [fma(alpha, x[i], y[i]) if i < 8 else y[i] + alpha * x[i] for i in range(n)]
    assert list(got) == want


def test_ger_fma_placement_and_skip(oracle):
    rng = np.random.default_rng(4)
    m, n, lda = 7, 5, 9
    x = rng.uniform(-1, 1, m)
    y = rng.uniform(-1, 1, 2 * n)
    y[0] = y[2] = 0.0  # column pair (0,1) is not skipped (y[2] is column 1 with ldy=2)
    a = rng.uniform(-1, 1, lda * n)
    got = a.copy()
    oracle.ger(m, n, -1.0, x, y, 2, got, lda)
    want = a.copy()
    for j in range(n):
        t = -1.0 * y[j * 2]
        pair_skip = (j < n - n % 2) and (-1.0 * y[(j - j % 2) * 2] == 0.0 and -1.0 * y[(j - j % 2 + 1) * 2] == 0.0)
        if pair_skip or (j == n - 1 and n % 2 == 1 and t == 0.0):
            continue
        for i in range(m):
            fused = (j < n - n % 2) and i < 4 * (m // 4)
            want[i + j * lda] = fma(x[i], t, a[i + j * lda]) if fused else a[i + j * lda] + t * x[i]
    assert list(got) == list(want)


def test_dot_product_order(oracle):
    rng = np.random.default_rng(5)
    n = 10
    x = rng.uniform(-1, 1, n)
    y = rng.uniform(-1, 1, n)
    lanes = [0.0] * 4
    for c in range(n // 4):
        for l in range(4):
            lanes[l] = fma(x[4 * c + l], y[4 * c + l], lanes[l])
    s = 0.0 + (((lanes[0] + lanes[1]) + lanes[2]) + lanes[3])
    for i in range(8, n):
        s = s + x[i] * y[i]
    assert oracle.dot_product(n, x, 1, y, 1) == s
    assert oracle.dot_product(0, x, 1, y, 1) == 0.0


# ----------------------------------------------------------------- GEMM

@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(1, 1, 1), (7, 5, 3), (33, 17, 65), (130, 140, 70)])
def test_gemm_integer_exact(oracle, ta, tb, m, n, k):
    rng = np.random.default_rng(m * 1000 + n * 10 + k)
    A = rng.integers(-8, 9, (m, k)).astype(np.float64)
    B = rng.integers(-8, 9, (k, n)).astype(np.float64)
    Cm = rng.integers(-8, 9, (m, n)).astype(np.float64)
    Aop = A if ta == "N" else A.T.copy()
    Bop = B if tb == "N" else B.T.copy()
    lda = Aop.shape[0] + 2
    ldb = Bop.shape[0] + 1
    ldc = m + 3
    a = np.zeros(lda * Aop.shape[1]); a.reshape(Aop.shape[1], lda).T[:Aop.shape[0]] = Aop
    b = np.zeros(ldb * Bop.shape[1]); b.reshape(Bop.shape[1], ldb).T[:Bop.shape[0]] = Bop
    c = np.zeros(ldc * n); c.reshape(n, ldc).T[:m] = Cm
    oracle.gemm(ta, tb, m, n, k, 2.0, a, lda, b, ldb, -1.0, c, ldc)
    assert np.array_equal(c.reshape(n, ldc).T[:m], 2.0 * A @ B - Cm)


def test_gemm_random_vs_numpy_and_fma_structure(oracle):
    rng = np.random.default_rng(6)
    m, n, k = 200, 150, 130
    A = rng.random((m, k)); B = rng.random((k, n)); Cm = rng.random((m, n))
    a = np.asfortranarray(A).ravel(order="F").copy()
    b = np.asfortranarray(B).ravel(order="F").copy()
    c = np.asfortranarray(Cm).ravel(order="F").copy()
    oracle.gemm("N", "N", m, n, k, 1.0, a, m, b, k, 1.0, c, m)
    ref = A @ B + Cm
    assert np.max(np.abs(c.reshape(n, m).T - ref) / ref) <= k * EPS
    # Interior element: per 64-slab acc = fma chain from 0, then C = fma(alpha, acc, C).
    i, j = 5, 7
    cc = Cm[i, j]
    for s in range(0, k, 64):
        acc = 0.0
        for l in range(s, min(k, s + 64)):
            acc = fma(A[i, l], B[l, j], acc)
        cc = fma(1.0, acc, cc)
    assert c[i + j * m] == cc


def test_gemm_edge_elements_unfused(oracle):
    # m % 8 != 0 -> the last m % 8 rows go through the unfused scalar kernel.
    rng = np.random.default_rng(7)
    m, n, k = 13, 4, 6
    A = rng.random((m, k)); B = rng.random((k, n)); Cm = rng.random((m, n))
    a = A.ravel(order="F").copy(); b = B.ravel(order="F").copy(); c = Cm.ravel(order="F").copy()
    oracle.gemm("N", "N", m, n, k, 0.5, a, m, b, k, 1.0, c, m)
    i, j = 12, 2
    cc = Cm[i, j]
    for l in range(k):
        cc = cc + (0.5 * B[l, j]) * A[i, l]
    assert c[i + j * m] == cc


def test_gemm_beta_alpha_rules(oracle):
    c = np.array([np.nan, 1.0])
    oracle.gemm("N", "N", 2, 1, 1, 0.0, np.ones(2), 2, np.ones(1), 1, 0.0, c, 2)
    assert list(c) == [0.0, 0.0]  # beta == 0 overwrites NaN; alpha == 0 returns
    c = np.array([1.0, 2.0])
    oracle.gemm("N", "N", 2, 1, 1, 0.0, np.ones(2), 2, np.ones(1), 1, 3.0, c, 2)
    assert list(c) == [3.0, 6.0]


# ----------------------------------------------------------------- TRSM

def _tri(rng, n, lower, unit):
    T = rng.uniform(-1, 1, (n, n)) / n
    T = np.tril(T) if lower else np.triu(T)
    np.fill_diagonal(T, 1.0 if unit else rng.uniform(1, 2, n))
    return T


@pytest.mark.parametrize("side", ["L", "R"])
@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["N", "U"])
@pytest.mark.parametrize("size", [5, 70, 150])
def test_trsm_all_variants_vs_scipy(oracle, side, uplo, trans, diag, size):
    rng = np.random.default_rng(hash((side, uplo, trans, diag, size)) % 2**32)
    m, n = (size, 9) if side == "L" else (9, size)
    na = m if side == "L" else n
    T = _tri(rng, na, uplo == "L", diag == "U")
    # garbage in the other triangle / unit diagonal must be ignored
    stored = T.copy()
    if uplo == "L":
        stored[np.triu_indices(na, 1)] = 7.0
    else:
        stored[np.tril_indices(na, -1)] = 7.0
    if diag == "U":
        np.fill_diagonal(stored, 3.0)
    X = rng.integers(-5, 6, (m, n)).astype(np.float64)
    op = T if trans == "N" else T.T
    Bm = op @ X if side == "L" else X @ op
    alpha = 2.0
    b = (Bm / alpha).ravel(order="F").copy()
    lda = na + 1
    a = np.zeros(lda * na); a.reshape(na, lda).T[:na] = stored
    oracle.trsm(side, uplo, trans, diag, m, n, alpha, a, lda, b, m)
    got = b.reshape(n, m).T
    assert np.max(np.abs(got - X)) < 1e-10 * max(1, na)


# ----------------------------------------------------------------- SYRK

@pytest.mark.parametrize("uplo,trans", [("L", "N"), ("U", "T")])
@pytest.mark.parametrize("n,k", [(3, 2), (33, 40), (100, 70)])
def test_syrk_supported_combos_vs_numpy(oracle, uplo, trans, n, k):
    rng = np.random.default_rng(n * 7 + k)
    A = rng.integers(-4, 5, (n, k) if trans == "N" else (k, n)).astype(np.float64)
    Cm = rng.integers(-4, 5, (n, n)).astype(np.float64)
    lda = A.shape[0]
    c = Cm.ravel(order="F").copy()
    oracle.syrk(uplo, trans, n, k, -1.0, A.ravel(order="F").copy(), lda, 2.0, c, n)
    full = -(A @ A.T if trans == "N" else A.T @ A) + 2.0 * Cm
    got = c.reshape(n, n).T
    mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    assert np.array_equal(got[mask], full[mask])
    assert np.array_equal(got[~mask], Cm[~mask])  # other triangle untouched


def test_syrk_quirk_un_lt_small_sizes_still_blas(oracle):
    # The (U,N) and (L,T) paths agree with BLAS while n, k <= 32 (one block, no offset).
    rng = np.random.default_rng(8)
    for uplo, trans in [("U", "N"), ("L", "T")]:
        n, k = 20, 30
        A = rng.integers(-3, 4, (n, k) if trans == "N" else (k, n)).astype(np.float64)
        c = np.zeros(n * n)
        oracle.syrk(uplo, trans, n, k, 1.0, A.ravel(order="F").copy(), A.shape[0], 0.0, c, n)
        full = A @ A.T if trans == "N" else A.T @ A
        mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
        assert np.array_equal(c.reshape(n, n).T[mask], full[mask])


# ----------------------------------------------------------------- GETRF

def _lu_backward_error(A, lu, ipiv):
    n = A.shape[0]
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    PA = A.copy()
    for i, p in enumerate(ipiv):
        PA[[i, p - 1]] = PA[[p - 1, i]]
    return np.linalg.norm(PA - L @ U, 1) / (np.linalg.norm(A, 1) * n * EPS)


@pytest.mark.parametrize("n", [3, 31, 32, 33, 100, 257])
def test_getrf_vs_lapack(oracle, n):
    rng = np.random.default_rng(n)
    A = rng.uniform(-1, 1, (n, n))
    a = A.ravel(order="F").copy()
    ipiv = np.zeros(n, np.int32)
    code, msg = oracle.getrf(n, n, a, n, ipiv)
    assert code == 0, msg
    lu_ref, piv_ref = sla.lu_factor(A)
    assert np.array_equal(ipiv, piv_ref + 1)
    lu = a.reshape(n, n).T
    assert np.max(np.abs(lu - lu_ref)) < 1e-10 * n
    assert _lu_backward_error(A, lu, ipiv) <= 10


@pytest.mark.parametrize("m,n", [(50, 20), (20, 50), (100, 70), (70, 100)])
def test_getrf_rectangular(oracle, m, n):
    rng = np.random.default_rng(m * n)
    A = rng.uniform(-1, 1, (m, n))
    a = A.ravel(order="F").copy()
    ipiv = np.zeros(min(m, n), np.int32)
    code, msg = oracle.getrf(m, n, a, m, ipiv)
    assert code == 0, msg
    lu_ref, piv_ref = sla.lu_factor(A) if m == n else (None, None)
    P, L, U = sla.lu(A)
    lu = a.reshape(n, m).T
    Lg = np.tril(lu, -1)[:, :min(m, n)] + np.eye(m, min(m, n))
    Ug = np.triu(lu)[:min(m, n), :]
    PA = A.copy()
    for i, p in enumerate(ipiv):
        PA[[i, p - 1]] = PA[[p - 1, i]]
    assert np.linalg.norm(PA - Lg @ Ug) <= 10 * max(m, n) * EPS * np.linalg.norm(A)


def test_getrf2_recursive_branch(oracle):
    rng = np.random.default_rng(9)
    n = 80
    A = rng.uniform(-1, 1, (n, n))
    a = A.ravel(order="F").copy()
    ipiv = np.zeros(n, np.int32)
    code, _ = oracle.getrf2(n, n, a, n, ipiv)
    assert code == 0
    _, piv_ref = sla.lu_factor(A)
    assert np.array_equal(ipiv, piv_ref + 1)
    assert _lu_backward_error(A, a.reshape(n, n).T, ipiv) <= 10


def test_getrf_errors(oracle):
    code, msg = oracle.getrf(4, 4, np.zeros(16), 3, np.zeros(4, np.int32))
    assert code == -4 and msg == "Argument 4 to getrf had an illegal value of 3"
    code, msg = oracle.getrf2(4, 4, np.zeros(16), 3, np.zeros(4, np.int32))
    assert msg == "Argument 4 to getrf2 had an illegal value of 3"
    # zero column 40 (0-based) of a 64x64 matrix: reported local to its 32-wide panel
    n = 64
    rng = np.random.default_rng(10)
    A = rng.uniform(-1, 1, (n, n))
    A[:, 40] = 0.0
    code, msg = oracle.getrf(n, n, A.ravel(order="F").copy(), n, np.zeros(n, np.int32))
    assert code == 9 and msg == "Matrix is singular. The pivot in column 9 is zero."


# ----------------------------------------------------------------- POTRF

@pytest.mark.parametrize("uplo", ["L", "U"])
@pytest.mark.parametrize("n", [1, 5, 32, 33, 100, 200])
def test_potrf_vs_lapack(oracle, uplo, n):
    rng = np.random.default_rng(n + (uplo == "U"))
    G = rng.uniform(-1, 1, (n, n))
    A = G @ G.T + n * np.eye(n)
    a = A.ravel(order="F").copy()
    code, msg = oracle.potrf(uplo, n, a, n)
    assert code == 0, msg
    F = a.reshape(n, n).T
    ref = sla.cholesky(A, lower=(uplo == "L"))
    mask = np.tril(np.ones((n, n), bool)) if uplo == "L" else np.triu(np.ones((n, n), bool))
    assert np.max(np.abs(F[mask] - ref[mask])) < 1e-12 * n
    Ft = np.where(mask, F, 0.0)
    R = A - (Ft @ Ft.T if uplo == "L" else Ft.T @ Ft)
    assert np.linalg.norm(R, 1) / (np.linalg.norm(A, 1) * n * EPS) <= 10
    assert np.array_equal(F[~mask], A[~mask])  # other triangle untouched


def test_potrf_errors(oracle):
    code, msg = oracle.potrf("X", 3, np.eye(3).ravel(), 3)
    assert code == -1 and msg == "Argument 1 to potrf had an illegal value"
    code, msg = oracle.potrf("l", 3, np.eye(3).ravel(), 3)  # case-sensitive
    assert code == -1
    code, msg = oracle.potrf("L", 3, np.eye(3).ravel(), 2)
    assert code == -4 and msg == "Argument 4 to potrf had an illegal value"
    # not SPD at global column 37 (0-based 36): reported local to its 32-wide block -> 5
    n = 64
    A = np.eye(n)
    A[36, 36] = -1.0
    code, msg = oracle.potrf("L", n, A.ravel(order="F").copy(), n)
    assert code == 5 and msg == "Matrix is not positive-definite. Failure at column 5."
    A = np.eye(3); A[1, 1] = np.nan
    code, msg = oracle.potrf("U", 3, A.ravel(order="F").copy(), 3)
    assert code == 2


# ----------------------------------------------------------------- solve compositions

def test_getrs_potrs_solve(oracle):
    rng = np.random.default_rng(11)
    n, nrhs = 150, 7
    A = rng.uniform(-1, 1, (n, n))
    X = rng.uniform(-1, 1, (n, nrhs))
    B = A @ X
    lu = A.ravel(order="F").copy()
    ipiv = np.zeros(n, np.int32)
    assert oracle.getrf(n, n, lu, n, ipiv)[0] == 0
    b = B.ravel(order="F").copy()
    oracle.getrs(n, nrhs, lu, n, ipiv, b, n)
    assert np.allclose(b.reshape(nrhs, n).T, np.linalg.solve(A, B), atol=1e-10)
    S = A @ A.T + n * np.eye(n)
    B = S @ X
    for uplo in "LU":
        f = S.ravel(order="F").copy()
        assert oracle.potrf(uplo, n, f, n)[0] == 0
        b = B.ravel(order="F").copy()
        oracle.potrs(uplo, n, nrhs, f, n, b, n)
        assert np.allclose(b.reshape(nrhs, n).T, X, atol=1e-12)


def test_thread_count_independence(oracle):
    rng = np.random.default_rng(12)
    n = 300
    A = rng.uniform(-1, 1, (n, n)).ravel(order="F")
    res = []
    saved = oracle.num_threads()
    try:
        for t in (1, 3):
            oracle.set_num_threads(t)
            a = A.copy()
            ipiv = np.zeros(n, np.int32)
            oracle.getrf(n, n, a, n, ipiv)
            res.append((a, ipiv))
    finally:
        oracle.set_num_threads(saved)
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
