"""The no-allocation LU entry point (no reference counterpart; for callers that cannot
allocate device memory): through the C ABI, on two calls in a row, the second smaller,
it must give exactly what bola_hip_dgetrf_dev gives (the factors, pivots and status bit
for bit), and it must refuse a factorization whose scratch exceeds its static array."""
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


def _pair(B, m, n, seed, zero_col=None):
    import torch
    rng = np.random.default_rng(seed)
    a0 = rng.uniform(-1, 1, (m, n))
    if zero_col is not None:
        a0[:, zero_col] = 0.0
    a0 = a0.ravel(order="F").copy()
    k = min(m, n)
    ref = torch.from_numpy(a0.copy()).cuda()
    rp = torch.full((k,), -5, dtype=torch.int32, device="cuda")
    got = torch.from_numpy(a0.copy()).cuda()
    gp = torch.full((k,), -5, dtype=torch.int32, device="cuda")
    err_ref = err_got = None
    try:
        B.getrf(m, n, ref, m, rp)
    except B.BolaError as e:
        err_ref = (e.code, e.message)
    try:
        B.getrf_noalloc(m, n, got, m, gp)
    except B.BolaError as e:
        err_got = (e.code, e.message)
    torch.cuda.synchronize()
    assert err_got == err_ref
    assert np.array_equal(gp.cpu().numpy(), rp.cpu().numpy()), (m, n)
    assert np.array_equal(got.cpu().numpy(), ref.cpu().numpy()), (m, n)


def test_noalloc_getrf_matches_getrf_dev_on_two_calls(B):
    # 256-column super-panels on three streams, then a smaller one in the same scratch
    _pair(B, 2600, 2600, 1)
    _pair(B, 600, 600, 2)


def test_noalloc_getrf_shapes_and_a_zero_pivot(B):
    _pair(B, 4096, 4096, 3)
    _pair(B, 1200, 800, 4)
    _pair(B, 100, 300, 5)       # the reference's own order (m % 8 != 0)
    _pair(B, 20, 20, 6)         # the unblocked path
    _pair(B, 600, 600, 7, zero_col=293)


def test_noalloc_getrf_refuses_more_than_its_scratch(B):
    import torch
    a = torch.zeros(16, dtype=torch.float64, device="cuda")
    p = torch.zeros(16, dtype=torch.int32, device="cuda")
    with pytest.raises(B.BolaError) as e:
        B.getrf_noalloc(65536, 65536, a, 65536, p)  # refused before any device work
    assert e.value.code == -2 and "scratch" in e.value.message
    assert float(a.abs().sum()) == 0.0
