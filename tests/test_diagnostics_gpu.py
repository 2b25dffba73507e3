"""Error norms (ErrNorms) and the tangency dot product (SphereTangentFunctor)
on device fields, through the C ABI, against the oracle's ErrNorms and numpy.

Tolerances: the norms are sums of products of the same FP64 values in a
different order: 1e-13 relative; the error field and u.x are elementwise
FP64 and compared exactly (u.x to 1e-14 absolute for O(1) values: the
three products may be fused differently).
"""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from lpm_amd import _abi  # noqa: E402


@pytest.fixture(scope="module")
def ctx():
    from lpm_amd.device import Context
    c = Context(0)
    yield c
    c.close()


def dev(a, dtype=torch.float64):
    return torch.tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


@pytest.mark.parametrize("n", [1, 255, 257, 100_000, 1_000_003])
def test_err_norms_match_oracle(ctx, ora, n):
    rng = np.random.default_rng(n)
    exact = rng.normal(size=(3, n))
    appx = exact + 1e-3 * rng.normal(size=(3, n))
    wt = rng.uniform(0.1, 1.0, size=n)
    got = ctx.err_norms(dev(appx), dev(exact), weight=dev(wt))
    want = ora.err_norms(appx.T, exact.T, wt)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)
    # scalars, no weights
    got = ctx.err_norms(dev(appx[0]), dev(exact[0]))
    want = ora.err_norms(appx[0], exact[0])
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)


def test_err_norms_mask_error_field_and_strides(ctx, ora):
    rng = np.random.default_rng(3)
    n = 5000
    exact = rng.normal(size=(3, n))
    appx = exact + rng.normal(size=(3, n))
    mask = (rng.uniform(size=n) < 0.3).astype(np.uint8)
    # AoS buffers passed as transposed (3, n) views
    a_aos = dev(appx.T).T
    x_aos = dev(exact.T).T
    err = torch.full((3, n), 7.0, dtype=torch.float64, device="cuda")
    got = ctx.err_norms(a_aos, x_aos, mask=dev(mask, torch.uint8), err=err)
    masked_appx = np.where(mask[None, :] != 0, exact, appx)
    want = ora.err_norms(masked_appx.T, exact.T)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)
    np.testing.assert_array_equal(err.cpu().numpy(), masked_appx - exact)


def test_err_norms_grid_stride_with_mask_weights_error_field(ctx, ora):
    """Above 4096 blocks x 256 threads of the partial-sum grid: every thread
    runs several grid-stride iterations and the final pass reduces 4096
    partials; with mask, weights, the error output and AoS strides."""
    rng = np.random.default_rng(11)
    n = 3_000_017
    exact = rng.normal(size=(3, n))
    appx = exact + 1e-2 * rng.normal(size=(3, n))
    mask = (rng.uniform(size=n) < 0.2).astype(np.uint8)
    wt = rng.uniform(0.1, 1.0, size=n)
    a_aos = dev(appx.T).T
    x_aos = dev(exact.T).T
    err = torch.full((n, 3), 7.0, dtype=torch.float64, device="cuda").T
    got = ctx.err_norms(a_aos, x_aos, weight=dev(wt),
                        mask=dev(mask, torch.uint8), err=err)
    masked_appx = np.where(mask[None, :] != 0, exact, appx)
    want = ora.err_norms(masked_appx.T, exact.T, wt)
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)
    np.testing.assert_array_equal(err.cpu().numpy(), masked_appx - exact)


def test_err_norms_exact_match_and_empty(ctx):
    x = dev(np.ones((3, 10)))
    assert ctx.err_norms(x, x) == (0.0, 0.0, 0.0)
    z = torch.zeros((3, 0), dtype=torch.float64, device="cuda")
    assert ctx.err_norms(z, z) == (0.0, 0.0, 0.0)


def test_sphere_tangency(ctx):
    rng = np.random.default_rng(4)
    n = 70_001
    x = rng.normal(size=(3, n))
    x /= np.linalg.norm(x, axis=0)
    u = rng.normal(size=(3, n))
    got = ctx.sphere_tangency(dev(u), dev(x)).cpu().numpy()
    np.testing.assert_allclose(got, (u * x).sum(axis=0), rtol=0, atol=1e-14)
    # a tangent field: u = x cross a
    t = np.cross(x.T, rng.normal(size=(n, 3))).T
    got = ctx.sphere_tangency(dev(t), dev(x)).cpu().numpy()
    assert np.abs(got).max() < 1e-15 * np.abs(t).max() * 8


def test_argument_errors(ctx):
    L = _abi.lib()
    x = dev(np.ones((3, 4)))
    out = (ctypes.c_double * 3)()
    # ncomp out of range
    rc = L.lpm_err_norms(ctx.h, x.data_ptr(), 1, 4, x.data_ptr(), 1, 4, None,
                         None, 4, 4, None, 1, 1, out, None)
    assert rc != 0
    # null output
    rc = L.lpm_err_norms(ctx.h, x.data_ptr(), 1, 4, x.data_ptr(), 1, 4, None,
                         None, 4, 3, None, 1, 1, None, None)
    assert rc != 0
    rc = L.lpm_sphere_tangency(ctx.h, x.data_ptr(), 1, 4, None, 1, 4, 4,
                               x.data_ptr(), None)
    assert rc != 0
