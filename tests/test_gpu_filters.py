"""GPU parity of the post-processing filters against the CPU oracle."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("shape,k", [((32, 32, 32), 5), ((17, 23, 29), 5), ((12, 10, 9), 3),
                                     ((8, 8, 8), 4)])
def test_laplacian(ctx, shape, k):
    rng = np.random.default_rng(0)
    v = rng.random(shape).astype(np.float32)
    ref = oracle.laplacian(v, k)
    out = ctx.laplacian(torch.from_numpy(v).cuda(), k)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), ref) < 1e-5


@pytest.mark.parametrize("shape,k,sigma", [((32, 32, 32), 5, 1.0), ((19, 21, 33), 5, 1.3),
                                           ((16, 16, 16), 7, 0.8), ((16, 16, 16), 5, 0.2),
                                           ((10, 12, 14), 4, 1.0)])
def test_log(ctx, shape, k, sigma):
    """Separable path (and dense path for sigma = 0.2, where taps are
    thresholded away, and even k) against the dense FP64 oracle."""
    rng = np.random.default_rng(1)
    v = rng.random(shape).astype(np.float32)
    w = np.empty(k ** 3, np.float32)
    from fasterbackprojection_amd import _lib
    import ctypes
    _lib.load().fbp_log_kernel_host(k, sigma, w.ctypes.data_as(ctypes.c_void_p))
    ref = oracle.log_filter(v, k, sigma, weights=w.astype(np.float64).reshape(k, k, k))
    out = ctx.log(torch.from_numpy(v).cuda(), k, sigma)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), ref) < 1e-4


def test_filters_at_the_timed_sizes(ctx):
    """The sizes the benchmark times: the separable LoG (k = 5, sigma = 1)
    on a 128^3 volume, whose fused kernel walks several z-chunks, and the
    Laplacian (k = 5) on 64^3, each against the FP64 oracle."""
    rng = np.random.default_rng(4)
    v = (rng.random((128, 128, 128)) ** 3).astype(np.float32)
    w = np.empty(125, np.float32)
    from fasterbackprojection_amd import _lib
    import ctypes
    _lib.load().fbp_log_kernel_host(5, 1.0, w.ctypes.data_as(ctypes.c_void_p))
    ref = oracle.log_filter(v, 5, 1.0, weights=w.astype(np.float64).reshape(5, 5, 5))
    out = ctx.log(torch.from_numpy(v).cuda(), 5, 1.0).cpu().numpy()
    assert rel_l2(out, ref) < 1e-4
    # a plane in each z-chunk boundary region and the borders, voxel by voxel
    for z in (0, 1, 2, 31, 32, 33, 63, 64, 95, 96, 126, 127):
        np.testing.assert_allclose(out[z], ref[z], rtol=0, atol=2e-5 * np.abs(ref).max())
    v64 = rng.random((64, 64, 64)).astype(np.float32)
    ref64 = oracle.laplacian(v64, 5)
    out64 = ctx.laplacian(torch.from_numpy(v64).cuda(), 5).cpu().numpy()
    assert rel_l2(out64, ref64) < 1e-5
    np.testing.assert_allclose(out64, ref64, rtol=0, atol=1e-4 * np.abs(ref64).max())


def test_normalize(ctx):
    rng = np.random.default_rng(2)
    v = (rng.random((20, 30, 40)) * 7 - 3).astype(np.float32)
    ref = oracle.normalize(v)
    t = torch.from_numpy(v).cuda()
    ctx.normalize(t)
    torch.cuda.synchronize()
    np.testing.assert_allclose(t.cpu().numpy(), ref, rtol=1e-6, atol=1e-7)
    mn, mx = ctx.minmax(t)
    assert mn == 0.0 and abs(mx - 1.0) < 1e-6


@pytest.mark.parametrize("shape,k,sigma", [((32, 32, 32), 5, 1.0), ((19, 21, 33), 5, 1.3),
                                           ((128, 128, 128), 5, 1.0), ((20, 24, 28), 15, 0.6)])
def test_log_normalize_equals_the_two_calls(ctx, shape, k, sigma):
    """fbp_filter_log_normalize (the LoG kernel reducing its own min / max)
    gives bit for bit what fbp_filter_log then fbp_normalize_minmax give;
    the last case (sigma 0.6, k 15) takes the dense LoG path."""
    rng = np.random.default_rng(sum(shape) + k)
    v = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).cuda()
    two = ctx.log(v, k, sigma)
    ctx.normalize(two)
    one = ctx.log_normalize(v, k, sigma)
    torch.cuda.synchronize()
    assert torch.equal(one, two)
    # a second call reuses the arrival counter
    one2 = ctx.log_normalize(v * 2.0, k, sigma)
    two2 = ctx.log(v * 2.0, k, sigma)
    ctx.normalize(two2)
    assert torch.equal(one2, two2)
