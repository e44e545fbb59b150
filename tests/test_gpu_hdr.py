"""GPU parity of the tone-mapping kernels (gdp_tm_*, gdp_percentile) against the oracle
(oracle/hdr.py).

Bit-exact: the binomial downsample, the attenuated divergence and the percentile's order
statistics.  Round-off: the log luminance (numpy's matmul may fuse), the factor (means are
summed in another order; pow and log are not correctly rounded).  The whole tone map
differs from the oracle's exact DCT integration by the solver tolerance only.
"""

import ctypes

import numpy as np
import pytest

from oracle import generators as G
from oracle import hdr as H

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gd():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gdpoisson

    return gdpoisson


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("shape", [(5, 7), (64, 96), (37, 50)])
def test_log_luminance_matches_oracle(gd, shape):
    import torch

    from gdpoisson import _lib

    rad = G.radiance_map(shape, seed=1)
    t = _dev(rad)
    out = torch.empty(shape, dtype=torch.float64, device="cuda")
    assert _lib.lib().gdp_tm_log_luminance(_p(t), shape[0], shape[1], _p(out), None) == 0
    assert rel(out.cpu().numpy(), H.log_luminance(rad)) < 1e-14


@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (37, 50), (64, 64), (17, 129)])
def test_downsample_is_bit_exact(gd, shape):
    import torch

    from gdpoisson import _lib

    x = np.random.default_rng(2).standard_normal(shape)
    out = torch.empty(((shape[0] + 1) // 2, (shape[1] + 1) // 2), dtype=torch.float64,
                      device="cuda")
    assert _lib.lib().gdp_tm_downsample(_p(_dev(x)), shape[0], shape[1], _p(out), None) == 0
    assert np.array_equal(out.cpu().numpy(), H.downsample(x))


@pytest.mark.parametrize("shape", [(2, 2), (31, 45), (64, 80)])
def test_attenuated_rhs_is_bit_exact(gd, shape):
    import torch

    from gdpoisson import _lib
    from oracle.operators import divergence, gradient

    rng = np.random.default_rng(3)
    L, f = rng.standard_normal(shape), rng.random(shape) + 0.5
    out = torch.empty(shape, dtype=torch.float64, device="cuda")
    assert _lib.lib().gdp_tm_rhs(_p(_dev(L)), _p(_dev(f)), shape[0], shape[1], _p(out),
                                 None) == 0
    gx, gy = gradient(L)
    assert np.array_equal(out.cpu().numpy(), divergence(gx * f, gy * f))


@pytest.mark.parametrize("levels", [1, 3, 8])
def test_attenuation_field_matches_oracle(gd, levels):
    L = H.log_luminance(G.radiance_map((150, 211), seed=4))
    cfg = gd.hdr.ToneMapConfig(levels=levels)
    got = gd.hdr.attenuation_field(L, cfg)
    want = H.attenuation_field(L, H.ToneMapConfig(levels=levels))
    assert got.shape == want.shape
    assert rel(got, want) < 1e-12


def test_flat_image_factor_on_the_gpu(gd):
    cfg = gd.hdr.ToneMapConfig()
    f = gd.hdr.attenuation_field(np.zeros((64, 64)), cfg)
    assert np.isfinite(f).all() and f.min() > 0.0
    assert np.allclose(f, cfg.alpha ** ((1.0 - cfg.beta) * 3), rtol=1e-12, atol=0)


@pytest.mark.parametrize("n", [1, 2, 3, 10, 1001, 300_007])
@pytest.mark.parametrize("q", [0.0, 1.0, 50.0, 99.5, 100.0])
def test_percentile_matches_numpy_exactly(gd, n, q):
    rng = np.random.default_rng(n)
    x = np.round(rng.standard_normal(n) * 3, 1)  # ties, negatives, signed zeros
    assert gd.hdr.percentile(x, q) == np.percentile(x, q)


def test_percentile_of_positive_wide_range(gd):
    x = np.exp(np.random.default_rng(9).normal(0.0, 4.0, 1_000_003))
    assert gd.hdr.percentile(x, 99.5) == np.percentile(x, 99.5)


@pytest.mark.parametrize("method", ["dct", "mgcg", "mgcg_mixed"])
def test_tone_map_matches_oracle(gd, method):
    rad = G.radiance_map((96, 128), seed=2)
    reports = []
    img, f = gd.tone_map(rad, method=method, reports=reports)
    want, wf = H.tone_map(rad)
    assert len(reports) == 1 and reports[0].converged
    assert reports[0].relative_residual <= (1e-12 if method == "dct" else 1e-8)
    assert img.shape == rad.shape and img.min() >= 0.0 and img.max() <= 1.0
    assert rel(f, wf) < 1e-12
    # the exact DCT integration agrees with the oracle's to round-off
    assert np.abs(img - want).max() < (1e-10 if method == "dct" else 1e-6)


def test_tone_map_rejects_a_grayscale_input(gd):
    with pytest.raises(ValueError):
        gd.tone_map(np.ones((16, 16)))


def test_tone_map_of_a_serving_request(gd):
    """3840 x 2160 log-normal radiance (the serving stream's tone-map request), mixed
    precision, against the oracle's exact DCT integration."""
    import torch

    rad = G.lognormal_radiance((2160, 3840), seed=5003)
    reports = []
    img, f = gd.tone_map(torch.from_numpy(rad).cuda(), method="mgcg_mixed", reports=reports)
    assert reports[0].converged and reports[0].relative_residual <= 1e-8
    want, _ = H.tone_map(rad)
    assert torch.abs(img.cpu() - torch.from_numpy(want)).max().item() < 1e-6


@pytest.mark.parametrize("shape", [(1, 1), (1, 6), (7, 1), (2, 3), (5, 17), (17, 5), (33, 64)])
def test_tone_map_of_small_and_odd_images(gd, shape):
    """Tiny and odd-sized radiance maps (pyramids of one or two levels, one-pixel
    solves) against the oracle."""
    rad = G.radiance_map(shape, seed=7)
    img, f = gd.tone_map(rad)
    want, wf = H.tone_map(rad)
    assert img.shape == want.shape
    assert rel(f, wf) < 1e-12
    assert np.abs(img - want).max() < 1e-10
