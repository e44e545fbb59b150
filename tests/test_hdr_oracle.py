"""The tone-mapping oracle (oracle/hdr.py) pinned by closed forms, by numpy's own
percentile and by the invariants of gradient-domain/tests/test_applications.py:178-206."""

import numpy as np
import pytest

from oracle import generators as G
from oracle import hdr as H
from oracle.solvers import solve_neumann


def _decades(image):
    # gradient-domain/tests/test_applications.py:36-40
    lum = np.maximum(image @ np.array(H.LUMA), 1e-8)
    lo, hi = np.percentile(lum, [1, 99])
    return float(np.log10(max(hi, 1e-8) / max(lo, 1e-8)))


@pytest.mark.parametrize("n", [1, 2, 3, 10, 1001])
@pytest.mark.parametrize("q", [0.0, 1.0, 50.0, 99.5, 100.0])
def test_percentile_matches_numpy(n, q):
    rng = np.random.default_rng(n)
    x = np.round(rng.standard_normal(n) * 3, 1)  # ties and negatives
    assert H.percentile_linear(x, q) == np.percentile(x, q)


def test_downsample_keeps_constants_and_linear_ramps():
    c = np.full((37, 50), 2.7)
    d = H.downsample(c)
    assert d.shape == (19, 25)
    assert np.allclose(d, 2.7, rtol=1e-15, atol=0)
    cols = np.arange(50, dtype=float)
    ramp = np.tile(0.3 * cols + 1.0, (37, 1))
    d = H.downsample(ramp)
    # the binomial blur reproduces a linear function away from the edge padding
    assert np.allclose(d[:, 1:-1], (0.3 * cols[::2] + 1.0)[None, 1:-1], rtol=0, atol=1e-12)


def test_upsample_reproduces_a_plane():
    hc, wc, h, w = 9, 13, 33, 50
    r, c = np.mgrid[0:hc, 0:wc].astype(float)
    coarse = 0.5 + 0.25 * r - 0.75 * c
    up = H.upsample(coarse, (h, w))
    R = np.linspace(0, hc - 1, h)[:, None]
    C = np.linspace(0, wc - 1, w)[None, :]
    assert np.allclose(up, 0.5 + 0.25 * R - 0.75 * C, rtol=0, atol=1e-12)


def test_central_gradient_of_a_ramp():
    img = np.tile(np.arange(20, dtype=float) * 1.5, (7, 1))
    gx, gy = H.central_gradient(img, spacing=2.0)
    assert np.allclose(gx[:, 1:-1], 1.5 / 2.0)
    assert np.allclose(gx[:, [0, -1]], 1.5 / 4.0)  # edge replication halves the ends
    assert np.array_equal(gy, np.zeros_like(gy))


def test_flat_image_factor_is_a_closed_form():
    # test_applications.py:196-199 and the formula: every level of a flat image has
    # g = eps and scale = alpha eps, so its factor is alpha^(1 - beta)
    cfg = H.ToneMapConfig()
    f = H.attenuation_field(np.zeros((64, 64)), cfg)
    levels = len(H.pyramid_shapes((64, 64), cfg.levels))
    assert levels == 3
    assert np.isfinite(f).all() and f.min() > 0.0
    assert np.allclose(f, cfg.alpha ** ((1.0 - cfg.beta) * levels), rtol=1e-12, atol=0)


def test_unit_beta_leaves_gradients_alone():
    L = np.log(G.radiance_map((40, 56), seed=3) @ np.array(H.LUMA))
    f = H.attenuation_field(L, H.ToneMapConfig(beta=1.0))
    assert np.allclose(f, 1.0, rtol=0, atol=1e-12)


def test_single_level_factor_is_the_power_law():
    L = np.log(G.radiance_map((48, 40), seed=4) @ np.array(H.LUMA))
    cfg = H.ToneMapConfig(levels=1)
    f = H.attenuation_field(L, cfg)
    gx, gy = np.gradient(np.pad(L, 1, mode="edge"))[1][1:-1, 1:-1], \
        np.gradient(np.pad(L, 1, mode="edge"))[0][1:-1, 1:-1]
    g = np.hypot(gx, gy)
    a = cfg.alpha * max(g.mean(), cfg.epsilon)
    want = np.exp((cfg.beta - 1.0) * np.log(np.maximum(g, cfg.epsilon) / a))
    assert np.allclose(f, want, rtol=1e-12, atol=0)


def test_unit_beta_and_saturation_tone_map_is_a_rescale():
    # factor 1: the solve returns log L - mean, so exp(c - max c) = L / max L and the
    # image is radiance / max L, normalised by its 99.5th percentile
    rad = G.radiance_map((36, 44), seed=5)
    img, f = H.tone_map(rad, H.ToneMapConfig(beta=1.0, saturation=1.0))
    lum = rad @ np.array(H.LUMA)
    base = rad / lum.max()
    want = np.clip(base / np.percentile(base, 99.5), 0.0, 1.0)
    assert np.allclose(img, want, rtol=0, atol=1e-9)


def test_tone_map_compresses_the_dynamic_range():
    # test_applications.py:178-186
    rad = G.radiance_map((128, 176), seed=2)
    mapped, factor = H.tone_map(rad)
    assert mapped.shape == rad.shape and factor.shape == rad.shape[:2]
    assert mapped.min() >= 0.0 and mapped.max() <= 1.0
    assert _decades(mapped) < 0.4 * _decades(rad)


def test_smaller_beta_compresses_more():
    # test_applications.py:189-193
    rad = G.radiance_map((96, 128), seed=2)
    gentle, _ = H.tone_map(rad, H.ToneMapConfig(beta=0.95))
    strong, _ = H.tone_map(rad, H.ToneMapConfig(beta=0.4))
    assert _decades(strong) < _decades(gentle)


def test_tone_map_rejects_a_grayscale_input():
    # test_applications.py:202-204
    with pytest.raises(ValueError):
        H.tone_map(np.ones((16, 16)))


def test_tone_map_solve_is_zero_mean_dct():
    rad = G.radiance_map((24, 30), seed=6)
    seen = []

    def spy(b):
        u = solve_neumann(b)
        seen.append(u)
        return u

    H.tone_map(rad, solve=spy)
    assert abs(seen[0].mean()) < 1e-12
