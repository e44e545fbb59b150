"""texture_flatten and illumination_change on the device (poisson.py:351-437) against
the oracle, with the reference's own property checks (test_applications.py:150-175)."""

import numpy as np
import pytest

from oracle import poisson as OP

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gd():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gdpoisson

    return gdpoisson


def _texture_example(shape=(96, 128), seed=1):
    rng = np.random.default_rng(seed)
    rows, cols = np.mgrid[0:shape[0], 0:shape[1]]
    regions = (rows > shape[0] // 2).astype(int) + 2 * (cols > shape[1] // 3).astype(int)
    image = rng.random((4, 3))[regions] + 0.05 * np.sin(1.7 * cols + 0.9 * rows)[..., None]
    edges = np.zeros(shape, dtype=bool)
    edges[:, :-1] |= regions[:, 1:] != regions[:, :-1]
    edges[:-1, :] |= regions[1:, :] != regions[:-1, :]
    return image, edges.astype(float)


@pytest.mark.parametrize("method", ["mgcg", "mgcg_mixed"])
@pytest.mark.parametrize("gray", [False, True])
def test_texture_flatten_matches_the_oracle(gd, method, gray):
    image, edges = _texture_example()
    if gray:
        image = image[..., 0].copy()
    got, reps = gd.texture_flatten(image, edges, method=method)
    want, _ = OP.texture_flatten(image, edges, method="direct")
    assert all(r.converged for r in reps)
    assert np.abs(got - want).max() < 1e-6


@pytest.mark.parametrize("method", ["mgcg", "mgcg_mixed"])
def test_masked_texture_flatten_matches_the_oracle(gd, method):
    image, edges = _texture_example((80, 100), seed=2)
    rows, cols = np.mgrid[0:80, 0:100]
    mask = (rows - 40) ** 2 + (cols - 50) ** 2 < 30.0 ** 2
    got, reps = gd.texture_flatten(image, edges, method=method, mask=mask)
    want, _ = OP.texture_flatten(image, edges, mask=mask)
    assert len(reps) == 3 and all(r.converged for r in reps)
    assert np.abs(got - want).max() < 1e-6
    assert np.array_equal(got[~mask], image[~mask])


def test_texture_flatten_with_unit_edges_returns_the_image(gd):
    """test_applications.py:150-152."""
    image, _ = _texture_example((64, 80))
    got, _ = gd.texture_flatten(image, np.ones(image.shape[:2]), tolerance=1e-11)
    assert np.abs(got - image).max() < 1e-9


@pytest.mark.parametrize("beta", [0.0, 0.2, 0.6])
@pytest.mark.parametrize("method", ["mgcg", "mgcg_mixed"])
def test_illumination_change_matches_the_oracle(gd, beta, method):
    rng = np.random.default_rng(8)
    image = rng.random((70, 90, 3)) * 0.3
    rows, cols = np.mgrid[0:70, 0:90]
    mask = (rows - 35) ** 2 + (cols - 45) ** 2 < 25.0 ** 2
    got, reps = gd.illumination_change(image, mask, alpha=0.4, beta=beta, method=method)
    want, _ = OP.illumination_change(image, mask, alpha=0.4, beta=beta)
    assert len(reps) == 3 and all(r.converged for r in reps)
    assert np.abs(got - want).max() < 1e-6
    assert np.array_equal(got[~mask], image[~mask])


def test_illumination_field_matches_numpy(gd):
    """The fused attenuated field against numpy's expression (poisson.py:433-437): the
    magnitude is exact; pow is not correctly rounded on either side."""
    import ctypes

    import torch

    from gdpoisson import _lib

    rng = np.random.default_rng(9)
    gx = rng.standard_normal((33, 47, 3)) * 0.1
    gy = rng.standard_normal((33, 47, 3)) * 0.1
    gx[0, :3] = gy[0, :3] = 0.0  # magnitudes under epsilon
    tx, ty = torch.from_numpy(gx).cuda(), torch.from_numpy(gy).cuda()
    ox, oy = torch.empty_like(tx), torch.empty_like(ty)
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    assert _lib.lib().gdp_illumination_field(p(tx), p(ty), tx.numel(), 0.2, 0.3, 1e-4, p(ox),
                                             p(oy), None) == 0
    f = (0.2 ** 0.3) * np.maximum(np.sqrt(gx ** 2 + gy ** 2), 1e-4) ** (-0.3)
    assert np.allclose(ox.cpu().numpy(), gx * f, rtol=1e-14, atol=0)
    assert np.allclose(oy.cpu().numpy(), gy * f, rtol=1e-14, atol=0)


def test_illumination_change_lifts_a_dark_selection(gd):
    """test_applications.py:167-175."""
    rows, cols = np.mgrid[0:96, 0:96]
    image = np.repeat((0.05 + 0.02 * np.sin(cols * 0.4))[..., None], 3, axis=2)
    mask = (rows - 48) ** 2 + (cols - 48) ** 2 < 30.0 ** 2
    inner = (rows - 48) ** 2 + (cols - 48) ** 2 < 24.0 ** 2
    out, _ = gd.illumination_change(image, mask, alpha=0.4, beta=0.6)
    ring = mask & ~inner
    assert out[ring].std() > image[ring].std()
    assert np.array_equal(out[~mask], image[~mask])
