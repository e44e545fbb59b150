"""Seeded inputs for the solve workloads (FP64 numpy).

* ``manufactured_problem`` restates ``gradient-domain/src/gradient_domain/benchmark.py:65-95``
  (smooth modes + 0.25 N(0,1), zero ring, b = laplacian(solution)); with seed 0 it is
  the problem behind ``gradient-domain/docs/benchmark.json``.
* ``blob_image`` is the smooth+noise target: a sum of 8 Gaussians (amplitude U[0,1],
  sigma U[10,60] px, centre uniform in the image) plus N(0, 0.01) noise, read as
  standard deviation 0.01.  Its Neumann RHS is ``divergence(*gradient(target))``, whose
  exact zero-mean solution is ``target - mean(target)`` (reference
  ``tests/test_operators.py:51-57`` pins that composition).
* ``blob_mask`` thresholds Gaussian-smoothed (sigma 40) white noise at its 70th
  percentile.
* ``compositing_problem`` builds the rectangle-domain compositing RHS of
  ``poisson.solve_dirichlet`` (``poisson.py:130-145``): guidance = source gradients
  inside the mask, target gradients outside; ``fold_boundary`` with the target ring.
* ``batch_shapes`` draws side lengths log-uniform in [128, 2048] and aspect U[0.5, 2]:
  ``H = round(s * sqrt(a))``, ``W = round(s / sqrt(a))``.
* ``hdr_log_luminance`` / ``attenuation_single_scale`` give the tone-map RHS with the
  single-scale factor ``(|grad|/alpha)^(beta-1)``, ``alpha = 0.1 mean|grad|``,
  ``beta = 0.85`` (``hdr.py:107-134`` with ``levels=1``; magnitudes from central
  differences ``hdr.py:99-104``, integrated field from forward differences ``hdr.py:174``).

Seeds: ``1000 * config + 10 * problem + channel`` unless a test says otherwise.
"""

from __future__ import annotations

import numpy as np

from .operators import divergence, fold_boundary, gradient, laplacian


def manufactured_problem(shape, seed=0):
    rng = np.random.default_rng(seed)
    rows, cols = np.mgrid[0 : shape[0], 0 : shape[1]] / max(shape)
    sol = (
        np.sin(3.0 * np.pi * cols) * np.sin(2.0 * np.pi * rows)
        + 0.4 * np.cos(11.0 * np.pi * cols) * np.sin(7.0 * np.pi * rows)
        + 0.25 * rng.standard_normal(shape)
    )
    sol[0] = sol[-1] = 0.0
    sol[:, 0] = sol[:, -1] = 0.0
    return laplacian(sol), sol


def blob_image(shape, seed, blobs=8, noise=0.01):
    rng = np.random.default_rng(seed)
    h, w = shape
    amp = rng.uniform(0.0, 1.0, blobs)
    sig = rng.uniform(10.0, 60.0, blobs)
    cy = rng.uniform(0.0, h, blobs)
    cx = rng.uniform(0.0, w, blobs)
    y = np.arange(h, dtype=np.float64)[:, None]
    x = np.arange(w, dtype=np.float64)[None, :]
    img = np.zeros(shape)
    for k in range(blobs):
        gy = np.exp(-((y - cy[k]) ** 2) / (2.0 * sig[k] ** 2))
        gx = np.exp(-((x - cx[k]) ** 2) / (2.0 * sig[k] ** 2))
        img += amp[k] * gy * gx
    img += noise * rng.standard_normal(shape)
    return img


def neumann_problem(shape, seed):
    """(b, exact zero-mean solution) for the blob target."""
    t = blob_image(shape, seed)
    b = divergence(*gradient(t))
    return b, t - t.mean()


def rgb_neumann_problem(shape, config, problem=0):
    bs, xs = [], []
    for c in range(3):
        b, x = neumann_problem(shape, 1000 * config + 10 * problem + c)
        bs.append(b)
        xs.append(x)
    return np.stack(bs), np.stack(xs)


def _gaussian_smooth(a, sigma):
    from scipy.ndimage import gaussian_filter

    return gaussian_filter(a, sigma, mode="reflect")


def blob_mask(shape, seed, sigma=40.0, percentile=70.0):
    rng = np.random.default_rng(seed)
    noise = _gaussian_smooth(rng.standard_normal(shape), sigma)
    return noise > np.percentile(noise, percentile)


def compositing_problem(shape, config=3, problem=0, channels=3):
    """Rectangle-domain seamless clone RHS per channel.

    Returns (folded RHS (c, H-2, W-2), target (c, H, W), mask (H, W)).
    """
    mask = blob_mask(shape, 1000 * config + 10 * problem + 9)
    rhs, tgt = [], []
    for c in range(channels):
        target = blob_image(shape, 1000 * config + 10 * problem + c)
        source = blob_image(shape, 1000 * config + 10 * problem + 5 + c)
        sgx, sgy = gradient(source)
        tgx, tgy = gradient(target)
        gx = np.where(mask, sgx, tgx)
        gy = np.where(mask, sgy, tgy)
        div = divergence(gx, gy)
        rhs.append(fold_boundary(div[1:-1, 1:-1], target))
        tgt.append(target)
    return np.stack(rhs), np.stack(tgt), mask


def compositing_images(shape, config=3, problem=0, channels=3):
    """(source, target) channels-last images and the blob mask of compositing_problem."""
    mask = blob_mask(shape, 1000 * config + 10 * problem + 9)
    target = np.stack([blob_image(shape, 1000 * config + 10 * problem + c)
                       for c in range(channels)], axis=2)
    source = np.stack([blob_image(shape, 1000 * config + 10 * problem + 5 + c)
                       for c in range(channels)], axis=2)
    return source, target, mask


def batch_shapes(count=64, seed=4000, low=128, high=2048):
    rng = np.random.default_rng(seed)
    s = np.exp(rng.uniform(np.log(low), np.log(high), count))
    a = rng.uniform(0.5, 2.0, count)
    return [(int(round(si * np.sqrt(ai))), int(round(si / np.sqrt(ai)))) for si, ai in zip(s, a)]


def hdr_log_luminance(shape, seed, sigma=2.0):
    """Log of the luminance of an i.i.d. log-normal (mu 0, sigma 2) RGB radiance map."""
    rng = np.random.default_rng(seed)
    radiance = np.exp(rng.normal(0.0, sigma, shape + (3,)))
    lum = radiance @ np.array([0.2126, 0.7152, 0.0722])
    return np.log(np.maximum(lum, np.finfo(float).tiny))


def central_gradient_magnitude(img, spacing=1.0):
    p = np.pad(img, 1, mode="edge")
    gx = (p[1:-1, 2:] - p[1:-1, :-2]) / (2.0 * spacing)
    gy = (p[2:, 1:-1] - p[:-2, 1:-1]) / (2.0 * spacing)
    return np.sqrt(gx * gx + gy * gy)


def attenuation_single_scale(log_lum, alpha=0.1, beta=0.85, eps=1e-4):
    mag = central_gradient_magnitude(log_lum)
    scale = alpha * max(float(mag.mean()), eps)
    g = np.maximum(mag, eps)
    return (scale / g) * (g / scale) ** beta


def tonemap_problem(shape, seed):
    """Neumann RHS div(factor * grad(log L)) of the tone-map solve."""
    log_l = hdr_log_luminance(shape, seed)
    f = attenuation_single_scale(log_l)
    gx, gy = gradient(log_l)
    return divergence(gx * f, gy * f)


def radiance_map(shape, seed, decades=5.0, noise=0.01):
    """Smooth HDR radiance (h, w, 3): the blob image rescaled so its luminance spans
    ``decades`` decades, times a smooth per-channel tint in [0.6, 1.4], times
    (1 + noise N(0, 1))."""
    base = blob_image(shape, seed)
    base = (base - base.min()) / max(base.max() - base.min(), 1e-12)
    lum = 10.0 ** (decades * base - 0.5 * decades)
    tints = []
    for c in range(3):
        t = blob_image(shape, seed + 11 + c)
        tints.append(0.6 + 0.8 * (t - t.min()) / max(t.max() - t.min(), 1e-12))
    rng = np.random.default_rng(seed + 17)
    rad = lum[..., None] * np.stack(tints, axis=2)
    return rad * (1.0 + noise * rng.standard_normal(rad.shape)).clip(0.5, None)


def lognormal_radiance(shape, seed, sigma=2.0):
    """i.i.d. log-normal (mu 0, sigma 2) radiance per pixel and channel, (h, w, 3): the
    tone-mapping requests of the serving stream (the log field is not smoothed)."""
    rng = np.random.default_rng(seed)
    return np.exp(rng.normal(0.0, sigma, tuple(shape) + (3,)))
