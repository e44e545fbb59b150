"""Gradient-domain tone mapping (FP64 numpy restatement; test infrastructure only).

Follows ``gradient-domain/src/gradient_domain/hdr.py``:

* ``ToneMapConfig``      hdr.py:37-70   alpha, beta, saturation, levels, epsilon
* ``downsample``         hdr.py:73-80   5-tap binomial blur (edge padding), rows of
                                        taps summed left to right, then every other sample
* ``upsample``           hdr.py:83-96   bilinear at numpy.linspace coordinates
* ``central_gradient``   hdr.py:99-104  edge-replicated central differences
* ``attenuation_field``  hdr.py:107-134 Gaussian-pyramid factor, coarse to fine
* ``tone_map``           hdr.py:137-185 log luminance, attenuated forward-difference
                                        field, zero-mean Neumann solve, exp, colour
                                        ratios, 99.5th-percentile normalisation
* ``percentile_linear``  numpy's default ("linear") percentile, used by hdr.py:184

Sums keep the reference's order (Python ``sum`` starts from 0 and adds the weighted taps
left to right) so a device restatement can match them bit for bit.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .operators import divergence, gradient
from .solvers import solve_neumann

LUMA = (0.2126, 0.7152, 0.0722)
BINOMIAL = (1.0 / 16.0, 4.0 / 16.0, 6.0 / 16.0, 4.0 / 16.0, 1.0 / 16.0)


@dataclass
class ToneMapConfig:
    alpha: float = 0.1
    beta: float = 0.85
    saturation: float = 0.5
    levels: int = 8
    epsilon: float = 1e-4


def log_luminance(radiance: np.ndarray) -> np.ndarray:
    """log(max(radiance . LUMA, tiny)) (hdr.py:165-167)."""
    lum = radiance @ np.array(LUMA)
    return np.log(np.maximum(lum, np.finfo(float).tiny))


def downsample(image: np.ndarray) -> np.ndarray:
    k = np.array([1.0, 4.0, 6.0, 4.0, 1.0]) / 16.0
    h, w = image.shape
    p = np.pad(image, ((0, 0), (2, 2)), mode="edge")
    blurred = sum(k[i] * p[:, i:i + w] for i in range(5))
    p = np.pad(blurred, ((2, 2), (0, 0)), mode="edge")
    blurred = sum(k[i] * p[i:i + h, :] for i in range(5))
    return blurred[::2, ::2]


def upsample(image: np.ndarray, shape) -> np.ndarray:
    rows = np.linspace(0, image.shape[0] - 1, shape[0])
    cols = np.linspace(0, image.shape[1] - 1, shape[1])
    r0 = np.floor(rows).astype(int)
    c0 = np.floor(cols).astype(int)
    r1 = np.minimum(r0 + 1, image.shape[0] - 1)
    c1 = np.minimum(c0 + 1, image.shape[1] - 1)
    rw = (rows - r0)[:, None]
    cw = (cols - c0)[None, :]
    top = image[r0][:, c0] * (1 - cw) + image[r0][:, c1] * cw
    bottom = image[r1][:, c0] * (1 - cw) + image[r1][:, c1] * cw
    return top * (1 - rw) + bottom * rw


def central_gradient(image: np.ndarray, spacing: float):
    p = np.pad(image, 1, mode="edge")
    gx = (p[1:-1, 2:] - p[1:-1, :-2]) / (2.0 * spacing)
    gy = (p[2:, 1:-1] - p[:-2, 1:-1]) / (2.0 * spacing)
    return gx, gy


def pyramid_shapes(shape, levels: int):
    """Shapes of the Gaussian pyramid of attenuation_field (hdr.py:121-123)."""
    shapes = [tuple(shape)]
    while len(shapes) < levels and min(shapes[-1]) > 16:
        h, w = shapes[-1]
        shapes.append(((h + 1) // 2, (w + 1) // 2))
    return shapes


def level_factor(level: np.ndarray, spacing: float, config: ToneMapConfig) -> np.ndarray:
    """(scale / g) (g / scale)^beta with g = max(|central grad|, eps), scale =
    alpha max(mean |grad|, eps) (hdr.py:127-132)."""
    gx, gy = central_gradient(level, spacing)
    mag = np.sqrt(gx * gx + gy * gy)
    scale = config.alpha * max(float(mag.mean()), config.epsilon)
    g = np.maximum(mag, config.epsilon)
    return (scale / g) * (g / scale) ** config.beta


def attenuation_field(log_lum: np.ndarray, config: ToneMapConfig) -> np.ndarray:
    pyramid = [np.asarray(log_lum, dtype=float)]
    while len(pyramid) < config.levels and min(pyramid[-1].shape) > 16:
        pyramid.append(downsample(pyramid[-1]))
    factor = None
    for lv in reversed(range(len(pyramid))):
        local = level_factor(pyramid[lv], 2.0 ** lv, config)
        factor = local if factor is None else upsample(factor, local.shape) * local
    return factor


def percentile_linear(values: np.ndarray, q: float) -> float:
    """numpy.percentile(values, q) with the default linear method, from the two order
    statistics around the virtual index (n - 1) q' (q' = q / 100), interpolated as
    numpy's _lerp does (b - (b - a)(1 - t) when t >= 0.5)."""
    x = np.sort(np.asarray(values, dtype=float).ravel())
    n = x.size
    qq = np.true_divide(q, np.float64(100))
    v = (n - 1) * qq
    lo = int(np.clip(np.floor(v), 0, n - 1))
    hi = int(np.clip(lo + 1, 0, n - 1))
    t = v - np.floor(v)
    a, b = x[lo], x[hi]
    d = b - a
    return float(b - d * (1 - t)) if t >= 0.5 else float(a + d * t)


def tone_map(radiance: np.ndarray, config: ToneMapConfig | None = None,
             solve=solve_neumann):
    """(image, factor); ``solve`` integrates the attenuated field (exact DCT by default)."""
    config = config or ToneMapConfig()
    radiance = np.asarray(radiance, dtype=float)
    if radiance.ndim != 3 or radiance.shape[2] != 3:
        raise ValueError("tone mapping expects a three-channel radiance map")
    lum = np.maximum(radiance @ np.array(LUMA), np.finfo(float).tiny)
    log_l = np.log(lum)
    factor = attenuation_field(log_l, config)
    gx, gy = gradient(log_l)
    compressed = solve(divergence(gx * factor, gy * factor))
    result = np.exp(compressed - compressed.max())
    image = ((radiance / lum[..., None]) ** config.saturation) * result[..., None]
    peak = percentile_linear(image, 99.5)
    return np.clip(image / max(peak, 1e-12), 0.0, 1.0), factor
