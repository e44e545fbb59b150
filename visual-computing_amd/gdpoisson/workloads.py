"""Seeded synthetic inputs of the benchmark workloads (numpy generation, device upload).

* ``blob_image``: a sum of 8 Gaussians (amplitude U[0,1], sigma U[10,60] px, centre
  uniform in the image) plus N(0, 0.01) noise (standard deviation 0.01).
* ``neumann_rhs``: ``divergence(gradient(target))`` on the device; its exact zero-mean
  solution is ``target - mean(target)``.
* ``blob_mask``: Gaussian-smoothed (sigma 40, reflected edges) white noise above its 70th
  percentile (about 30% of the pixels).
* ``compositing_images``: source and target blob images (channels-last) and a blob mask.
* ``batch_shapes``: sides log-uniform in [128, 2048], aspect U[0.5, 2],
  H = round(s sqrt(a)), W = round(s / sqrt(a)).
* ``lognormal_radiance``: i.i.d. log-normal (mu 0, sigma 2) radiance per pixel and channel.
* ``serving_stream``: Poisson arrivals at a fixed rate, 70% compositing requests (square
  sides uniform in {512 .. 2048}, crops of blob images and of a blob mask) and 30%
  3840 x 2160 tone-mapping requests.

Seeds: 1000 * config + 10 * problem + channel unless stated.
"""

from __future__ import annotations

import numpy as np
import torch

from .serving import Request, poisson_arrivals


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
        img += amp[k] * np.exp(-((y - cy[k]) ** 2) / (2.0 * sig[k] ** 2)) * \
            np.exp(-((x - cx[k]) ** 2) / (2.0 * sig[k] ** 2))
    img += noise * rng.standard_normal(shape)
    return img


def neumann_rhs(target) -> torch.Tensor:
    """divergence(gradient(target)) of a device or host plane, on the device."""
    from .operators import divergence, gradient

    t = target if isinstance(target, torch.Tensor) else torch.from_numpy(np.asarray(target))
    t = t.to(dtype=torch.float64).cuda().contiguous()
    return divergence(*gradient(t))


def rgb_neumann_problem(shape, config=2, problem=0):
    """([b per channel] on the device, [exact zero-mean solution per channel] on the host)."""
    bs, exact = [], []
    for c in range(3):
        t = blob_image(shape, 1000 * config + 10 * problem + c)
        bs.append(neumann_rhs(t))
        exact.append(t - t.mean())
    return bs, exact


def blob_mask(shape, seed, sigma=40.0, percentile=70.0):
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    noise = gaussian_filter(rng.standard_normal(shape), sigma, mode="reflect")
    return noise > np.percentile(noise, percentile)


def compositing_images(shape, config=3, problem=0, channels=3):
    """(source, target) channels-last and the mask, as in the 2048 x 2048 RGB workload."""
    base = 1000 * config + 10 * problem
    mask = blob_mask(shape, base + 9)
    target = np.stack([blob_image(shape, base + c) for c in range(channels)], axis=2)
    source = np.stack([blob_image(shape, base + 5 + c) for c in range(channels)], axis=2)
    return source, target, mask


def batch_shapes(count=64, seed=4000, low=128, high=2048):
    rng = np.random.default_rng(seed)
    s = np.exp(rng.uniform(np.log(low), np.log(high), count))
    a = rng.uniform(0.5, 2.0, count)
    return [(int(round(si * np.sqrt(ai))), int(round(si / np.sqrt(ai)))) for si, ai in zip(s, a)]


def lognormal_radiance(shape, seed, sigma=2.0):
    rng = np.random.default_rng(seed)
    return np.exp(rng.normal(0.0, sigma, tuple(shape) + (3,)))


def serving_stream(count: int, qps: float, seed: int = 5000, sides=(512, 2048),
                   side_step: int = 256, tonemap_shape=(2160, 3840), tonemap_fraction=0.3,
                   base_side: int = 2048, radiance_maps: int = 6, tonemap_levels: int = 1,
                   device=None) -> list[Request]:
    """Requests of the mixed serving workload with device-resident inputs.

    Compositing request: a square whose side is drawn uniformly from the multiples of
    ``side_step`` in ``sides`` (the standard sizes an image stream repeats), cut at a
    random place from one pair of base_side x base_side RGB blob images and from a blob
    mask (source and target the same size, offset 0).  Tone-mapping request: one of
    ``radiance_maps`` log-normal radiance maps of ``tonemap_shape`` (mu 0, sigma 2, i.i.d.
    per pixel), attenuated by the single-scale factor (|grad| / alpha)^(beta - 1)
    (``tonemap_levels`` = 1: attenuation_field with one level, hdr.py:122, 127-132).
    Arrivals: a Poisson process of rate ``qps``.
    """
    from .hdr import ToneMapConfig

    dev = torch.device("cuda", torch.cuda.current_device() if device is None else device)
    rng = np.random.default_rng(seed)
    src, tgt, mask = compositing_images((base_side, base_side), config=5, problem=0)
    src_d = torch.from_numpy(src).to(dev)
    tgt_d = torch.from_numpy(tgt).to(dev)
    mask_d = torch.from_numpy(mask).to(dev)
    rads = [torch.from_numpy(lognormal_radiance(tonemap_shape, seed + 100 + k)).to(dev)
            for k in range(radiance_maps)]
    config = ToneMapConfig(levels=int(tonemap_levels))
    side_set = list(range(sides[0], min(sides[1], base_side) + 1, max(1, int(side_step))))
    arrivals = poisson_arrivals(count, qps, seed + 1)
    out = []
    for i, t in enumerate(arrivals):
        if rng.random() < tonemap_fraction:
            out.append(Request("tonemap", {"radiance": rads[int(rng.integers(radiance_maps))],
                                           "config": config}, float(t), i))
            continue
        s = int(side_set[int(rng.integers(len(side_set)))])
        r0 = int(rng.integers(0, base_side - s + 1))
        c0 = int(rng.integers(0, base_side - s + 1))
        win = (slice(r0, r0 + s), slice(c0, c0 + s))
        out.append(Request("composite", {
            "source": src_d[win], "target": tgt_d[win], "mask": mask_d[win],
            "offset": (0, 0), "mode": "import"}, float(t), i))
    return out
