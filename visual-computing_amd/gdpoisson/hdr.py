"""Gradient-domain tone mapping on the GPU, with the interface of
``gradient-domain/src/gradient_domain/hdr.py``.

``attenuation_field`` (hdr.py:107-134) builds the Gaussian pyramid of the log luminance
(gdp_tm_downsample) and the per-level factor from coarse to fine (gdp_tm_level_factor);
``tone_map`` (hdr.py:137-185) integrates the attenuated forward-difference field with a
Neumann solve on the device -- by default the reference's exact DCT diagonalisation
(``solve_neumann``), or multigrid-preconditioned CG (``method="mgcg"`` / ``"mgcg_mixed"``,
to the stated relative residual) -- then composes the display image and normalises it by
its 99.5th percentile, all without leaving the GPU.
numpy in gives numpy out; device tensors in give device tensors out.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from . import _lib
from ._device import to_device, to_host_like
from .operators import _stream
from .solvers import (SolverReport, dct_supported, relative_residual, solve_neumann,
                      solve_planes)

__all__ = ["ToneMapConfig", "attenuation_field", "tone_map", "prepare_tone_map",
           "finish_tone_map", "ToneJob", "percentile"]


@dataclass
class ToneMapConfig:
    """Attenuation parameters (hdr.py:37-70): gradients of ``alpha`` times the level's
    mean magnitude are left alone, the factor is ``(g / alpha) ** (beta - 1)``,
    ``saturation`` is the exponent of the colour ratios, ``levels`` the pyramid depth
    (stopping before a level would drop to 16 pixels or fewer), ``epsilon`` the floor of
    the magnitudes."""

    alpha: float = 0.1
    beta: float = 0.85
    saturation: float = 0.5
    levels: int = 8
    epsilon: float = 1e-4


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _work_tensor(h: int, w: int, device) -> torch.Tensor:
    nbytes = ctypes.c_size_t()
    _lib.check(_lib.lib().gdp_tm_work_bytes(h, w, ctypes.byref(nbytes)), "gdp_tm_work_bytes")
    return torch.empty((nbytes.value + 7) // 8, dtype=torch.float64, device=device)


def pyramid_shapes(shape, levels: int):
    """Level shapes of attenuation_field's pyramid (hdr.py:121-123)."""
    shapes = [tuple(int(s) for s in shape)]
    while len(shapes) < levels and min(shapes[-1]) > 16:
        h, w = shapes[-1]
        shapes.append(((h + 1) // 2, (w + 1) // 2))
    return shapes


def _attenuation(L: torch.Tensor, config: ToneMapConfig, work: torch.Tensor) -> torch.Tensor:
    lib, s = _lib.lib(), _stream()
    shapes = pyramid_shapes(L.shape, int(config.levels))
    pyramid = [L]
    for (h, w), (ho, wo) in zip(shapes[:-1], shapes[1:]):
        nxt = torch.empty((ho, wo), dtype=torch.float64, device=L.device)
        _lib.check(lib.gdp_tm_downsample(_p(pyramid[-1]), h, w, _p(nxt), s), "gdp_tm_downsample")
        pyramid.append(nxt)
    factor = None
    for lv in reversed(range(len(pyramid))):
        h, w = shapes[lv]
        f = torch.empty((h, w), dtype=torch.float64, device=L.device)
        hc, wc = (factor.shape if factor is not None else (0, 0))
        _lib.check(lib.gdp_tm_level_factor(
            _p(pyramid[lv]), h, w, 2.0 ** lv, float(config.alpha), float(config.beta),
            float(config.epsilon), _p(factor), hc, wc, _p(f), _p(work), s),
            "gdp_tm_level_factor")
        factor = f
    return factor


def attenuation_field(log_luminance, config: ToneMapConfig):
    """Per-pixel multi-scale attenuation factor of a log-luminance image (hdr.py:107-134)."""
    L, host = to_device(log_luminance)
    if L.ndim != 2:
        raise ValueError("attenuation_field expects one (h, w) plane")
    work = _work_tensor(*L.shape, L.device)
    return to_host_like(_attenuation(L, config, work), host)


@dataclass
class ToneJob:
    """A prepared tone_map: the radiance, its attenuation factor and the right-hand side
    (1 x h x w) of the Neumann solve."""

    radiance: torch.Tensor
    host: bool
    config: ToneMapConfig
    factor: torch.Tensor
    rhs: torch.Tensor
    work: torch.Tensor


def prepare_tone_map(radiance, config: ToneMapConfig | None = None) -> ToneJob:
    """Log luminance, attenuation factor and attenuated divergence (hdr.py:160-175), on
    the current stream."""
    config = config or ToneMapConfig()
    rad, host = to_device(radiance)
    if rad.ndim != 3 or rad.shape[2] != 3:
        raise ValueError("tone mapping expects a three-channel radiance map")
    h, w = int(rad.shape[0]), int(rad.shape[1])
    if h < 1 or w < 1:
        raise ValueError("tone mapping needs a non-empty radiance map")
    lib, s, dev = _lib.lib(), _stream(), rad.device
    work = _work_tensor(h, w, dev)
    L = torch.empty((h, w), dtype=torch.float64, device=dev)
    _lib.check(lib.gdp_tm_log_luminance(_p(rad), h, w, _p(L), s), "gdp_tm_log_luminance")
    factor = _attenuation(L, config, work)
    rhs = torch.empty((1, h, w), dtype=torch.float64, device=dev)
    _lib.check(lib.gdp_tm_rhs(_p(L), _p(factor), h, w, _p(rhs), s), "gdp_tm_rhs")
    return ToneJob(rad, host, config, factor, rhs, work)


def finish_tone_map(job: ToneJob, compressed: torch.Tensor):
    """Display image from the integrated log luminance (hdr.py:180-185)."""
    lib, s, dev = _lib.lib(), _stream(), job.radiance.device
    h, w = int(job.radiance.shape[0]), int(job.radiance.shape[1])
    image = torch.empty((h, w, 3), dtype=torch.float64, device=dev)
    _lib.check(lib.gdp_tm_compose(_p(job.radiance), _p(compressed), h, w,
                                  float(job.config.saturation), _p(image), _p(job.work), s),
               "gdp_tm_compose")
    nsel = ctypes.c_size_t()
    _lib.check(lib.gdp_percentile_work_bytes(ctypes.byref(nsel)), "gdp_percentile_work_bytes")
    sel = torch.empty((nsel.value + 7) // 8, dtype=torch.float64, device=dev)
    peak = torch.empty(1, dtype=torch.float64, device=dev)
    _lib.check(lib.gdp_percentile(_p(image), image.numel(), 99.5, _p(peak), _p(sel), s),
               "gdp_percentile")
    _lib.check(lib.gdp_tm_normalise(_p(image), image.numel(), _p(peak), s), "gdp_tm_normalise")
    return to_host_like(image, job.host), to_host_like(job.factor, job.host)


def tone_map(radiance, config: ToneMapConfig | None = None, method: str = "dct",
             tolerance: float = 1e-8, max_iterations: int = 2000, reports: list | None = None):
    """Compress a radiance map to a displayable image (hdr.py:137-185).

    Returns ``(image in [0, 1] of shape (h, w, 3), factor (h, w))``.  The Neumann solve
    is the exact DCT (``method="dct"``, as the reference's ``solve_neumann``; sides whose
    prime factors are not all 2, 3, 5 use FP64 MG-PCG) or MG-PCG (``"mgcg"`` FP64,
    ``"mgcg_mixed"``) to ``tolerance``; its SolverReport (with the true relative residual)
    is appended to ``reports`` when a list is given.
    """
    import time

    job = prepare_tone_map(radiance, config)
    if method == "dct":
        t0 = time.perf_counter()
        exact = dct_supported(*job.rhs.shape[1:])
        u = solve_neumann(job.rhs[0], method="dct" if exact else "mgcg", tolerance=1e-12)
        if reports is not None:
            rr = relative_residual(u, job.rhs[0], 1.0, "neumann")
            reports.append(SolverReport(method="dct" if exact else "mgcg", iterations=1,
                                        relative_residual=rr,
                                        seconds=time.perf_counter() - t0, converged=True))
        return finish_tone_map(job, u)
    sol, reps = solve_planes(job.rhs, method=method, tolerance=tolerance,
                             max_iterations=max_iterations, boundary="neumann")
    if reports is not None:
        reports.extend(reps)
    return finish_tone_map(job, sol[0])


def percentile(x, q: float) -> float:
    """numpy.percentile(x, q) (linear method) of a device tensor or array, on the GPU."""
    t, _ = to_device(x)
    t = t.reshape(-1)
    nsel = ctypes.c_size_t()
    _lib.check(_lib.lib().gdp_percentile_work_bytes(ctypes.byref(nsel)), "gdp_percentile_work_bytes")
    sel = torch.empty((nsel.value + 7) // 8, dtype=torch.float64, device=t.device)
    out = torch.empty(1, dtype=torch.float64, device=t.device)
    _lib.check(_lib.lib().gdp_percentile(_p(t), t.numel(), float(q), _p(out), _p(sel), _stream()),
               "gdp_percentile")
    return float(out.item())
