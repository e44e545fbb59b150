"""Gradient-domain compositing on the GPU, with the interface of
``gradient-domain/src/gradient_domain/poisson.py``.

``GuidanceField`` (poisson.py:52-92) holds a target gradient field; ``solve_dirichlet``
(poisson.py:102-147) integrates a field over a rectangle whose one-pixel ring is fixed;
``seamless_clone`` (poisson.py:261-348) pastes a source region into a target.  Images
are ``(h, w)`` or channels-last ``(h, w, c)``; numpy in gives numpy out, device tensors
in give device tensors out.  Every channel of a call is solved by one batched plan.

``seamless_clone`` takes either route of the reference.  Rectangle (``domain=
"rectangle"``): the unknowns are a window around the selection, the guidance is the
chosen combination inside the mask and the target's gradient outside, and the window's
ring supplies Dirichlet values; one fused kernel builds the right-hand side
(gdp_clone_rhs).  Mask (``domain="mask"``, ``solve_masked``, poisson.py:154-244): the
unknowns are the selected pixels off the image border and every other pixel keeps its
value.  The reference assembles that sparse system and factorises it with SuperLU; here
it is the Dirichlet problem of the selection's bounding box with a mask (gdp_plan_set_mask):
masked MG-PCG whose operator and residual act on the selected pixels only, preconditioned
by the box's V-cycle masked on both sides, to the stated relative residual.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._device import to_device, to_host_like
from .operators import _planes, _stream, _unplanes, divergence, fold_boundary, gradient
from .solvers import SolverReport, solve_planes

__all__ = ["GuidanceField", "solve_dirichlet", "solve_masked", "seamless_clone",
           "prepare_clone", "finish_clone", "CloneJob"]

MODES = {"import": 0, "mixed": 1, "average": 2}


def _mask_tensor(mask, shape, device) -> torch.Tensor:
    t = mask if isinstance(mask, torch.Tensor) else torch.from_numpy(np.asarray(mask))
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"mask shape {tuple(t.shape)} does not match {tuple(shape)}")
    return (t != 0).to(device=device, dtype=torch.uint8).contiguous()


@dataclass
class GuidanceField:
    """A target gradient field (gx, gy), each ``(h, w)`` or ``(h, w, c)``."""

    gx: object
    gy: object

    @classmethod
    def of(cls, image) -> "GuidanceField":
        """The field that reproduces ``image`` exactly."""
        return cls(*gradient(image))

    def _combine(self, other, mask, factor, scale: float, mode: int) -> "GuidanceField":
        ax, host = to_device(self.gx)
        ay, _ = to_device(self.gy)
        pax, cl = _planes(ax)
        pay, _ = _planes(ay)
        c, h, w = pax.shape
        pbx = pby = None
        if other is not None:
            pbx, _ = _planes(to_device(other.gx)[0])
            pby, _ = _planes(to_device(other.gy)[0])
            if pbx.shape != pax.shape:
                raise ValueError("the two fields differ in shape")
        m = _mask_tensor(mask, (h, w), pax.device) if mask is not None else None
        f = None
        if factor is not None:
            f = to_device(factor)[0].contiguous()
            if tuple(f.shape) != (h, w):
                raise ValueError(f"factor shape {tuple(f.shape)} does not match {(h, w)}")
        ox, oy = torch.empty_like(pax), torch.empty_like(pay)
        ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
        _lib.check(_lib.lib().gdp_guidance(
            ptr(pax), ptr(pay), ptr(pbx), ptr(pby), ptr(m), ptr(f), float(scale), mode, h, w, c,
            ptr(ox), ptr(oy), _stream()), "gdp_guidance")
        return GuidanceField(to_host_like(_unplanes(ox, cl), host),
                             to_host_like(_unplanes(oy, cl), host))

    def select(self, mask, other: "GuidanceField") -> "GuidanceField":
        """This field inside ``mask`` and ``other`` outside it (poisson.py:71-76)."""
        return self._combine(other, mask, None, 1.0, 0)

    def mix_by_magnitude(self, other: "GuidanceField") -> "GuidanceField":
        """Per component, the field of larger magnitude (this one on ties; :78-88)."""
        return self._combine(other, None, None, 1.0, 1)

    def scaled(self, factor) -> "GuidanceField":
        """Both components times a scalar or a per-pixel factor (poisson.py:90-92)."""
        if isinstance(factor, (int, float)):
            return self._combine(None, None, None, float(factor), 3)
        return self._combine(None, None, factor, 1.0, 3)


def _solve_window(rhs_planes: torch.Tensor, method, tolerance, max_iterations,
                  mask: torch.Tensor | None = None):
    if mask is None:
        return solve_planes(rhs_planes, method=method, tolerance=tolerance,
                            max_iterations=max_iterations)
    return _solve_masked_planes(rhs_planes, mask, method, tolerance, max_iterations)


def _solve_masked_planes(rhs: torch.Tensor, mask: torch.Tensor, method, tolerance,
                         max_iterations):
    """Masked Dirichlet MG-PCG of every plane of rhs (c, m, n) over the unknowns of mask."""
    import dataclasses
    import time

    from .plan import Plan
    from .solvers import SolverReport, _pcg_options

    if method not in ("mgcg", "mgcg_mixed"):
        raise KeyError(f"the masked solve needs 'mgcg' or 'mgcg_mixed', not {method!r}")
    c, m, n = rhs.shape
    precision = "mixed" if method == "mgcg_mixed" else "fp64"
    opts = dataclasses.replace(_pcg_options(precision, tolerance, max_iterations, "dirichlet"),
                               use_graph=False)
    t0 = time.perf_counter()
    plan = Plan([(m, n)] * c, "dirichlet", 1.0, opts, device=rhs.device, masks=[mask] * c)
    us, reps = plan.mgcg([rhs[k] for k in range(c)])
    torch.cuda.current_stream().synchronize()
    dt = time.perf_counter() - t0
    return torch.stack(us), [SolverReport(method=f"{method} (masked)", iterations=int(r.iterations),
                                          relative_residual=float(r.relative_residual),
                                          seconds=dt, converged=bool(r.converged)) for r in reps]


def _interior(mask_hw: torch.Tensor) -> torch.Tensor:
    """The selected pixels off the image border (poisson.py:182-184), as uint8."""
    inner = mask_hw.clone()
    inner[0, :] = 0
    inner[-1, :] = 0
    inner[:, 0] = 0
    inner[:, -1] = 0
    return inner


def _box(inner: torch.Tensor):
    """Window [top, bottom) x [left, right) of the unknowns plus their fixed neighbours."""
    nz = torch.nonzero(inner)
    if nz.numel() == 0:
        return None
    (r0, c0), (r1, c1) = nz.min(dim=0).values.tolist(), nz.max(dim=0).values.tolist()
    return r0 - 1, r1 + 2, c0 - 1, c1 + 2


def solve_masked(field: "GuidanceField", image, mask, method: str = "mgcg_mixed",
                 tolerance: float = 1e-8, max_iterations: int = 2000):
    """Integrate ``field`` over the pixels ``mask`` selects; every other pixel keeps its
    value in ``image`` and acts as Dirichlet data (poisson.py:154-244).  Returns
    ``(result, reports)``, one report per channel."""
    img, host = to_device(image)
    pi, cl = _planes(img)
    c, h, w = pi.shape
    inner = _interior(_mask_tensor(mask, (h, w), pi.device))
    box = _box(inner)
    if box is None:
        return to_host_like(_unplanes(pi.clone(), cl), host), []
    top, bottom, left, right = box
    hh, ww = bottom - top, right - left
    gx, _ = to_device(field.gx)
    gy, _ = to_device(field.gy)
    pgx, _ = _planes(gx)
    pgy, _ = _planes(gy)
    if pgx.shape != pi.shape or pgy.shape != pi.shape:
        raise ValueError("the field and the image differ in shape")
    rhs = torch.empty((c, hh - 2, ww - 2), dtype=torch.float64, device=pi.device)
    _lib.check(_lib.lib().gdp_masked_rhs(
        ctypes.c_void_p(pgx.data_ptr()), ctypes.c_void_p(pgy.data_ptr()),
        ctypes.c_void_p(pi.data_ptr()), ctypes.c_void_p(inner.data_ptr()), h, w, c, top, left,
        hh, ww, ctypes.c_void_p(rhs.data_ptr()), _stream()), "gdp_masked_rhs")
    unknown = inner[top + 1:bottom - 1, left + 1:right - 1].contiguous()
    sol, reports = _solve_masked_planes(rhs, unknown, method, tolerance, max_iterations)
    out = pi.clone()
    win = out[:, top + 1:bottom - 1, left + 1:right - 1]
    win.copy_(torch.where(unknown.bool(), sol, win))
    return to_host_like(_unplanes(out, cl), host), reports


def solve_dirichlet(field: GuidanceField, boundary, method: str = "mgcg",
                    tolerance: float = 1e-8, max_iterations: int = 2000):
    """Integrate ``field`` over a rectangle whose ring is fixed to ``boundary``
    (poisson.py:102-147).  Returns ``(image, reports)``, one report per channel."""
    bd, host = to_device(boundary)
    pb, cl = _planes(bd)
    c, h, w = pb.shape
    if h < 3 or w < 3:
        raise ValueError("the rectangle needs an interior (at least 3 x 3)")
    rhs = divergence(to_device(field.gx)[0], to_device(field.gy)[0])
    prhs, _ = _planes(rhs)
    folded = fold_boundary(_unplanes(prhs[:, 1:-1, 1:-1].contiguous(), cl), bd)
    pf, _ = _planes(folded)
    sol, reports = _solve_window(pf, method, tolerance, max_iterations)
    out = pb.clone()
    out[:, 1:-1, 1:-1] = sol
    return to_host_like(_unplanes(out, cl), host), reports


@dataclass
class CloneJob:
    """A prepared seamless_clone: the target planes, the window and its right-hand side
    (planes x (hh - 2) x (ww - 2)); ``rhs`` is None for an empty selection.  ``unknown``
    is the window interior's mask of unknowns for the mask route (None: rectangle)."""

    target: torch.Tensor
    channels_last: bool
    host: bool
    window: tuple
    rhs: torch.Tensor | None
    unknown: torch.Tensor | None = None


def prepare_clone(source, target, mask, offset=(0, 0), mode: str = "import",
                  domain: str = "mask", margin: int = 24) -> CloneJob:
    """Validation, window and fused right-hand side of seamless_clone (device work on
    the current stream; one small host read of the selection's bounding box)."""
    if mode not in MODES:
        raise ValueError(f"unknown mode {mode!r}")
    if domain not in ("rectangle", "mask"):
        raise ValueError(f"unknown domain {domain!r}")
    tgt, host = to_device(target)
    src, _ = to_device(source)
    if src.ndim != tgt.ndim:
        raise ValueError("source and target need the same number of dimensions")
    pt, cl = _planes(tgt)
    ps, _ = _planes(src)
    c, h, w = pt.shape
    if ps.shape[0] != c:
        raise ValueError("source and target need the same number of channels")
    sh, sw = ps.shape[1:]
    row, col = int(offset[0]), int(offset[1])
    if row < 0 or col < 0 or row + sh > h or col + sw > w:
        raise ValueError("the source crop does not fit in the target at this offset")
    m = _mask_tensor(mask, (sh, sw), pt.device)
    lib, args = _lib.lib(), None
    if domain == "mask":
        placed = torch.zeros((h, w), dtype=torch.uint8, device=pt.device)
        placed[row:row + sh, col:col + sw] = m
        inner = _interior(placed)
        box = _box(inner)
        if box is None:
            return CloneJob(pt, cl, host, (0, h, 0, w), None)
        top, bottom, left, right = box
    else:
        nz = torch.nonzero(m)
        if nz.numel() == 0:
            return CloneJob(pt, cl, host, (0, h, 0, w), None)
        r0, c0 = (nz.min(dim=0).values + torch.tensor([row, col], device=nz.device)).tolist()
        r1, c1 = (nz.max(dim=0).values + torch.tensor([row, col], device=nz.device)).tolist()
        top, bottom = max(r0 - margin, 0), min(r1 + margin + 1, h)
        left, right = max(c0 - margin, 0), min(c1 + margin + 1, w)
    hh, ww = bottom - top, right - left
    if hh < 3 or ww < 3:
        raise ValueError("the window around the selection has no interior")
    rhs = torch.empty((c, hh - 2, ww - 2), dtype=torch.float64, device=pt.device)
    common = (ctypes.c_void_p(ps.data_ptr()), sh, sw, row, col, ctypes.c_void_p(m.data_ptr()),
              ctypes.c_void_p(pt.data_ptr()), h, w, c, top, left, hh, ww, MODES[mode])
    if domain == "mask":
        _lib.check(lib.gdp_clone_rhs_masked(*common, ctypes.c_void_p(inner.data_ptr()),
                                            ctypes.c_void_p(rhs.data_ptr()), _stream()),
                   "gdp_clone_rhs_masked")
        unknown = inner[top + 1:bottom - 1, left + 1:right - 1].contiguous()
        return CloneJob(pt, cl, host, (top, bottom, left, right), rhs, unknown)
    _lib.check(lib.gdp_clone_rhs(*common, ctypes.c_void_p(rhs.data_ptr()), _stream()),
               "gdp_clone_rhs")
    return CloneJob(pt, cl, host, (top, bottom, left, right), rhs)


def finish_clone(job: CloneJob, solution: torch.Tensor | None):
    """The composited image: the target with the window's unknowns replaced."""
    out = job.target.clone()
    if solution is not None:
        top, bottom, left, right = job.window
        win = out[:, top + 1:bottom - 1, left + 1:right - 1]
        if job.unknown is None:
            win.copy_(solution)
        else:
            win.copy_(torch.where(job.unknown.bool(), solution, win))
    return to_host_like(_unplanes(out, job.channels_last), job.host)


def seamless_clone(source, target, mask, offset=(0, 0), mode: str = "import",
                   domain: str = "mask", method: str = "mgcg", margin: int = 24,
                   tolerance: float = 1e-8, max_iterations: int = 2000):
    """Paste the masked region of ``source`` into ``target`` at ``offset`` without a seam
    (poisson.py:261-348, rectangle or mask route).  Returns ``(image, reports)``."""
    job = prepare_clone(source, target, mask, offset, mode, domain, margin)
    if job.rhs is None:
        return finish_clone(job, None), []
    sol, reports = _solve_window(job.rhs, method, tolerance, max_iterations, job.unknown)
    return finish_clone(job, sol), reports


def _masked_method(method: str) -> str:
    # the reference solves masked systems directly whatever method is named; here the
    # masked solve is FP64 MG-PCG unless mixed precision is asked for
    return method if method in ("mgcg", "mgcg_mixed") else "mgcg"


def texture_flatten(image, edges, method: str = "mgcg", mask=None, tolerance: float = 1e-8,
                    max_iterations: int = 2000):
    """Remove texture and keep the edges that define shapes (poisson.py:351-388): the
    image's gradient times the ``edges`` weights (h, w) in [0, 1], integrated over the
    whole image with its ring fixed, or over ``mask`` with every other pixel fixed.
    Returns ``(result, reports)``."""
    field = GuidanceField.of(image).scaled(edges)
    if mask is not None:
        return solve_masked(field, image, mask, method=_masked_method(method),
                            tolerance=tolerance, max_iterations=max_iterations)
    return solve_dirichlet(field, image, method=method, tolerance=tolerance,
                           max_iterations=max_iterations)


def illumination_change(image, mask, alpha: float = 0.2, beta: float = 0.2,
                        epsilon: float = 1e-4, method: str = "mgcg", tolerance: float = 1e-8,
                        max_iterations: int = 2000):
    """Compress the local dynamic range inside a selection (poisson.py:391-437): every
    gradient of the image times ``alpha ** beta * max(|g|, epsilon) ** -beta`` (fused on
    the device, per channel), integrated over ``mask`` with every other pixel fixed.
    Returns ``(result, reports)``."""
    field = GuidanceField.of(image)
    gx, host = to_device(field.gx)
    gy, _ = to_device(field.gy)
    gx, gy = gx.contiguous(), gy.contiguous()
    ox, oy = torch.empty_like(gx), torch.empty_like(gy)
    _lib.check(_lib.lib().gdp_illumination_field(
        ctypes.c_void_p(gx.data_ptr()), ctypes.c_void_p(gy.data_ptr()), gx.numel(),
        float(alpha), float(beta), float(epsilon), ctypes.c_void_p(ox.data_ptr()),
        ctypes.c_void_p(oy.data_ptr()), _stream()), "gdp_illumination_field")
    attenuated = GuidanceField(to_host_like(ox, host), to_host_like(oy, host))
    return solve_masked(attenuated, image, mask, method=_masked_method(method),
                        tolerance=tolerance, max_iterations=max_iterations)
