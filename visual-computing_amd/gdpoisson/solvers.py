"""Poisson solvers on the GPU with the interface of
``gradient-domain/src/gradient_domain/solvers.py``.

``SOLVERS`` keeps the reference's plug-in signature
``fn(b, spacing, tolerance, max_iterations) -> (u, iterations, converged)``
(solvers.py:163-168) and ``solve_system`` its contract (solvers.py:171-221):
``KeyError`` for an unknown method, ``b`` taken as float64, a zero right-hand side
returns zeros with residual 0.0, the report carries the true relative residual.

Methods:

``direct``      the exact solve (solvers.py:85-90, SuperLU there): DST-I diagonalisation of
                the rectangle operator on the device, for sides m with 2 (m + 1) a product
                of 2, 3 and 5 up to 4096; other sides are solved by FP64 MG-PCG to a
                relative residual of 1e-13.
``mgcg``        V-cycle-preconditioned CG, FP64 throughout (solvers.py:134-160).
``mgcg_mixed``  the same CG with an FP32 V-cycle (r, x, d, q, dots and the coarsest
                solve stay FP64) and the Polak-Ribiere beta.
``multigrid``   repeated V-cycles, nu = 2/2 (solvers.py:121-131).
``cg``          Jacobi-preconditioned CG (solvers.py:93-118).

``boundary="neumann"`` (keyword extension) solves ``div(grad u) / h^2 = b`` for the
zero-mean compatible part of ``b`` and returns the zero-mean solution, the
convention of ``solve_neumann`` (solvers.py:224-262).
"""

from __future__ import annotations

import ctypes
import threading

import time
from dataclasses import dataclass

import torch

from . import _lib
from ._device import cached_plan, to_device, to_host_like
from .multigrid import MultigridOptions
from .plan import PlanOptions


@dataclass
class SolverReport:
    method: str
    iterations: int
    relative_residual: float
    seconds: float
    converged: bool

    def __str__(self) -> str:
        return (
            f"{self.method}: {self.iterations} iterations, "
            f"residual {self.relative_residual:.2e}, {self.seconds * 1e3:.1f} ms"
        )


def _pcg_options(precision, tolerance, max_iterations, boundary):
    return PlanOptions(
        precision=precision,
        max_iterations=int(max_iterations),
        tolerance=float(tolerance),
    )


# the fallback of "direct" for sides the DST does not cover: FP64 MG-PCG this far
DIRECT_FALLBACK_TOLERANCE = 1e-13


def _run(kind, b, spacing, tolerance, max_iterations, boundary="dirichlet"):
    """Solve one grid on the device; returns (u tensor, iterations, converged, true residual)."""
    if kind == "direct":
        if boundary != "dirichlet":
            raise ValueError("the direct method solves the Dirichlet rectangle operator "
                             "(use solve_neumann for the Neumann one)")
        m, n = int(b.shape[0]), int(b.shape[1])
        if dst_supported(m, n):
            u = _dst_solve(b, float(spacing))
            return u, 0, True, relative_residual(u, b, spacing)
        return _run("mgcg", b, spacing, DIRECT_FALLBACK_TOLERANCE, max(int(max_iterations), 200))
    if kind == "multigrid":
        opts = MultigridOptions(max_cycles=int(max_iterations), tolerance=float(tolerance))
        plan = cached_plan([tuple(b.shape)], boundary, float(spacing), PlanOptions(
            precision="fp64", pre_smoothing=opts.pre_smoothing,
            post_smoothing=opts.post_smoothing, coarsest_size=opts.coarsest_size,
            coarsest_sweeps=opts.coarsest_sweeps, max_iterations=opts.max_cycles,
            tolerance=opts.tolerance, use_graph=False))
        (u,), (rep,), _ = plan.mg_solve([b])
        return u, int(rep.iterations), bool(rep.converged), float(rep.relative_residual)
    precision = "mixed" if kind == "mgcg_mixed" else "fp64"
    plan = cached_plan([tuple(b.shape)], boundary, float(spacing),
                       _pcg_options(precision, tolerance, max_iterations, boundary))
    (u,), (rep,) = plan.mgcg([b], jacobi=(kind == "cg"))
    return u, int(rep.iterations), bool(rep.converged), float(rep.relative_residual)


def solve_planes(planes: torch.Tensor, method: str = "mgcg", tolerance: float = 1e-8,
                 max_iterations: int = 2000, spacing: float = 1.0,
                 boundary: str = "dirichlet"):
    """Solve every plane of a (c, m, n) device tensor with one batched plan.

    Returns (solutions (c, m, n), list of SolverReport); ``seconds`` is the batch's device
    time (the planes are solved together).
    """
    if method not in SOLVERS:
        raise KeyError(f"unknown solver {method!r}; choose from {sorted(SOLVERS)}")
    c, m, n = planes.shape
    bs = [planes[k] for k in range(c)]
    torch.cuda.current_stream().synchronize()
    t0 = time.perf_counter()
    if method == "direct":
        res = [_run("direct", x, spacing, tolerance, max_iterations, boundary) for x in bs]
        us = [r[0] for r in res]
        reps = [_lib.Report(r[1], int(r[2]), r[3], 0.0) for r in res]
    elif method == "multigrid":
        opts = MultigridOptions(max_cycles=int(max_iterations), tolerance=float(tolerance))
        plan = cached_plan([(m, n)] * c, boundary, float(spacing), PlanOptions(
            precision="fp64", pre_smoothing=opts.pre_smoothing,
            post_smoothing=opts.post_smoothing, coarsest_size=opts.coarsest_size,
            coarsest_sweeps=opts.coarsest_sweeps, max_iterations=opts.max_cycles,
            tolerance=opts.tolerance, use_graph=False))
        us, reps, _ = plan.mg_solve(bs)
    else:
        precision = "mixed" if method == "mgcg_mixed" else "fp64"
        plan = cached_plan([(m, n)] * c, boundary, float(spacing),
                           _pcg_options(precision, tolerance, max_iterations, boundary))
        us, reps = plan.mgcg(bs, jacobi=(method == "cg"))
    torch.cuda.current_stream().synchronize()
    dt = time.perf_counter() - t0
    reports = [SolverReport(method=method, iterations=int(r.iterations),
                            relative_residual=float(r.relative_residual), seconds=dt,
                            converged=bool(r.converged)) for r in reps]
    return torch.stack(us), reports


def _registry_entry(kind):
    def fn(b, spacing, tolerance, max_iterations):
        t, host = to_device(b)
        u, it, conv, _ = _run(kind, t, spacing, tolerance, max_iterations)
        return to_host_like(u, host), it, conv

    fn.__name__ = f"_solve_{kind}"
    return fn


SOLVERS = {
    "direct": _registry_entry("direct"),
    "cg": _registry_entry("cg"),
    "multigrid": _registry_entry("multigrid"),
    "mgcg": _registry_entry("mgcg"),
    "mgcg_mixed": _registry_entry("mgcg_mixed"),
}


def solve_system(
    b,
    method: str = "mgcg",
    spacing: float = 1.0,
    tolerance: float = 1e-8,
    max_iterations: int = 2000,
    boundary: str = "dirichlet",
):
    """Solve ``laplacian(u) = b`` (solvers.py:171-221); returns ``(u, SolverReport)``.

    ``seconds`` covers the device solve from a device-resident ``b`` to a
    device-resident ``u`` (host copies of numpy inputs are outside it).
    """
    if method not in SOLVERS:
        raise KeyError(f"unknown solver {method!r}; choose from {sorted(SOLVERS)}")
    if boundary not in ("dirichlet", "neumann"):
        raise ValueError(f"unknown boundary {boundary!r}")
    t, host = to_device(b)
    if t.ndim != 2:
        raise ValueError(f"b must be 2-D, got shape {tuple(t.shape)}")
    # a zero right-hand side converges at set-up with zero iterations and u = 0
    torch.cuda.current_stream().synchronize()
    t0 = time.perf_counter()
    u, it, conv, rel = _run(method, t, spacing, tolerance, max_iterations, boundary)
    torch.cuda.current_stream().synchronize()
    dt = time.perf_counter() - t0
    return to_host_like(u, host), SolverReport(
        method=method, iterations=it, relative_residual=rel, seconds=dt, converged=conv
    )


_DCT_TABLES: dict = {}
_DCT_LOCK = threading.Lock()


def relative_residual(u: torch.Tensor, b: torch.Tensor, spacing: float = 1.0,
                      boundary: str = "dirichlet") -> float:
    """True ``||b - A u|| / ||b||`` of one device grid (solvers.py:78-82), 0 when b = 0."""
    out = ctypes.c_double()
    _lib.check(_lib.lib().gdp_relative_residual(
        ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(b.data_ptr()), int(u.shape[0]),
        int(u.shape[1]), 1, float(spacing), _lib.NEUMANN if boundary == "neumann" else
        _lib.DIRICHLET, ctypes.byref(out), _lib.current_stream()), "gdp_relative_residual")
    return float(out.value)


def dct_supported(m: int, n: int) -> bool:
    """Whether the exact device DCT handles an m x n grid (sides <= 4096 whose prime
    factors are 2, 3 and 5)."""
    return bool(_lib.lib().gdp_dct_supported(int(m), int(n)))


def dst_supported(m: int, n: int) -> bool:
    """Whether the exact device DST-I (the direct Dirichlet solve) handles an m x n grid:
    2 (side + 1) at most 4096 with prime factors 2, 3 and 5."""
    return bool(_lib.lib().gdp_dst_supported(int(m), int(n)))


def _transform_tables(kind: str, m: int, n: int, device) -> torch.Tensor:
    """The device tables of the exact DCT ("dct") or DST ("dst") solve of an m x n grid,
    built once per shape and device and shared by every stream."""
    key = (kind, m, n, device)
    tab = _DCT_TABLES.get(key)
    if tab is None:
        L = _lib.lib()
        nbytes = ctypes.c_size_t()
        _lib.check(getattr(L, f"gdp_{kind}_tables_bytes")(m, n, ctypes.byref(nbytes)),
                   f"gdp_{kind}")
        tab = torch.empty(nbytes.value // 8, dtype=torch.float64, device=device)
        _lib.check(getattr(L, f"gdp_{kind}_tables")(m, n, ctypes.c_void_p(tab.data_ptr()),
                                                    _lib.current_stream()), f"gdp_{kind}_tables")
        # the cache is shared by every stream (serving workers): a table is published only
        # once it is written, and evicted only when no stream can still be reading it
        torch.cuda.current_stream(device).synchronize()
        with _DCT_LOCK:
            if len(_DCT_TABLES) >= 32:
                torch.cuda.synchronize(device)
                _DCT_TABLES.clear()
            _DCT_TABLES[key] = tab
    return tab


def _dst_solve(t: torch.Tensor, spacing: float) -> torch.Tensor:
    m, n = int(t.shape[0]), int(t.shape[1])
    tab = _transform_tables("dst", m, n, t.device)
    work = torch.empty(2 * m * n, dtype=torch.float64, device=t.device)
    u = torch.empty_like(t)
    _lib.check(_lib.lib().gdp_dirichlet_dst_solve(
        ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(u.data_ptr()), m, n, float(spacing),
        ctypes.c_void_p(tab.data_ptr()), ctypes.c_void_p(work.data_ptr()),
        _lib.current_stream()), "gdp_dirichlet_dst_solve")
    tab.record_stream(torch.cuda.current_stream(t.device))
    return u


def _dct_solve(t: torch.Tensor) -> torch.Tensor:
    m, n = int(t.shape[0]), int(t.shape[1])
    tab = _transform_tables("dct", m, n, t.device)
    work = torch.empty(2 * m * n, dtype=torch.float64, device=t.device)
    u = torch.empty_like(t)
    _lib.check(_lib.lib().gdp_neumann_dct_solve(
        ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(u.data_ptr()), m, n,
        ctypes.c_void_p(tab.data_ptr()), ctypes.c_void_p(work.data_ptr()),
        _lib.current_stream()), "gdp_neumann_dct_solve")
    # the table was allocated on the stream that built it: if another worker evicts it
    # while this stream still reads it, the allocator must not hand its memory out before
    # this stream's solve has run
    tab.record_stream(torch.cuda.current_stream(t.device))
    return u


def solve_neumann(b, method: str = "dct", tolerance: float = 1e-12):
    """Zero-mean solution of ``div(grad u) = b`` (solvers.py:224-262).

    ``method="dct"`` (default) is the reference's exact diagonalisation on the device:
    ``idctn(dctn(b, II, ortho) / lambda)`` with the constant mode dropped, by FFTs, for
    sides of at most 4096 points whose prime factors are 2, 3 and 5; other sides, and
    ``method="mgcg"`` / ``"mgcg_mixed"``, use the Neumann MG-PCG to ``tolerance``.  Any
    constant component of ``b`` is discarded, as in the reference.
    """
    t, host = to_device(b)
    if t.ndim != 2:
        raise ValueError(f"b must be 2-D, got shape {tuple(t.shape)}")
    if method not in ("dct", "mgcg", "mgcg_mixed"):
        raise KeyError(f"unknown method {method!r}")
    if method == "dct" and dct_supported(*t.shape):
        return to_host_like(_dct_solve(t), host)
    u, _, _, _ = _run("mgcg_mixed" if method == "mgcg_mixed" else "mgcg", t, 1.0, tolerance,
                      2000, boundary="neumann")
    return to_host_like(u, host)
