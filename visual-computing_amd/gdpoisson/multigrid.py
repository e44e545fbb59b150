"""Geometric multigrid on the GPU with the interface of
``gradient-domain/src/gradient_domain/multigrid.py``.

``MultigridOptions`` / ``MultigridReport`` keep the reference fields and defaults
(multigrid.py:46-108).  ``v_cycle`` and ``solve`` run on libgdpoisson plans; the
keyword extensions ``boundary`` ("dirichlet" as in the reference, or "neumann") and
``precision`` ("fp64", or "mixed" for an FP32 cycle with FP64 residuals) default to
the reference behaviour.
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from . import _lib
from ._device import cached_plan, to_device, to_host_like
from .plan import PlanOptions


@dataclass
class MultigridOptions:
    pre_smoothing: int = 2
    post_smoothing: int = 2
    coarsest_size: int = 8
    coarsest_sweeps: int = 60
    max_cycles: int = 60
    tolerance: float = 1e-8


@dataclass
class MultigridReport:
    residuals: list
    cycles: int
    converged: bool

    @property
    def convergence_factor(self) -> float:
        """Geometric mean of the per-cycle residual reduction (multigrid.py:91-102)."""
        if len(self.residuals) < 2 or self.residuals[0] <= 0:
            return float("nan")
        ratio = self.residuals[-1] / self.residuals[0]
        return float(ratio ** (1.0 / (len(self.residuals) - 1)))

    def __str__(self) -> str:
        return (
            f"{self.cycles} cycles, relative residual {self.residuals[-1]:.2e}, "
            f"convergence factor {self.convergence_factor:.3f}"
        )


def _plan_options(options: MultigridOptions, precision: str, boundary: str,
                  max_iterations: int | None = None, tolerance: float | None = None):
    cs = options.coarsest_size
    return PlanOptions(
        precision=precision,
        pre_smoothing=options.pre_smoothing,
        post_smoothing=options.post_smoothing,
        coarsest_size=cs,
        coarsest_sweeps=options.coarsest_sweeps,
        max_iterations=options.max_cycles if max_iterations is None else max_iterations,
        tolerance=options.tolerance if tolerance is None else tolerance,
        use_graph=False,
    )


def _coarse_shape(shape, boundary: str = "dirichlet"):
    if boundary == "neumann":
        return ((shape[0] + 1) // 2, (shape[1] + 1) // 2)
    return ((shape[0] - 1) // 2, (shape[1] - 1) // 2)


def _smooth(u, b, spacing: float, sweeps: int, boundary: str = "dirichlet", order: int = 0):
    """Red-black Gauss-Seidel sweeps (multigrid.py:147-166); returns the new iterate."""
    tu, host = to_device(u)
    tb, _ = to_device(b)
    out = tu.clone()
    m, n = out.shape
    bc = _lib.NEUMANN if boundary == "neumann" else _lib.DIRICHLET
    _lib.check(_lib.lib().gdp_smooth(out.data_ptr(), tb.data_ptr(), m, n, float(spacing),
                                     int(sweeps), bc, int(order), _lib.current_stream()),
               "smooth")
    return to_host_like(out, host)


def _restrict(residual, boundary: str = "dirichlet"):
    """Full-weighting restriction ``0.25 P_r^T r P_c`` (multigrid.py:169-174)."""
    t, host = to_device(residual)
    m, n = t.shape
    mc, nc = _coarse_shape((m, n), boundary)
    out = torch.empty((mc, nc), dtype=torch.float64, device=t.device)
    bc = _lib.NEUMANN if boundary == "neumann" else _lib.DIRICHLET
    _lib.check(_lib.lib().gdp_restrict(t.data_ptr(), out.data_ptr(), m, n, mc, nc, bc,
                                       _lib.current_stream()), "restrict")
    return to_host_like(out, host)


def _prolong(correction, shape, boundary: str = "dirichlet"):
    """Bilinear prolongation ``P_r e P_c^T`` (multigrid.py:177-181)."""
    t, host = to_device(correction)
    mc, nc = t.shape
    m, n = int(shape[0]), int(shape[1])
    out = torch.empty((m, n), dtype=torch.float64, device=t.device)
    bc = _lib.NEUMANN if boundary == "neumann" else _lib.DIRICHLET
    _lib.check(_lib.lib().gdp_prolong(t.data_ptr(), out.data_ptr(), m, n, mc, nc, bc,
                                      _lib.current_stream()), "prolong")
    return to_host_like(out, host)


def v_cycle(u, b, spacing: float, options: MultigridOptions, boundary: str = "dirichlet",
            precision: str = "fp64"):
    """One V-cycle on ``laplacian(u) = b`` (multigrid.py:184-215)."""
    tu, host = to_device(u)
    tb, _ = to_device(b)
    plan = cached_plan([tuple(tb.shape)], boundary, float(spacing),
                       _plan_options(options, precision, boundary, max_iterations=1))
    (out,) = plan.vcycle([tu], [tb])
    return to_host_like(out, host)


def solve(b, spacing: float = 1.0, options: MultigridOptions | None = None, initial=None,
          boundary: str = "dirichlet", precision: str = "fp64"):
    """Repeated V-cycles until the relative residual < tolerance (multigrid.py:218-255)."""
    options = options or MultigridOptions()
    tb, host = to_device(b)
    plan = cached_plan([tuple(tb.shape)], boundary, float(spacing),
                       _plan_options(options, precision, boundary))
    init = None if initial is None else [to_device(initial)[0]]
    (u,), (rep,), (hist,) = plan.mg_solve([tb], init)
    if rep.iterations == 0 and rep.converged:  # zero right-hand side (multigrid.py:245-247)
        hist = [0.0]
    return to_host_like(u, host), MultigridReport(
        residuals=list(hist), cycles=int(rep.iterations), converged=bool(rep.converged)
    )
