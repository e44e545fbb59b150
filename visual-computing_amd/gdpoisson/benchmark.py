"""Solver-scaling study with the interface of
``gradient-domain/src/gradient_domain/benchmark.py``: the same manufactured problem at
growing sizes, every method to the same residual, on the device solvers.

``manufactured_problem`` (benchmark.py:65-95) builds the reference's right-hand side on
the host: smooth modes plus white noise with a zero ring, ``b = laplacian(solution)``
with the five-point operator and zeros outside, so ``solution`` is the exact discrete
answer.  ``solver_scaling`` (benchmark.py:98-149) times ``solve_system`` per method and
size; the seconds are device time from a resident right-hand side to a resident
solution (``SolverReport.seconds``), the host copies of the numpy input excluded.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .solvers import solve_system

__all__ = ["BenchmarkRecord", "BenchmarkTable", "manufactured_problem", "solver_scaling"]


@dataclass
class BenchmarkRecord:
    """One measurement (benchmark.py:28-62)."""

    method: str
    shape: tuple[int, int]
    unknowns: int
    seconds: float
    iterations: int
    relative_residual: float
    max_error: float

    def as_dict(self) -> dict:
        return {
            "method": self.method,
            "shape": list(self.shape),
            "unknowns": self.unknowns,
            "seconds": self.seconds,
            "iterations": self.iterations,
            "relative_residual": self.relative_residual,
            "max_error": self.max_error,
        }


def manufactured_problem(shape, seed: int = 0):
    """(b, solution) with ``laplacian(solution) = b`` exactly (benchmark.py:65-95)."""
    rng = np.random.default_rng(seed)
    rows, columns = np.mgrid[0 : shape[0], 0 : shape[1]] / max(shape)
    solution = (
        np.sin(3.0 * np.pi * columns) * np.sin(2.0 * np.pi * rows)
        + 0.4 * np.cos(11.0 * np.pi * columns) * np.sin(7.0 * np.pi * rows)
        + 0.25 * rng.standard_normal(shape)
    )
    solution[0] = solution[-1] = 0.0
    solution[:, 0] = solution[:, -1] = 0.0
    # five-point Laplacian with zeros outside (operators.py:86-107), spacing 1
    p = np.pad(solution, 1)
    b = (p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:]) - 4.0 * solution
    return b, solution


def solver_scaling(sizes, methods=None, tolerance: float = 1e-8, max_iterations: int = 5000,
                   direct_limit: int = 4095, seed: int = 0) -> dict:
    """Time each solver on square grids of the given sides (benchmark.py:98-149).

    ``methods`` defaults to the reference's four (direct, cg, multigrid, mgcg).  The
    reference skips the direct method above 600 (SuperLU's fill-in); the device direct
    solve has none, so its limit here is the DST's largest side.
    """
    methods = list(methods or ["direct", "cg", "multigrid", "mgcg"])
    results: dict = {m: [] for m in methods}
    for size in sizes:
        b, solution = manufactured_problem((size, size), seed=seed)
        for method in methods:
            if method == "direct" and size > direct_limit:
                continue
            estimate, report = solve_system(b, method=method, tolerance=tolerance,
                                            max_iterations=max_iterations)
            results[method].append(BenchmarkRecord(
                method=method, shape=(size, size), unknowns=size * size,
                seconds=report.seconds, iterations=report.iterations,
                relative_residual=report.relative_residual,
                max_error=float(np.abs(estimate - solution).max())))
    return results


@dataclass
class BenchmarkTable:
    """Markdown view of a scaling study (benchmark.py:152-177)."""

    results: dict = field(default_factory=dict)

    def to_markdown(self) -> str:
        rows = [r for records in self.results.values() for r in records]
        rows.sort(key=lambda r: (r.unknowns, r.method))
        lines = ["| Grid | Unknowns | Method | Iterations | Seconds | Max error |",
                 "| --- | ---: | --- | ---: | ---: | ---: |"]
        for r in rows:
            it = "-" if r.iterations == 0 else str(r.iterations)
            lines.append(f"| {r.shape[0]} x {r.shape[1]} | {r.unknowns} | {r.method} | {it} | "
                         f"{r.seconds:.4f} | {r.max_error:.2e} |")
        return "\n".join(lines)
