"""Linear solvers for the interior Poisson system (FP64 numpy restatement).

Follows ``gradient-domain/src/gradient_domain/solvers.py``:

* ``SolverReport``        solvers.py:49-75
* ``relative_residual``   solvers.py:78-82   (0 when b == 0)
* ``solve_direct``        solvers.py:85-90   (SuperLU of ``sparse_laplacian``)
* ``solve_cg``            solvers.py:93-118  (Jacobi-PCG on -L u = -b)
* ``solve_multigrid``     solvers.py:121-131 (V-cycles, nu = 2/2)
* ``solve_mgcg``          solvers.py:134-160 (V-cycle-PCG, nu = 1/1, Fletcher-Reeves beta,
                                              stop on recursive ||r|| <= tol ||-b||)
* ``SOLVERS``/``solve_system`` solvers.py:163-221 (KeyError, zero-RHS shortcut, report)
* ``solve_neumann``       solvers.py:224-262 (DCT-II diagonalisation, zero-mean solution)

``mgcg`` generalises ``solve_mgcg`` with three keyword extensions that the reference
does not have: ``bc`` (Neumann uses the cell-centred cycle of ``oracle.multigrid``,
projects b onto zero mean and returns the zero-mean solution), ``precision``
("mixed" runs the V-cycle in FP32 while r, x, d, q, the dots and the coarsest solve
stay FP64) and ``flexible`` (Polak-Ribiere beta, which for this recurrence is
``beta = -alpha <z_new, q> / delta``).
"""

from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np
import scipy.sparse.linalg as spla
from scipy.fft import dctn, idctn

from .multigrid import (
    DIRICHLET,
    NEUMANN,
    MultigridOptions,
    apply_operator,
    v_cycle,
)
from .operators import laplacian, sparse_laplacian


@dataclass
class SolverReport:
    method: str
    iterations: int
    relative_residual: float
    seconds: float
    converged: bool


def relative_residual(u, b, spacing, bc=DIRICHLET) -> float:
    if bc == NEUMANN:
        b = b - b.mean()
    scale = float(np.linalg.norm(b))
    if scale == 0.0:
        return 0.0
    return float(np.linalg.norm(b - apply_operator(u, spacing, bc))) / scale


def solve_direct(b, spacing, tolerance, max_iterations):
    op = sparse_laplacian(b.shape, spacing).tocsc()
    return spla.splu(op).solve(b.reshape(-1)).reshape(b.shape), 0, True


def solve_cg(b, spacing, tolerance, max_iterations):
    diag = 4.0 / (spacing * spacing)
    rhs = -b
    u = np.zeros_like(b)
    r = rhs.copy()
    d = r / diag
    delta = float(np.sum(r * d))
    target = tolerance * float(np.linalg.norm(rhs))
    for it in range(1, max_iterations + 1):
        q = -laplacian(d, spacing)
        alpha = delta / float(np.sum(d * q))
        u += alpha * d
        r -= alpha * q
        if float(np.linalg.norm(r)) <= target:
            return u, it, True
        z = r / diag
        new_delta = float(np.sum(r * z))
        d = z + (new_delta / delta) * d
        delta = new_delta
    return u, max_iterations, False


def solve_multigrid(b, spacing, tolerance, max_iterations):
    options = MultigridOptions(max_cycles=max_iterations, tolerance=tolerance)
    u = np.zeros_like(b)
    scale = float(np.linalg.norm(b)) or 1.0
    for cycle in range(1, max_iterations + 1):
        u = v_cycle(u, b, spacing, options)
        if float(np.linalg.norm(b - laplacian(u, spacing))) / scale < tolerance:
            return u, cycle, True
    return u, max_iterations, False


def default_pcg_options(bc: str, precision: str = "fp64") -> MultigridOptions:
    """The cycle of each solver configuration.

    Dirichlet FP64: the reference mgcg cycle (nu = 1/1, solvers.py:137; coarsest <= 8
    with 60 sweeps).  Neumann: nu = 1/1 (FP64) or 2/2 (mixed), coarsening to 16 cells
    per side with the exact DCT coarsest solve.  Dirichlet mixed: nu = 2/2, coarsening
    to 16 with the exact DST coarsest solve.
    """
    nu = 2 if precision == "mixed" else 1
    if bc == NEUMANN:
        return MultigridOptions(pre_smoothing=nu, post_smoothing=nu, coarsest_size=16)
    if precision == "mixed":
        return MultigridOptions(pre_smoothing=2, post_smoothing=2, coarsest_size=16,
                                coarsest_sweeps=0)
    return MultigridOptions(pre_smoothing=1, post_smoothing=1)


def mgcg(
    b,
    spacing=1.0,
    tolerance=1e-8,
    max_iterations=2000,
    bc=DIRICHLET,
    precision="fp64",
    flexible=False,
    options=None,
    history=None,
):
    """MG-preconditioned CG on ``-A u = -b`` (solvers.py:134-160 structure)."""
    options = options or default_pcg_options(bc, precision)
    vdt = np.float32 if precision == "mixed" else np.float64
    b = np.asarray(b, dtype=np.float64)
    if bc == NEUMANN:
        b = b - b.mean()

    def precondition(res):
        z = v_cycle(np.zeros(res.shape, dtype=vdt), (-res).astype(vdt), spacing, options, bc)
        return z.astype(np.float64)

    rhs = -b
    u = np.zeros_like(b)
    r = rhs.copy()
    z = precondition(r)
    d = z.copy()
    delta = float(np.sum(r * z))
    target = tolerance * float(np.linalg.norm(rhs))
    converged = False
    it = 0
    for it in range(1, max_iterations + 1):
        q = -apply_operator(d, spacing, bc)
        alpha = delta / float(np.sum(d * q))
        u += alpha * d
        r -= alpha * q
        rn = float(np.linalg.norm(r))
        if history is not None:
            history.append(rn / max(float(np.linalg.norm(rhs)), 1e-300))
        if rn <= target:
            converged = True
            break
        z = precondition(r)
        new_delta = float(np.sum(r * z))
        if flexible:
            beta = -alpha * float(np.sum(z * q)) / delta
        else:
            beta = new_delta / delta
        d = z + beta * d
        delta = new_delta
    if bc == NEUMANN:
        u -= u.mean()
    return u, it, converged


def solve_mgcg(b, spacing, tolerance, max_iterations):
    """The reference ``_solve_mgcg`` (Dirichlet, FP64, Fletcher-Reeves)."""
    return mgcg(b, spacing, tolerance, max_iterations)


SOLVERS = {
    "direct": solve_direct,
    "cg": solve_cg,
    "multigrid": solve_multigrid,
    "mgcg": solve_mgcg,
}


def solve_system(b, method="mgcg", spacing=1.0, tolerance=1e-8, max_iterations=2000):
    """Dispatcher of solvers.py:171-221 (Dirichlet)."""
    if method not in SOLVERS:
        raise KeyError(f"unknown solver {method!r}; choose from {sorted(SOLVERS)}")
    b = np.asarray(b, dtype=float)
    if not np.any(b):
        return np.zeros_like(b), SolverReport(method, 0, 0.0, 0.0, True)
    t0 = time.perf_counter()
    u, it, conv = SOLVERS[method](b, spacing, tolerance, max_iterations)
    dt = time.perf_counter() - t0
    return u, SolverReport(method, it, relative_residual(u, b, spacing), dt, conv)


def solve_neumann(b: np.ndarray) -> np.ndarray:
    """Exact zero-mean Neumann solve by DCT-II (solvers.py:224-262)."""
    b = np.asarray(b, dtype=float)
    m, n = b.shape
    lam = (
        2.0 * np.cos(np.pi * np.arange(m) / m)[:, None]
        + 2.0 * np.cos(np.pi * np.arange(n) / n)[None, :]
        - 4.0
    )
    lam[0, 0] = 1.0
    t = dctn(b, type=2, norm="ortho") / lam
    t[0, 0] = 0.0
    return idctn(t, type=2, norm="ortho")
