"""Geometric multigrid V-cycle (FP64 numpy restatement + Neumann variant).

Dirichlet part follows ``gradient-domain/src/gradient_domain/multigrid.py``:

* ``MultigridOptions``  multigrid.py:46-71   (pre/post 2, coarsest 8, 60 sweeps, 60 cycles, 1e-8)
* ``MultigridReport``   multigrid.py:74-108  (residual history, geometric-mean factor)
* ``prolongation_1d``   multigrid.py:111-133 (coarse I -> fine 2I+1, halves between, zero outside)
* ``parity_mask``       multigrid.py:136-140 ((i + j) % 2 == parity)
* ``coarse_shape``      multigrid.py:143-144 ((m - 1) // 2)
* ``smooth``            multigrid.py:147-166 (RB Gauss-Seidel, colour 0 then colour 1)
* ``restrict``          multigrid.py:169-174 (0.25 * P_r^T r P_c)
* ``prolong``           multigrid.py:177-181 (P_r e P_c^T)
* ``v_cycle``           multigrid.py:184-215 (coarsest: min side <= coarsest_size or
                                              min coarse side < 2 -> coarsest_sweeps sweeps)
* ``solve``             multigrid.py:218-255 (repeated V-cycles, true residual per cycle)

Neumann part (new; the reference has no Neumann multigrid, it solves Neumann
problems with the DCT in ``solvers.py:224-262``):

* cell-centred coarsening ``m -> (m + 1) // 2``; coarse cell I covers fine cells 2I, 2I+1;
* prolongation: 1-D linear interpolation at cell centres (3/4 parent, 1/4 next-nearest
  coarse cell, reflected at the ends so constants are reproduced exactly), 2-D tensor;
* restriction ``0.25 * P^T`` (scaled adjoint, as in the Dirichlet pair);
* operator ``neumann_laplacian`` (in-grid neighbours only), rediscretised at 2h;
* coarsest level (``max side <= coarsest_size`` or ``min side < 3``): exact FP64 solve by
  the DCT-II diagonalisation of ``solvers.solve_neumann`` (zero-mean solution);

Dirichlet extension: ``coarsest_sweeps == 0`` replaces the coarsest sweeps by the exact
FP64 solve of the coarsest Dirichlet system, diagonalised by the DST-I (eigenvalues
``2 cos(pi k/(m+1)) + 2 cos(pi l/(n+1)) - 4``).
* symmetric cycle: pre-smoothing colour 0 then 1, post-smoothing colour 1 then 0.

Every function takes a ``dtype`` so the same code also describes the FP32 V-cycle of
the mixed-precision solver; the coarsest solve always runs in FP64.
"""

from __future__ import annotations

from dataclasses import dataclass
from functools import lru_cache

import numpy as np
import scipy.sparse as sp
from scipy.fft import dctn, dstn, idctn, idstn

from .operators import laplacian, neumann_laplacian

DIRICHLET = "dirichlet"
NEUMANN = "neumann"


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
        if len(self.residuals) < 2 or self.residuals[0] <= 0:
            return float("nan")
        ratio = self.residuals[-1] / self.residuals[0]
        return float(ratio ** (1.0 / (len(self.residuals) - 1)))


# ----------------------------------------------------------------------------- transfers
@lru_cache(maxsize=128)
def prolongation_1d(fine: int, coarse: int) -> sp.csr_matrix:
    """Dirichlet (vertex-centred) 1-D interpolation, multigrid.py:111-133."""
    rows, cols, vals = [], [], []
    for i in range(fine):
        if i % 2 == 1:
            s = (i - 1) // 2
            if 0 <= s < coarse:
                rows.append(i), cols.append(s), vals.append(1.0)
        else:
            for s in (i // 2 - 1, i // 2):
                if 0 <= s < coarse:
                    rows.append(i), cols.append(s), vals.append(0.5)
    return sp.csr_matrix((vals, (rows, cols)), shape=(fine, coarse))


@lru_cache(maxsize=128)
def neumann_prolongation_1d(fine: int, coarse: int) -> sp.csr_matrix:
    """Cell-centred 1-D linear interpolation with reflection at both ends.

    Fine cell i has parent I = i // 2 (weight 3/4) and the next-nearest coarse cell
    J = I - 1 for even i, I + 1 for odd i (weight 1/4).  A J outside [0, coarse)
    is reflected onto I, so every row sums to one.
    """
    rows, cols, vals = [], [], []
    for i in range(fine):
        parent = i // 2
        other = parent - 1 if i % 2 == 0 else parent + 1
        if other < 0 or other >= coarse:
            rows.append(i), cols.append(parent), vals.append(1.0)
        else:
            rows.append(i), cols.append(parent), vals.append(0.75)
            rows.append(i), cols.append(other), vals.append(0.25)
    return sp.csr_matrix((vals, (rows, cols)), shape=(fine, coarse))


@lru_cache(maxsize=128)
def parity_mask(shape: tuple, parity: int) -> np.ndarray:
    r = np.arange(shape[0])[:, None]
    c = np.arange(shape[1])[None, :]
    return (r + c) % 2 == parity


@lru_cache(maxsize=128)
def neighbour_count(shape: tuple) -> np.ndarray:
    """Number of in-grid 4-neighbours of every cell (the Neumann diagonal)."""
    cnt = np.full(shape, 4.0)
    cnt[0, :] -= 1
    cnt[-1, :] -= 1
    cnt[:, 0] -= 1
    cnt[:, -1] -= 1  # a 1-wide axis is hit twice above: both neighbours missing
    return cnt


def coarse_shape(shape: tuple, bc: str = DIRICHLET) -> tuple:
    if bc == NEUMANN:
        return ((shape[0] + 1) // 2, (shape[1] + 1) // 2)
    return ((shape[0] - 1) // 2, (shape[1] - 1) // 2)


def is_coarsest(shape: tuple, options: MultigridOptions, bc: str = DIRICHLET) -> bool:
    if bc == NEUMANN:
        return max(shape) <= options.coarsest_size or min(shape) < 3
    return min(shape) <= options.coarsest_size or min(coarse_shape(shape)) < 2


def level_shapes(shape: tuple, options: MultigridOptions, bc: str = DIRICHLET) -> list:
    """Shapes of every level from fine to coarsest."""
    shapes = [tuple(shape)]
    while not is_coarsest(shapes[-1], options, bc):
        shapes.append(coarse_shape(shapes[-1], bc))
    return shapes


def apply_operator(u: np.ndarray, spacing: float, bc: str = DIRICHLET) -> np.ndarray:
    if bc == NEUMANN:
        return neumann_laplacian(u, spacing)
    return laplacian(u, spacing)


# ----------------------------------------------------------------------------- smoother
def smooth(u, b, spacing, sweeps, bc=DIRICHLET, order=(0, 1)):
    """Red-black Gauss-Seidel on ``A u = b`` (multigrid.py:147-166).

    Dirichlet: ``u <- (sum of 4 neighbours, zero outside - h^2 b) / 4``.
    Neumann:   ``u <- (sum of in-grid neighbours - h^2 b) / count``.
    Arithmetic is done in ``u.dtype``.
    """
    dt = u.dtype
    h2 = dt.type(spacing * spacing)
    b = b.astype(dt, copy=False)
    if bc == NEUMANN:
        cnt = neighbour_count(u.shape).astype(dt)
    for _ in range(sweeps):
        for parity in order:
            if bc == NEUMANN:
                p = np.pad(u, 1, mode="edge")
                nb = p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:]
                nb = nb - (dt.type(4) - cnt) * u
                upd = (nb - h2 * b) / cnt
            else:
                p = np.pad(u, 1)
                nb = p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:]
                upd = dt.type(0.25) * (nb - h2 * b)
            u = np.where(parity_mask(u.shape, parity), upd, u)
    return u


def restrict(residual: np.ndarray, bc: str = DIRICHLET) -> np.ndarray:
    """``0.25 * P_r^T r P_c`` (multigrid.py:169-174; cell-centred P for Neumann)."""
    cm, cn = coarse_shape(residual.shape, bc)
    pfun = neumann_prolongation_1d if bc == NEUMANN else prolongation_1d
    pr = pfun(residual.shape[0], cm)
    pc = pfun(residual.shape[1], cn)
    out = (pr.T @ residual.astype(np.float64) @ pc) * 0.25
    return out.astype(residual.dtype)


def prolong(correction: np.ndarray, shape: tuple, bc: str = DIRICHLET) -> np.ndarray:
    """``P_r e P_c^T`` (multigrid.py:177-181; cell-centred P for Neumann)."""
    pfun = neumann_prolongation_1d if bc == NEUMANN else prolongation_1d
    pr = pfun(shape[0], correction.shape[0])
    pc = pfun(shape[1], correction.shape[1])
    out = pr @ correction.astype(np.float64) @ pc.T
    return out.astype(correction.dtype)


def dct_neumann_solve(b: np.ndarray, spacing: float = 1.0) -> np.ndarray:
    """Zero-mean solution of ``neumann_laplacian(u, h) = b`` (solvers.py:224-262, scaled)."""
    b = np.asarray(b, dtype=np.float64) * (spacing * spacing)
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


def dst_dirichlet_solve(b: np.ndarray, spacing: float = 1.0) -> np.ndarray:
    """Exact solution of ``laplacian(u, h) = b`` (zero outside) by DST-I."""
    b = np.asarray(b, dtype=np.float64) * (spacing * spacing)
    m, n = b.shape
    lam = (
        2.0 * np.cos(np.pi * np.arange(1, m + 1) / (m + 1))[:, None]
        + 2.0 * np.cos(np.pi * np.arange(1, n + 1) / (n + 1))[None, :]
        - 4.0
    )
    return idstn(dstn(b, type=1, norm="ortho") / lam, type=1, norm="ortho")


def coarsest_solve(u, b, spacing, options, bc):
    if bc == NEUMANN:
        res = b.astype(np.float64) - neumann_laplacian(u.astype(np.float64), spacing)
        return (u.astype(np.float64) + dct_neumann_solve(res, spacing)).astype(u.dtype)
    if options.coarsest_sweeps == 0:
        res = b.astype(np.float64) - laplacian(u.astype(np.float64), spacing)
        return (u.astype(np.float64) + dst_dirichlet_solve(res, spacing)).astype(u.dtype)
    # Dirichlet: coarsest_sweeps RB-GS sweeps (multigrid.py:206-207), done in FP64.
    out = smooth(u.astype(np.float64), b.astype(np.float64), spacing, options.coarsest_sweeps)
    return out.astype(u.dtype)


# ----------------------------------------------------------------------------- V-cycle
def v_cycle(u, b, spacing, options, bc=DIRICHLET):
    """One V-cycle on ``A u = b`` (multigrid.py:184-215).

    Dirichlet: pre and post smoothing both sweep colour 0 then colour 1, as the
    reference does.  Neumann: post smoothing runs colour 1 then colour 0 so the cycle
    is a symmetric operator.
    """
    if is_coarsest(u.shape, options, bc):
        return coarsest_solve(u, b, spacing, options, bc)
    post_order = (1, 0) if bc == NEUMANN else (0, 1)
    u = smooth(u, b, spacing, options.pre_smoothing, bc, (0, 1))
    res = b - apply_operator(u, spacing, bc).astype(u.dtype)
    cs = coarse_shape(u.shape, bc)
    ec = v_cycle(np.zeros(cs, dtype=u.dtype), restrict(res, bc), 2.0 * spacing, options, bc)
    u = u + prolong(ec, u.shape, bc)
    return smooth(u, b, spacing, options.post_smoothing, bc, post_order)


def solve(b, spacing=1.0, options=None, initial=None, bc=DIRICHLET):
    """Repeated V-cycles until the true relative residual < tolerance (multigrid.py:218-255)."""
    options = options or MultigridOptions()
    b = np.asarray(b, dtype=float)
    u = np.zeros_like(b) if initial is None else np.array(initial, dtype=float, copy=True)
    scale = float(np.linalg.norm(b))
    if scale == 0.0:
        return u, MultigridReport(residuals=[0.0], cycles=0, converged=True)
    res = [float(np.linalg.norm(b - apply_operator(u, spacing, bc))) / scale]
    for cycle in range(1, options.max_cycles + 1):
        u = v_cycle(u, b, spacing, options, bc)
        res.append(float(np.linalg.norm(b - apply_operator(u, spacing, bc))) / scale)
        if res[-1] < options.tolerance:
            return u, MultigridReport(residuals=res, cycles=cycle, converged=True)
    return u, MultigridReport(residuals=res, cycles=options.max_cycles, converged=False)
