"""Discrete operators on a regular grid (FP64 numpy restatement).

Follows ``gradient-domain/src/gradient_domain/operators.py``:

* ``gradient``        operators.py:30-50   forward differences, last row/col zero
* ``divergence``      operators.py:53-83   negative adjoint of ``gradient``
* ``laplacian``       operators.py:86-107  5-point, zero outside (homogeneous Dirichlet)
* ``sparse_laplacian`` operators.py:110-136 CSR Kronecker form of ``laplacian``
* ``fold_boundary``   operators.py:139-169 Dirichlet ring moved into the RHS

``neumann_laplacian`` is new: ``divergence(gradient(u)) / h**2``, i.e. the 5-point
stencil whose diagonal is minus the number of in-grid neighbours (the composition
the reference pins at ``tests/test_operators.py:51-57``).
"""

from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def gradient(image: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Forward differences; the last column of gx and last row of gy are zero."""
    gx = np.zeros_like(image, dtype=float)
    gy = np.zeros_like(image, dtype=float)
    gx[:, :-1] = image[:, 1:] - image[:, :-1]
    gy[:-1, :] = image[1:, :] - image[:-1, :]
    return gx, gy


def divergence(gx: np.ndarray, gy: np.ndarray) -> np.ndarray:
    """Negative adjoint of :func:`gradient` (backward differences, ends special)."""
    dx = np.zeros_like(gx, dtype=float)
    dy = np.zeros_like(gy, dtype=float)
    if gx.shape[1] > 1:
        dx[:, 0] = gx[:, 0]
        dx[:, 1:-1] = gx[:, 1:-1] - gx[:, :-2]
        dx[:, -1] = -gx[:, -2]
    if gy.shape[0] > 1:
        dy[0, :] = gy[0, :]
        dy[1:-1, :] = gy[1:-1, :] - gy[:-2, :]
        dy[-1, :] = -gy[-2, :]
    return dx + dy


def laplacian(u: np.ndarray, spacing: float = 1.0) -> np.ndarray:
    """``(sum of 4 neighbours - 4 u) / h^2`` with zeros outside the array."""
    pad = ((1, 1), (1, 1)) + ((0, 0),) * (u.ndim - 2)
    p = np.pad(u, pad)
    nb = p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:]
    return (nb - 4.0 * u) / (spacing * spacing)


def neumann_laplacian(u: np.ndarray, spacing: float = 1.0) -> np.ndarray:
    """``sum over in-grid neighbours of (u_nbr - u) / h^2`` (edge-padded stencil)."""
    pad = ((1, 1), (1, 1)) + ((0, 0),) * (u.ndim - 2)
    p = np.pad(u, pad, mode="edge")
    nb = p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:]
    return (nb - 4.0 * u) / (spacing * spacing)


def sparse_laplacian(shape: tuple[int, int], spacing: float = 1.0) -> sp.csr_matrix:
    """CSR matrix of :func:`laplacian` in C (row-major) order."""
    rows, cols = shape

    def second_difference(size: int) -> sp.csr_matrix:
        return sp.diags(
            [np.ones(size - 1), np.full(size, -2.0), np.ones(size - 1)], [-1, 0, 1], format="csr"
        )

    op = sp.kron(second_difference(rows), sp.identity(cols, format="csr")) + sp.kron(
        sp.identity(rows, format="csr"), second_difference(cols)
    )
    return (op / (spacing * spacing)).tocsr()


def fold_boundary(rhs: np.ndarray, boundary: np.ndarray, spacing: float = 1.0) -> np.ndarray:
    """Subtract the Dirichlet ring of ``boundary`` (divided by h^2) from the interior RHS."""
    folded = np.array(rhs, dtype=float, copy=True)
    w = 1.0 / (spacing * spacing)
    folded[0] -= w * boundary[0, 1:-1]
    folded[-1] -= w * boundary[-1, 1:-1]
    folded[:, 0] -= w * boundary[1:-1, 0]
    folded[:, -1] -= w * boundary[1:-1, -1]
    return folded
