"""Discrete operators on the GPU, with the signatures of
``gradient-domain/src/gradient_domain/operators.py``.

numpy in -> numpy out (the drop-in form); a CUDA (HIP) tensor in -> tensor out on
the same device.  Arrays are ``(h, w)`` or channels-last ``(h, w, c)`` like the
reference; channels are processed as independent planes by one launch.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from ._device import to_device, to_host_like


def _planes(t: torch.Tensor) -> tuple[torch.Tensor, bool]:
    """(planes, h, w) contiguous view of a 2-D or channels-last 3-D tensor."""
    if t.ndim == 2:
        return t.unsqueeze(0).contiguous(), False
    if t.ndim == 3:
        return t.permute(2, 0, 1).contiguous(), True
    raise ValueError(f"expected a 2-D or 3-D array, got shape {tuple(t.shape)}")


def _unplanes(p: torch.Tensor, channels_last: bool) -> torch.Tensor:
    return p.permute(1, 2, 0).contiguous() if channels_last else p[0]


def _stream():
    return _lib.current_stream()


def gradient(image):
    """Forward differences, last column / row zero (operators.py:30-50)."""
    t, host = to_device(image)
    p, cl = _planes(t)
    gx = torch.empty_like(p)
    gy = torch.empty_like(p)
    c, h, w = p.shape
    _lib.check(_lib.lib().gdp_gradient(p.data_ptr(), gx.data_ptr(), gy.data_ptr(), h, w, c,
                                       _stream()), "gradient")
    return to_host_like(_unplanes(gx, cl), host), to_host_like(_unplanes(gy, cl), host)


def divergence(gx, gy):
    """Negative adjoint of :func:`gradient` (operators.py:53-83)."""
    tx, host = to_device(gx)
    ty, _ = to_device(gy)
    if tx.shape != ty.shape:
        raise ValueError("gx and gy must have the same shape")
    px, cl = _planes(tx)
    py, _ = _planes(ty)
    out = torch.empty_like(px)
    c, h, w = px.shape
    _lib.check(_lib.lib().gdp_divergence(px.data_ptr(), py.data_ptr(), out.data_ptr(), h, w, c,
                                         _stream()), "divergence")
    return to_host_like(_unplanes(out, cl), host)


def _apply(u, spacing, bc):
    t, host = to_device(u)
    p, cl = _planes(t)
    out = torch.empty_like(p)
    c, m, n = p.shape
    _lib.check(_lib.lib().gdp_laplacian(p.data_ptr(), out.data_ptr(), m, n, c, float(spacing),
                                        bc, _stream()), "laplacian")
    return to_host_like(_unplanes(out, cl), host)


def laplacian(u, spacing: float = 1.0):
    """5-point Laplacian with zeros outside (operators.py:86-107)."""
    return _apply(u, spacing, _lib.DIRICHLET)


def neumann_laplacian(u, spacing: float = 1.0):
    """``divergence(gradient(u)) / h^2``: the in-grid (edge-padded) 5-point stencil."""
    return _apply(u, spacing, _lib.NEUMANN)


def fold_boundary(right_hand_side, boundary, spacing: float = 1.0):
    """Move the Dirichlet ring of ``boundary`` into the interior RHS (operators.py:139-169)."""
    r, host = to_device(right_hand_side)
    bd, _ = to_device(boundary)
    pr, cl = _planes(r)
    pb, _ = _planes(bd)
    c, h, w = pb.shape
    if pr.shape != (c, h - 2, w - 2):
        raise ValueError("right_hand_side must be the interior (h - 2, w - 2) of boundary")
    out = torch.empty_like(pr)
    _lib.check(_lib.lib().gdp_fold_boundary(pr.data_ptr(), pb.data_ptr(), out.data_ptr(), h, w,
                                            c, float(spacing), _stream()), "fold_boundary")
    return to_host_like(_unplanes(out, cl), host)


def sparse_laplacian(shape, spacing: float = 1.0):
    """The five-point operator of :func:`laplacian` as a scipy CSR matrix
    (operators.py:110-136): ``(T_m kron I_n + I_m kron T_n) / h^2`` in row-major order,
    T = tridiag(1, -2, 1).  A host helper (the reference's tests and the direct method's
    definition use it); the device solvers never form it."""
    import scipy.sparse as sp

    m, n = int(shape[0]), int(shape[1])

    def tri(k):
        return sp.diags([np.ones(k - 1), np.full(k, -2.0), np.ones(k - 1)], [-1, 0, 1],
                        format="csr")

    a = sp.kron(tri(m), sp.identity(n, format="csr")) + sp.kron(sp.identity(m, format="csr"),
                                                               tri(n))
    return (a / (spacing * spacing)).tocsr()
