"""Gradient-domain compositing, rectangle route (FP64 numpy restatement; test
infrastructure only).

Follows ``gradient-domain/src/gradient_domain/poisson.py``:

* ``GuidanceField``  poisson.py:52-92    of / select / mix_by_magnitude / scaled
* ``solve_dirichlet`` poisson.py:102-147  divergence, fold the ring, solve per channel
* ``seamless_clone``  poisson.py:245-348  place the source crop, guidance by mode
                                          (import / mixed / average), window = mask
                                          bounding box + margin, solve_dirichlet there

* ``solve_masked``    poisson.py:154-244  the system over the selected pixels off the
                                          border, fixed neighbours folded in, SuperLU
* ``seamless_clone(domain="mask")``  poisson.py:337-338
* ``texture_flatten``      poisson.py:351-388  gradient times edge weights, rectangle or
                                               masked solve
* ``illumination_change``  poisson.py:391-437  gradient times alpha**beta |g|**-beta
                                               (floored at epsilon), masked solve
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .operators import divergence, fold_boundary, gradient
from .solvers import solve_system


@dataclass
class GuidanceField:
    gx: np.ndarray
    gy: np.ndarray

    @classmethod
    def of(cls, image):
        return cls(*gradient(np.asarray(image, dtype=float)))

    def select(self, mask, other):
        sel = mask[..., None] if self.gx.ndim == 3 else mask
        return GuidanceField(np.where(sel, self.gx, other.gx), np.where(sel, self.gy, other.gy))

    def mix_by_magnitude(self, other):
        return GuidanceField(
            np.where(np.abs(self.gx) >= np.abs(other.gx), self.gx, other.gx),
            np.where(np.abs(self.gy) >= np.abs(other.gy), self.gy, other.gy),
        )

    def scaled(self, factor):
        f = factor[..., None] if (np.ndim(factor) == 2 and self.gx.ndim == 3) else factor
        return GuidanceField(self.gx * f, self.gy * f)


def _channels(image):
    image = np.asarray(image, dtype=float)
    return (image[..., None], True) if image.ndim == 2 else (image, False)


def clone_rhs(field: GuidanceField, boundary: np.ndarray) -> np.ndarray:
    """The folded interior right-hand side of every channel, (h-2, w-2, c)."""
    bd, _ = _channels(boundary)
    rhs = divergence(field.gx, field.gy)
    if rhs.ndim == 2:
        rhs = rhs[..., None]
    return np.stack([fold_boundary(rhs[1:-1, 1:-1, k], bd[..., k]) for k in range(bd.shape[2])],
                    axis=2)


def solve_dirichlet(field, boundary, method="mgcg", tolerance=1e-8, max_iterations=2000):
    bd, gray = _channels(boundary)
    folded = clone_rhs(field, bd)
    out = bd.copy()
    reports = []
    for k in range(bd.shape[2]):
        u, rep = solve_system(folded[..., k], method=method, tolerance=tolerance,
                              max_iterations=max_iterations)
        out[1:-1, 1:-1, k] = u
        reports.append(rep)
    return (out[..., 0] if gray else out), reports


_OFFSETS = ((-1, 0), (1, 0), (0, -1), (0, 1))


def solve_masked(field, image, mask):
    """Unknowns: mask without the image border; operator -4 / +1 over them; neighbours
    that are not unknowns fixed to the image (poisson.py:154-244)."""
    chans, gray = _channels(image)
    interior = np.zeros(mask.shape, dtype=bool)
    interior[1:-1, 1:-1] = np.asarray(mask, dtype=bool)[1:-1, 1:-1]
    if not interior.any():
        return np.asarray(image, dtype=float).copy(), []
    index = np.full(mask.shape, -1, dtype=np.int64)
    index[interior] = np.arange(int(interior.sum()))
    nu = int(interior.sum())
    rows, cols, vals = [np.arange(nu)], [np.arange(nu)], [np.full(nu, -4.0)]
    ii, jj = np.nonzero(interior)
    for di, dj in _OFFSETS:
        nb = index[ii + di, jj + dj]
        ok = nb >= 0
        rows.append(np.arange(nu)[ok])
        cols.append(nb[ok])
        vals.append(np.ones(int(ok.sum())))
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(nu, nu)).tocsc()
    lu = spla.splu(A)
    rhs = divergence(field.gx, field.gy)
    if rhs.ndim == 2:
        rhs = rhs[..., None]
    out = chans.copy()
    for k in range(chans.shape[2]):
        b = rhs[..., k][interior].astype(float)
        for di, dj in _OFFSETS:
            fixed = index[ii + di, jj + dj] < 0
            b[fixed] -= chans[ii + di, jj + dj, k][fixed]
        out[..., k][interior] = lu.solve(b)
    return (out[..., 0] if gray else out), []


def clone_window(source, target, mask, offset=(0, 0), mode="import", margin=24):
    """(field, window slices) of the rectangle route of seamless_clone."""
    target = np.asarray(target, dtype=float)
    source = np.asarray(source, dtype=float)
    row, col = offset
    sh, sw = source.shape[:2]
    canvas = np.zeros(target.shape)
    canvas[row:row + sh, col:col + sw] = source
    placed = np.zeros(target.shape[:2], dtype=bool)
    placed[row:row + sh, col:col + sw] = mask
    sf, tf = GuidanceField.of(canvas), GuidanceField.of(target)
    if mode == "import":
        inside = sf
    elif mode == "mixed":
        inside = sf.mix_by_magnitude(tf)
    elif mode == "average":
        inside = GuidanceField(0.5 * (sf.gx + tf.gx), 0.5 * (sf.gy + tf.gy))
    else:
        raise ValueError(mode)
    field = inside.select(placed, tf)
    rows, cols = np.nonzero(placed)
    top, bottom = max(int(rows.min()) - margin, 0), min(int(rows.max()) + margin + 1,
                                                        target.shape[0])
    left, right = max(int(cols.min()) - margin, 0), min(int(cols.max()) + margin + 1,
                                                        target.shape[1])
    return field, (slice(top, bottom), slice(left, right))


def placed_mask(source, target, mask, offset=(0, 0)):
    placed = np.zeros(np.asarray(target).shape[:2], dtype=bool)
    row, col = offset
    sh, sw = np.asarray(source).shape[:2]
    placed[row:row + sh, col:col + sw] = mask
    return placed


def seamless_clone(source, target, mask, offset=(0, 0), mode="import", method="direct",
                   margin=24, tolerance=1e-8, domain="mask"):
    target = np.asarray(target, dtype=float)
    field, win = clone_window(source, target, mask, offset, mode, margin)
    if domain == "mask":
        return solve_masked(field, target, placed_mask(source, target, mask, offset))
    patch, reports = solve_dirichlet(GuidanceField(field.gx[win], field.gy[win]), target[win],
                                     method=method, tolerance=tolerance)
    out = target.copy()
    out[win] = patch
    return out, reports


def texture_flatten(image, edges, method="mgcg", mask=None):
    """poisson.py:351-388: the image's field scaled by the edge weights (broadcast over
    channels), integrated with the image as boundary (or over ``mask``)."""
    image = np.asarray(image, dtype=float)
    weights = np.asarray(edges, dtype=float)
    if image.ndim == 3:
        weights = weights[..., None]
    field = GuidanceField.of(image).scaled(weights)
    if mask is not None:
        return solve_masked(field, image, mask)
    return solve_dirichlet(field, image, method=method)


def illumination_change(image, mask, alpha=0.2, beta=0.2, epsilon=1e-4):
    """poisson.py:391-437: gradient magnitudes remapped by alpha**beta * g**-beta (g floored
    at epsilon), integrated over ``mask`` with the other pixels fixed."""
    image = np.asarray(image, dtype=float)
    field = GuidanceField.of(image)
    magnitude = np.sqrt(field.gx ** 2 + field.gy ** 2)
    factor = (alpha ** beta) * np.maximum(magnitude, epsilon) ** (-beta)
    return solve_masked(field.scaled(factor), image, mask)
