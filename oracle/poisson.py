"""Gradient-domain compositing, rectangle route (FP64 numpy restatement; test
infrastructure only).

Follows ``gradient-domain/src/gradient_domain/poisson.py``:

* ``GuidanceField``  poisson.py:52-92    of / select / mix_by_magnitude / scaled
* ``solve_dirichlet`` poisson.py:102-147  divergence, fold the ring, solve per channel
* ``seamless_clone``  poisson.py:245-348  place the source crop, guidance by mode
                                          (import / mixed / average), window = mask
                                          bounding box + margin, solve_dirichlet there

The mask-domain route (SuperLU over the irregular selection, poisson.py:154-244) is not
restated: the GPU path does not offer it.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

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


def seamless_clone(source, target, mask, offset=(0, 0), mode="import", method="direct",
                   margin=24, tolerance=1e-8):
    target = np.asarray(target, dtype=float)
    field, win = clone_window(source, target, mask, offset, mode, margin)
    patch, reports = solve_dirichlet(GuidanceField(field.gx[win], field.gy[win]), target[win],
                                     method=method, tolerance=tolerance)
    out = target.copy()
    out[win] = patch
    return out, reports
