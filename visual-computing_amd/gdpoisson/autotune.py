"""Tile-shape autotuner for batched multigrid plans.

Every tiled kernel of a plan comes from a compiled menu of tile shapes
(``gdp_tile_menu``): the down and up strip kernels of each tiled level (kinds 0 and 1,
rows per strip) and the Krylov kernels (kind 2).  The best entry depends on the level's
size, the number of problems and the sweep count, so a plan is tuned for what it
solves: a coordinate search that times every menu entry of one kernel at a time with
the plan's own launch profiler (HIP events around each launch) on a seeded synthetic
right-hand side, keeps the fastest, and moves to the next kernel.  Results are cached
per (shapes, boundary, options) for the life of the process.
"""

from __future__ import annotations

import re
from typing import Sequence

import torch

from .plan import Plan

_CACHE: dict = {}


def menu_size(kind: int) -> int:
    """Entries of a compiled tile menu (gdp_tile_menu answers EINVAL past the last)."""
    import ctypes

    from . import _lib

    th, tw = ctypes.c_int(), ctypes.c_int()
    n = 0
    while _lib.lib().gdp_tile_menu(kind, n, 1, ctypes.byref(th), ctypes.byref(tw)) == 0:
        n += 1
    return n


def _kernel_times(plan: Plan, b: Sequence[torch.Tensor], solves: int) -> dict[str, float]:
    """Mean time per launch of every kernel kind (ms); the time per solve of each kind
    under the key ("solve", kind)."""
    out = [torch.empty_like(x) for x in b]
    plan.mgcg(b, out)  # settle the new tables
    plan.profile(True)
    for _ in range(solves):
        plan.mgcg(b, out)
    rep = plan.profile_report()
    plan.profile(False)
    times: dict = {k: v["ms"] / max(v["launches"], 1) for k, v in rep.items()}
    times.update({("solve", k): v["ms"] / solves for k, v in rep.items()})
    return times


def _synthetic_rhs(plan: Plan, seed: int = 1234) -> list[torch.Tensor]:
    g = torch.Generator(device="cpu").manual_seed(seed)
    bs = []
    for (m, n), bc in zip(plan.shapes, plan.boundary):
        x = torch.randn(m, n, generator=g, dtype=torch.float64)
        if bc == "neumann":
            x -= x.mean()
        bs.append(x.to(plan.device))
    return bs


def _key(plan: Plan):
    o = plan.options
    return (tuple(plan.shapes), tuple(plan.boundary), o.precision, plan.smoothing,
            o.coarsest_size, o.coarsest_sweeps, o.flexible)


def autotune(plan: Plan, b: Sequence[torch.Tensor] | None = None, solves: int = 2,
             use_cache: bool = True) -> dict:
    """Pick the fastest tile shape of every tiled kernel of ``plan`` and bind it.

    Returns {(kind, level): menu index} (kind 0 down, 1 up, 2 Krylov with level 0).
    """
    key = _key(plan)
    if use_cache and key in _CACHE:
        plan.set_tiles(_CACHE[key])
        return dict(_CACHE[key])
    b = list(b) if b is not None else _synthetic_rhs(plan)
    base = _kernel_times(plan, b, solves)
    levels = sorted({int(mt.group(1)) for k in base if isinstance(k, str)
                     for mt in [re.match(r"down L(\d+)", k)] if mt})
    choice: dict = {}

    def cost(times: dict, kind: int, level: int) -> float:
        # a tile choice also costs the kernels that reduce its partial sums: every block
        # of the apply reduces the finest up kernel's <r,z>, <q,z> slots, and the first
        # down kernel's convergence test reduces the update's ||r||^2 slots (one slot
        # per block of the producer); the x-update blocks of the tail launch walk the
        # Krylov tiles
        if kind == 2:
            return sum(t for k, t in times.items()
                       if k in ("apply", "update", "update (x deferred)", "down L0", "tail"))
        if kind == 0 and level == 0:
            # the fused update-down entries replace the update of most iterations, move
            # its convergence test to the tail kernels and its r store to the finest up
            # kernel: compare the time per solve of all of these
            return sum(v for k, v in times.items()
                       if isinstance(k, tuple) and k[1] in (
                           "down L0", "down L0+update", "update", "update (x deferred)",
                           "tail", "up L0+dots"))
        pre = "down" if kind == 0 else "up"
        t = sum(v for k, v in times.items()
                if isinstance(k, str) and re.fullmatch(rf"{pre} L{level}(\+dots)?", k))
        if kind == 1 and level == 0:
            t += times.get("apply", 0.0)
        return t

    for kind, level in [(k, l) for l in levels for k in (0, 1)] + [(2, 0)]:
        best_idx, best_t = None, float("inf")
        for idx in range(menu_size(kind)):
            try:
                plan.set_tiles({(kind, level): idx})
            except ValueError:
                continue  # an entry this level cannot take (fused update: finest only)
            t = cost(_kernel_times(plan, b, solves), kind, level)
            if t < best_t:
                best_idx, best_t = idx, t
        choice[(kind, level)] = best_idx
        plan.set_tiles({(kind, level): best_idx})
    _CACHE[key] = dict(choice)
    return choice


def clear_cache() -> None:
    _CACHE.clear()
