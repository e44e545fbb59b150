"""Batched multigrid plans: a set of independent grids solved by shared launches.

A :class:`Plan` owns one libgdpoisson plan and its device workspace (a torch
``uint8`` tensor).  Every kernel launch of a solve covers all of the plan's
problems, so the three channels of an RGB image, or a batch of images of different
sizes, cost the launches of one solve.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Sequence

import torch

from . import _lib

BOUNDARY = {"dirichlet": _lib.DIRICHLET, "neumann": _lib.NEUMANN}
PRECISION = {"fp64": _lib.FP64, "mixed": _lib.MIXED}


@dataclass(frozen=True)
class PlanOptions:
    """Solver knobs (``MultigridOptions`` of multigrid.py:46-71 plus precision etc.)."""

    precision: str = "fp64"
    # default: 1/1 for FP64 (the reference mgcg cycle, solvers.py:137); 2/2 for mixed,
    # where the extra sweeps run in LDS and cut the PCG iterations from 9 to 6
    pre_smoothing: int | None = None
    post_smoothing: int | None = None
    # defaults: Neumann coarsens to 16 cells with the exact DCT coarsest solve; Dirichlet
    # FP64 keeps the reference (coarsest <= 8, 60 sweeps, multigrid.py:66-69); Dirichlet
    # mixed coarsens to 16 with the exact DST coarsest solve (coarsest_sweeps = 0)
    coarsest_size: int | None = None
    coarsest_sweeps: int | None = None
    max_iterations: int = 2000
    tolerance: float = 1e-8
    flexible: bool | None = None  # default: Polak-Ribiere for mixed, Fletcher-Reeves for fp64
    use_graph: bool = True


def _device(device) -> torch.device:
    if device is None:
        if not torch.cuda.is_available():
            raise RuntimeError("gdpoisson needs a HIP GPU (torch.cuda.is_available() is False)")
        return torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)


def plan_workspace_bytes(shapes, boundary: str = "dirichlet", precision: str = "fp64") -> int:
    """Device workspace bytes of a plan with the default options of ``precision`` for
    these shapes; takes no device memory."""
    p = Plan(shapes, boundary, 1.0, PlanOptions(precision=precision), allocate=False)
    return p.workspace_bytes


class Plan:
    """A bound libgdpoisson plan for ``shapes`` (one (m, n) per problem)."""

    def __init__(
        self,
        shapes: Sequence[tuple[int, int]],
        boundary: str | Sequence[str] = "dirichlet",
        spacing: float | Sequence[float] = 1.0,
        options: PlanOptions | None = None,
        device=None,
        tiles: dict | None = None,
        allocate: bool = True,
        masks: Sequence | None = None,
    ):
        opts = options or PlanOptions()
        self.options = opts
        self.shapes = [tuple(int(x) for x in s) for s in shapes]
        n = len(self.shapes)
        bcs = [boundary] * n if isinstance(boundary, str) else list(boundary)
        hs = [float(spacing)] * n if isinstance(spacing, (int, float)) else list(spacing)
        if len(bcs) != n or len(hs) != n:
            raise ValueError("boundary / spacing lists must match the number of shapes")
        for b in bcs:
            if b not in BOUNDARY:
                raise ValueError(f"unknown boundary {b!r}")
        if opts.precision not in PRECISION:
            raise ValueError(f"unknown precision {opts.precision!r}")
        self.boundary = bcs
        self.device = _device(device)
        probs = (_lib.Problem * n)()
        masked = [False] * n if masks is None else [mk is not None for mk in masks]
        if len(masked) != n:
            raise ValueError("masks must match the number of shapes")
        for i, ((m, nn), b, h) in enumerate(zip(self.shapes, bcs, hs)):
            # a masked problem keeps every level but the coarsest out of the coarse tail,
            # so they are smoothed with its mask
            probs[i] = _lib.Problem(m, nn, BOUNDARY[b], 1 if masked[i] else 0, h)
        mixed = opts.precision == "mixed"
        cs = opts.coarsest_size
        if cs is None:
            cs = 16 if (bcs[0] == "neumann" or mixed) else 8
        sweeps = opts.coarsest_sweeps
        if sweeps is None:
            sweeps = 0 if (mixed and bcs[0] == "dirichlet") else 60
        flex = opts.flexible
        if flex is None:
            flex = opts.precision == "mixed"
        nu_default = 2 if opts.precision == "mixed" else 1
        pre = nu_default if opts.pre_smoothing is None else opts.pre_smoothing
        post = nu_default if opts.post_smoothing is None else opts.post_smoothing
        self.smoothing = (pre, post)
        copt = _lib.Options(
            PRECISION[opts.precision],
            pre,
            post,
            cs,
            sweeps,
            opts.max_iterations,
            int(bool(flex)),
            int(bool(opts.use_graph)),
            opts.tolerance,
        )
        L = _lib.lib()
        self._plan = ctypes.c_void_p()
        _lib.check(L.gdp_plan_create(probs, n, ctypes.byref(copt), ctypes.byref(self._plan)),
                   "gdp_plan_create")
        self.max_iterations = opts.max_iterations
        if tiles:
            for (kind, level), idx in tiles.items():
                _lib.check(L.gdp_plan_set_tile(self._plan, kind, level, idx), "gdp_plan_set_tile")
        # masked Dirichlet domains: one (m, n) uint8 device tensor per problem (or None),
        # kept alive with the plan
        self.masks = list(masks) if masks is not None else [None] * n
        if len(self.masks) != n:
            raise ValueError("masks must match the number of shapes")
        for i, mk in enumerate(self.masks):
            if mk is None:
                continue
            if (tuple(mk.shape) != self.shapes[i] or mk.dtype != torch.uint8
                    or not mk.is_contiguous() or mk.device.type != "cuda"):
                raise ValueError(f"mask {i}: need a contiguous uint8 device tensor of shape "
                                 f"{self.shapes[i]}")
            _lib.check(L.gdp_plan_set_mask(self._plan, i, ctypes.c_void_p(mk.data_ptr())),
                       "gdp_plan_set_mask")
        nbytes = ctypes.c_size_t()
        _lib.check(L.gdp_plan_workspace_bytes(self._plan, ctypes.byref(nbytes)))
        self.workspace_bytes = int(nbytes.value)
        if not allocate:  # sizing only: no device memory, not usable for solves
            self.workspace = None
            return
        with torch.cuda.device(self.device):
            self.workspace = torch.empty(self.workspace_bytes + 256, dtype=torch.uint8,
                                         device=self.device)
        self._bind()

    def _bind(self):
        base = self.workspace.data_ptr()
        aligned = (base + 255) & ~255
        with torch.cuda.device(self.device):
            _lib.check(
                _lib.lib().gdp_plan_bind(self._plan, ctypes.c_void_p(aligned),
                                         self.workspace_bytes),
                "gdp_plan_bind",
            )

    def set_tiles(self, tiles: dict):
        """Re-bind with {(kind, level): menu index} tile choices (see gdp_tile_menu)."""
        L = _lib.lib()
        for (kind, level), idx in tiles.items():
            _lib.check(L.gdp_plan_set_tile(self._plan, kind, level, idx), "gdp_plan_set_tile")
        self._bind()

    def __del__(self):
        p = getattr(self, "_plan", None)
        if p is not None and p.value:
            try:
                _lib.lib().gdp_plan_destroy(p)
            except Exception:
                pass
            self._plan = ctypes.c_void_p()

    # -------------------------------------------------------------- helpers
    @property
    def count(self) -> int:
        return len(self.shapes)

    def levels(self, problem: int = 0) -> list[tuple[int, int]]:
        L = _lib.lib()
        nl = ctypes.c_int()
        _lib.check(L.gdp_plan_levels(self._plan, problem, ctypes.byref(nl)))
        out = []
        for lv in range(nl.value):
            m, n = ctypes.c_int(), ctypes.c_int()
            _lib.check(L.gdp_plan_level_shape(self._plan, problem, lv, ctypes.byref(m),
                                              ctypes.byref(n)))
            out.append((m.value, n.value))
        return out

    def _check_inputs(self, arrays, name):
        if len(arrays) != self.count:
            raise ValueError(f"{name}: expected {self.count} arrays, got {len(arrays)}")
        for a, s in zip(arrays, self.shapes):
            if tuple(a.shape) != s:
                raise ValueError(f"{name}: shape {tuple(a.shape)} does not match {s}")
            if a.dtype != torch.float64 or not a.is_contiguous() or a.device != self.device:
                raise ValueError(f"{name}: need contiguous float64 tensors on {self.device}")

    def _ptrs(self, arrays):
        """ctypes array of device pointers, reused when the same tensors come back."""
        key = tuple(x.data_ptr() for x in arrays)
        cache = self.__dict__.setdefault("_ptr_cache", {})
        arr = cache.get(key)
        if arr is None:
            if len(cache) > 64:
                cache.clear()
            arr = cache[key] = _lib.ptr_array(list(key))
        return arr

    # -------------------------------------------------------------- solves
    def mgcg(self, b: Sequence[torch.Tensor], out: Sequence[torch.Tensor] | None = None,
             jacobi: bool = False):
        """MG-preconditioned CG (Jacobi-CG if ``jacobi``) of every problem.

        Returns (solutions, list of _lib.Report).
        """
        b = list(b)
        self._check_inputs(b, "b")
        if out is None:
            out = [torch.empty_like(x) for x in b]
        out = list(out)
        self._check_inputs(out, "out")
        reps = (_lib.Report * self.count)()
        bp = self._ptrs(b)
        up = self._ptrs(out)
        fn = _lib.lib().gdp_cg_solve if jacobi else _lib.lib().gdp_mgcg_solve
        with torch.cuda.device(self.device):
            _lib.check(fn(self._plan, bp, up, reps, _lib.current_stream()), "solve")
        return out, list(reps)

    def vcycle(self, u: Sequence[torch.Tensor], b: Sequence[torch.Tensor],
               out: Sequence[torch.Tensor] | None = None):
        u, b = list(u), list(b)
        self._check_inputs(u, "u")
        self._check_inputs(b, "b")
        if out is None:
            out = [torch.empty_like(x) for x in b]
        self._check_inputs(list(out), "out")
        with torch.cuda.device(self.device):
            _lib.check(
                _lib.lib().gdp_vcycle(
                    self._plan,
                    _lib.ptr_array([x.data_ptr() for x in u]),
                    _lib.ptr_array([x.data_ptr() for x in b]),
                    _lib.ptr_array([x.data_ptr() for x in out]),
                    _lib.current_stream(),
                ),
                "gdp_vcycle",
            )
        return list(out)

    def mg_solve(self, b: Sequence[torch.Tensor], initial: Sequence[torch.Tensor] | None = None):
        """Repeated V-cycles (multigrid.solve).  Returns (solutions, reports, histories)."""
        b = list(b)
        self._check_inputs(b, "b")
        out = [torch.empty_like(x) for x in b]
        ip = None
        if initial is not None:
            initial = list(initial)
            self._check_inputs(initial, "initial")
            ip = _lib.ptr_array([x.data_ptr() for x in initial])
        reps = (_lib.Report * self.count)()
        hist = (ctypes.c_double * (self.count * (self.max_iterations + 1)))()
        with torch.cuda.device(self.device):
            _lib.check(
                _lib.lib().gdp_mg_solve(
                    self._plan,
                    _lib.ptr_array([x.data_ptr() for x in b]),
                    ip,
                    _lib.ptr_array([x.data_ptr() for x in out]),
                    reps,
                    hist,
                    _lib.current_stream(),
                ),
                "gdp_mg_solve",
            )
        histories = []
        stride = self.max_iterations + 1
        for p in range(self.count):
            k = reps[p].iterations
            histories.append([hist[p * stride + i] for i in range(k + 1)])
        return out, list(reps), histories

    # -------------------------------------------------------------- kernel timing
    def profile(self, enable: bool = True) -> None:
        """Bracket every launch with HIP events on its stream (graphs are bypassed)."""
        _lib.check(_lib.lib().gdp_plan_profile(self._plan, int(bool(enable))), "profile")

    def profile_report(self) -> dict:
        """{kernel kind: {"ms", "launches", "bytes"}} aggregated since profile(True)."""
        L = _lib.lib()
        n = ctypes.c_int()
        with torch.cuda.device(self.device):
            _lib.check(L.gdp_plan_profile_count(self._plan, ctypes.byref(n)), "profile_count")
        out = {}
        buf = ctypes.create_string_buffer(64)
        ms, launches, nbytes = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_double()
        for i in range(n.value):
            _lib.check(L.gdp_plan_profile_entry(self._plan, i, buf, 64, ctypes.byref(ms),
                                                ctypes.byref(launches), ctypes.byref(nbytes)))
            out[buf.value.decode()] = {"ms": ms.value, "launches": int(launches.value),
                                       "bytes": nbytes.value}
        return out
