"""Serving loop for gradient-domain requests: continuous batching over HIP streams with
admission by HBM budget.

The reference runs one solve at a time on the caller's thread (solvers.py:171-221,
called per channel by poisson.solve_dirichlet, poisson.py:141-143, and once by
hdr.tone_map, hdr.py:175).  Here a request is a whole compositing job (seamless_clone,
rectangle route unless the request asks for ``domain="mask"``) or tone-mapping job
(tone_map), and a :class:`Server` replays a stream of timed arrivals:

* requests wait in arrival order; each holds an estimate of the device bytes it needs
  while admitted (inputs it copies, temporaries, outputs and its share of a plan's
  workspace, see :func:`request_bytes`);
* whenever one of ``slots`` workers is free, the oldest waiting request is admitted if
  its bytes fit what is left of the budget, together with every later waiting request of
  the same kind and image shape that still fits (up to ``max_batch`` problems);
* a worker runs its batch on its own HIP stream: each request's front-end kernels, one
  batched plan per plane shape holding every channel of every request of the batch with
  that shape (one launch per kernel for all of them), then each request's back-end
  kernels; the only host waits are the worker's own (its stream and each plan's single
  end-of-solve read);
* the plans stay resident: bound plans (level pyramid, Krylov vectors, tile tables and
  the hipGraphs of their PCG iterations) wait between batches in a pool shared by the
  workers, keyed by plane shape and a capacity of planes (the batch's plane count rounded
  up to 1, 2, 3, 4, 6, 8, 12, ...; the planes a batch leaves free hold a zero right-hand
  side and finish at set-up), within a byte cap, so a batch of a shape seen before reuses
  a workspace and replays its captured graphs instead of building and capturing anew;
* batches on different workers overlap on the device, and requests that arrive while
  others run are admitted as soon as a worker and budget free up.

Each problem of a batched plan is solved with the arithmetic it gets alone (tile shapes
chosen by its shape only, per-problem reductions in a fixed order, per-problem
stopping), so a request's result does not depend on what it was batched or overlapped
with: the same bits alone and in the mix.  Latency = completion - arrival, queueing
included.
"""

from __future__ import annotations

import dataclasses
import threading
import time
from collections import OrderedDict, deque
from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
from dataclasses import dataclass, field
from typing import Callable, Sequence

import numpy as np

KINDS = ("composite", "tonemap")
BOUNDARY_OF = {"composite": "dirichlet", "tonemap": "neumann"}


@dataclass(eq=False)
class Request:
    """``kind`` "composite": ``args`` = source, target, mask, offset, mode (and optionally
    margin); "tonemap": ``args`` = radiance (and optionally config).  Images are
    channels-last.  ``arrival`` is in seconds from the start of the stream.  Requests
    compare by identity (their arguments hold device tensors)."""

    kind: str
    args: dict
    arrival: float = 0.0
    rid: int = -1


@dataclass
class Completion:
    rid: int
    kind: str
    arrival: float
    admitted: float
    finished: float
    batch: int
    nbytes: int
    pixels: int
    output: object = None
    reports: list = field(default_factory=list)
    error: str | None = None

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def _shape(x) -> tuple:
    return tuple(int(s) for s in (x.shape if hasattr(x, "shape") else x))


def request_pixels(req: Request) -> int:
    a = req.args
    if "nbytes" in a and "pixels" in a:
        return int(a["pixels"])
    img = a["target"] if req.kind == "composite" else a["radiance"]
    s = _shape(img)
    return s[0] * s[1]


_WS_CACHE: dict = {}


def plan_capacity(planes: int) -> int:
    """Planes of a pooled plan for a batch of ``planes``: the next of 1, 2, 3, 4, 6, 8,
    12, 16, 24, ... (powers of two and three times powers of two), so batches of nearby
    sizes share one resident plan and at most a third of its planes idle."""
    c = 1
    while True:
        for v in (c, 3 * c // 2):
            if v >= planes:
                return max(v, 1)
        c *= 2


def _plan_bytes_per_point(boundary: str, precision: str) -> float:
    """Workspace bytes per unknown of a one-problem plan (measured on a 512 x 512 plan:
    the workspace is the level pyramid plus Krylov vectors, linear in the points)."""
    key = (boundary, precision)
    if key not in _WS_CACHE:
        from .plan import plan_workspace_bytes

        b = plan_workspace_bytes([(512, 512)], boundary, precision)
        _WS_CACHE[key] = b / (512 * 512)
    return _WS_CACHE[key]


def request_bytes(req: Request, precision: str = "mixed") -> int:
    """Upper bound of the device bytes a request holds while admitted.

    composite: the target's planes (a copy) and the output image, the source's planes and
    mask, the right-hand side and solution of the window (bounded by the target), the plan
    workspace of its channels and, by the mask route, its byte masks; tonemap: log luminance, pyramid (4/3 of a plane),
    factor, right-hand side, solution, the output image and the plan workspace.  An
    ``nbytes`` entry in ``args`` overrides the estimate.
    """
    a = req.args
    if "nbytes" in a:
        return int(a["nbytes"])
    per = 1.1 * _plan_bytes_per_point(BOUNDARY_OF[req.kind], precision)
    if req.kind == "composite":
        h, w = _shape(a["target"])[:2]
        c = _shape(a["target"])[2] if len(_shape(a["target"])) == 3 else 1
        sh, sw = _shape(a["source"])[:2]
        # the mask route's placed, interior and window masks
        masks = 3 * h * w if a.get("domain", "rectangle") == "mask" else 0
        return int(8 * c * (2 * h * w + sh * sw) + sh * sw + 16 * c * h * w + per * c * h * w
                   + masks)
    h, w = _shape(a["radiance"])[:2]
    return int(8 * h * w * (3 + 1 + 4 / 3 + 1 + 1 + 1) + per * h * w + (1 << 20))


class _Pooled:
    """A bound plan of ``capacity`` planes of one shape (key: its pool key, None for a
    plan built for one batch).  ``solve`` runs MG-PCG on up to ``capacity`` planes: the
    rest get a zero right-hand side, which the solver finishes at set-up (their outputs,
    zeros, go to one shared scratch plane that nothing reads)."""

    def __init__(self, key, plan, capacity: int):
        self.key, self.plan, self.capacity = key, plan, capacity
        self._pad_b = self._pad_u = None

    def solve(self, planes: list):
        import torch

        n = len(planes)
        if n > self.capacity:
            raise ValueError(f"{n} planes for a plan of {self.capacity}")
        if n == self.capacity:
            return self.plan.mgcg(planes)
        if self._pad_b is None:
            self._pad_b = torch.zeros_like(planes[0])
            self._pad_u = torch.empty_like(planes[0])
        outs = [torch.empty_like(p) for p in planes] + [self._pad_u] * (self.capacity - n)
        us, rs = self.plan.mgcg(list(planes) + [self._pad_b] * (self.capacity - n), outs)
        return us[:n], rs[:n]


class Server:
    """Continuous-batching request server on one GPU (see the module docstring).

    ``budget_bytes``: device bytes the server may hold (default: half of the free HBM);
    ``plan_pool_bytes`` of it (default an eighth) is kept for the resident plans, the rest
    is the admission budget of the requests.  ``slots``: workers, one HIP stream each.
    ``max_batch``: problems (channels) per batched plan.  The resident plans of all
    workers form one pool (a plan is used by one worker at a time).  ``method``: "mgcg_mixed" or
    "mgcg".  ``tonemap_method``: "dct" integrates each tone map exactly by the device DCT,
    as the reference's tone_map does (sides whose prime factors are not all 2, 3, 5 join
    the batched plan), "mg" puts every tone map in the batched MG-PCG plan.  ``graphs``:
    capture the pooled plans' PCG iterations in hipGraphs (a pooled plan replays them on
    every later use; mask-route plans, whose masks change per request, are built per
    batch and run without).  ``tiles``: {(boundary, (m, n)): {(kind, level): menu index}},
    the tile shapes of planes of that shape (see :meth:`tune`); other shapes get the
    default tiles.  ``executor`` replaces the device pipeline (tests of the scheduling):
    called as ``executor(batch)`` in a worker, it returns one ``(output, reports, error)``
    per request.
    """

    def __init__(self, budget_bytes: int | None = None, slots: int = 4, max_batch: int = 12,
                 method: str = "mgcg_mixed", tolerance: float = 1e-8,
                 max_iterations: int = 2000, graphs: bool = True, keep_outputs: bool = True,
                 device=None, executor: Callable | None = None, tonemap_method: str = "dct",
                 plan_pool_bytes: int | None = None, tiles: dict | None = None):
        if method not in ("mgcg", "mgcg_mixed"):
            raise ValueError("method must be 'mgcg' or 'mgcg_mixed'")
        if tonemap_method not in ("dct", "mg"):
            raise ValueError("tonemap_method must be 'dct' or 'mg'")
        self.tonemap_method = tonemap_method
        if slots < 1 or max_batch < 1:
            raise ValueError("slots and max_batch must be positive")
        self.slots, self.max_batch = int(slots), int(max_batch)
        self.method, self.tolerance, self.max_iterations = method, float(tolerance), int(max_iterations)
        self.precision = "mixed" if method == "mgcg_mixed" else "fp64"
        self.graphs, self.keep_outputs = bool(graphs), bool(keep_outputs)
        self.executor = executor
        self.device = None
        if executor is None:
            import torch

            self.device = torch.device("cuda", torch.cuda.current_device() if device is None
                                       else int(device))
            if budget_bytes is None:
                free, _ = torch.cuda.mem_get_info(self.device)
                budget_bytes = free // 2
        if budget_bytes is None:
            raise ValueError("budget_bytes is required with a custom executor")
        if plan_pool_bytes is None:
            plan_pool_bytes = 0 if executor is not None else int(budget_bytes) // 8
        if not 0 <= plan_pool_bytes < int(budget_bytes):
            raise ValueError("plan_pool_bytes must be smaller than the budget")
        self.total_budget = int(budget_bytes)
        self.pool_bytes = int(plan_pool_bytes)
        self.budget = self.total_budget - self.pool_bytes  # admission budget of requests
        self.tiles = dict(tiles or {})
        self._local = threading.local()
        self._lock = threading.Lock()
        self.peak_bytes = 0
        self.batches: list[tuple[str, int, int]] = []  # (kind, requests, problems)
        self.plan_hits = 0     # batches served by a resident plan
        self.plan_builds = 0   # plans built (and, when pooled, kept)
        self.unpooled_builds = 0  # of which built for one batch (mask route, no pool)
        self.pool_peak = 0     # largest bytes of resident plans (in use or waiting)
        # resident plans waiting for a batch: key -> [_Pooled, ...] (least recently
        # returned first), and the bytes of every pooled plan, in use or waiting
        self._pool: OrderedDict = OrderedDict()
        self._pool_used = 0
        self._workers = None

    def close(self) -> None:
        """Stop the workers; their streams and the resident plans are released."""
        if self._workers is not None:
            self._workers.shutdown(wait=True)
            self._workers = None
        with self._lock:
            self._pool.clear()
            self._pool_used = 0

    # ------------------------------------------------------------ scheduling
    def _bytes(self, req: Request) -> int:
        return request_bytes(req, self.precision)

    def _problems(self, req: Request) -> int:
        if req.kind == "tonemap":
            return 1
        t = req.args.get("target")
        s = _shape(t) if t is not None else (1, 1, 3)
        return s[2] if len(s) == 3 else 1

    @staticmethod
    def _signature(req: Request) -> tuple:
        """Requests batch together only with the same kind and image shape (and route), so
        a batch's planes repeat shapes that the resident plans were built for."""
        a = req.args
        if req.kind == "composite" and "target" in a:
            return ("composite", _shape(a["target"]), a.get("domain", "rectangle"))
        if req.kind == "tonemap" and "radiance" in a:
            return ("tonemap", _shape(a["radiance"]))
        return (req.kind,)

    def _form_batch(self, waiting: list, free: int):
        head = waiting[0]
        need = self._bytes(head)
        if need > free:
            return [], 0
        sig = self._signature(head)
        batch, total, probs = [head], need, self._problems(head)
        for r in waiting[1:]:
            if self._signature(r) != sig:
                continue
            p, b = self._problems(r), self._bytes(r)
            if probs + p > self.max_batch:
                break
            if total + b > free:
                continue
            batch.append(r)
            total += b
            probs += p
        return batch, total

    def run(self, requests: Sequence[Request], realtime: bool = True) -> list[Completion]:
        """Serve ``requests`` (arrival times honoured when ``realtime``, else all arrive at
        once) and return their completions in request order."""
        for i, r in enumerate(requests):
            if r.kind not in KINDS:
                raise ValueError(f"unknown request kind {r.kind!r}")
            if r.rid < 0:
                r.rid = i
        if self.executor is None:
            import torch

            torch.cuda.synchronize(self.device)  # inputs made on other streams are ready
        pending = deque(sorted(requests, key=lambda r: (r.arrival if realtime else 0.0, r.rid)))
        waiting: list[Request] = []
        running: dict = {}
        done: list[Completion] = []
        in_use = 0
        bid = 0
        t0 = time.perf_counter()
        clock = lambda: time.perf_counter() - t0  # noqa: E731
        # the workers (and with them their streams and resident plans) live as long as the
        # server, so later runs reuse what earlier ones built
        if self._workers is None:
            self._workers = ThreadPoolExecutor(max_workers=self.slots,
                                               thread_name_prefix="gdp-slot")
        pool = self._workers
        while pending or waiting or running:
            now = clock()
            while pending and (not realtime or pending[0].arrival <= now):
                waiting.append(pending.popleft())
            while waiting and self._bytes(waiting[0]) > self.budget:
                r = waiting.pop(0)
                done.append(Completion(r.rid, r.kind, r.arrival if realtime else 0.0, now, now,
                                       -1, self._bytes(r), request_pixels(r),
                                       error="needs more device memory than the budget"))
            while waiting and len(running) < self.slots:
                batch, nbytes = self._form_batch(waiting, self.budget - in_use)
                if not batch:
                    break
                for r in batch:
                    waiting.remove(r)
                in_use += nbytes
                self.peak_bytes = max(self.peak_bytes, in_use)
                self.batches.append((batch[0].kind, len(batch),
                                     sum(self._problems(r) for r in batch)))
                fut = pool.submit(self._execute, batch, bid, clock(), clock, realtime)
                running[fut] = nbytes
                bid += 1
            timeout = None
            if pending and realtime:
                timeout = max(0.0, pending[0].arrival - clock())
            if running:
                finished, _ = wait(list(running), timeout=timeout, return_when=FIRST_COMPLETED)
                for f in finished:
                    in_use -= running.pop(f)
                    done.extend(f.result())
            elif timeout:
                time.sleep(timeout)
        return sorted(done, key=lambda c: c.rid)

    # ------------------------------------------------------------ execution
    def _execute(self, batch: list, bid: int, admitted: float, clock, realtime: bool):
        # a failure of the batch as a whole (an executor that raises, a plan that cannot be
        # built or bound) fails that batch's requests only: the loop keeps serving and the
        # batch's budget is still handed back
        try:
            if self.executor is not None:
                results = list(self.executor(batch))
            else:
                results = self._device_batch(batch)
            if len(results) != len(batch):
                raise RuntimeError(f"executor returned {len(results)} results for "
                                   f"{len(batch)} requests")
        except Exception as e:  # noqa: BLE001 -- reported per request
            msg = f"{type(e).__name__}: {e}"
            results = [(None, [], msg)] * len(batch)
        fin = clock()
        out = []
        for r, (output, reports, err) in zip(batch, results):
            out.append(Completion(r.rid, r.kind, r.arrival if realtime else 0.0, admitted, fin,
                                  bid, self._bytes(r), request_pixels(r),
                                  output if self.keep_outputs else None, reports, err))
        return out

    def _stream(self):
        import torch

        s = getattr(self._local, "stream", None)
        if s is None:
            s = torch.cuda.Stream(self.device)  # default priority
            self._local.stream = s
        return s

    def _plan_options(self, boundary: str, graphs: bool):
        from .solvers import _pcg_options

        return dataclasses.replace(_pcg_options(self.precision, self.tolerance,
                                                self.max_iterations, boundary),
                                   use_graph=graphs)

    def _tiles_of(self, boundary: str, shapes: list):
        """Tile choices of a plan whose planes all have one shape listed in ``tiles``."""
        if shapes and all(sh == shapes[0] for sh in shapes):
            return self.tiles.get((boundary, tuple(shapes[0])))
        return None

    def _plan(self, shapes: list, boundary: str, masks):
        """A bound plan for these planes, as a _Pooled (its ``solve`` pads the planes to
        the plan's capacity): a resident one of the same plane shape and capacity when one
        is waiting in the pool (its workspace and captured graphs are reused), else built.
        Hand it back with :meth:`_release`.  Mask-route plans are built per batch."""
        from .plan import Plan

        tiles = self._tiles_of(boundary, shapes)
        if masks is not None or self.pool_bytes == 0:
            with self._lock:
                self.plan_builds += 1
                self.unpooled_builds += 1
            plan = Plan(shapes, boundary, 1.0, self._plan_options(boundary, False),
                        device=self.device, tiles=tiles, masks=masks)
            return _Pooled(None, plan, len(shapes))
        shape = tuple(shapes[0])
        if any(tuple(sh) != shape for sh in shapes):
            raise ValueError("a pooled plan holds planes of one shape")
        cap = plan_capacity(len(shapes))
        key = (boundary, shape, cap)
        with self._lock:
            waiting = self._pool.get(key)
            if waiting:
                pooled = waiting.pop()
                if not waiting:
                    del self._pool[key]
                self.plan_hits += 1
                return pooled
        plan = Plan([shape] * cap, boundary, 1.0, self._plan_options(boundary, self.graphs),
                    device=self.device, tiles=tiles)
        with self._lock:
            self.plan_builds += 1
            self._pool_used += plan.workspace_bytes
            self.pool_peak = max(self.pool_peak, self._pool_used)
        return _Pooled(key, plan, cap)

    def _release(self, pooled: "_Pooled") -> None:
        """Return a pooled plan after its batch (its last solve has finished: the solve is
        synchronous); plans waiting longest are dropped past ``plan_pool_bytes``."""
        if pooled.key is None:
            return
        with self._lock:
            self._pool.setdefault(pooled.key, []).insert(0, pooled)
            self._pool.move_to_end(pooled.key)
            while self._pool_used > self.pool_bytes and self._pool:
                key, waiting = next(iter(self._pool.items()))
                old = waiting.pop()  # the longest waiting of the least recent key
                if not waiting:
                    del self._pool[key]
                self._pool_used -= old.plan.workspace_bytes
                if old is pooled:
                    break

    def tune(self, requests: Sequence[Request], solves: int = 2) -> dict:
        """Autotune the tile shapes of every plane shape these requests solve by MG-PCG (a
        plan of one request's planes per shape, timed by :func:`gdpoisson.autotune`), add
        them to ``tiles`` and return the table.  The choice depends on the shape only, so
        a request keeps its bits alone and batched."""
        import torch

        from .autotune import autotune
        from .hdr import prepare_tone_map
        from .plan import Plan
        from .poisson import prepare_clone
        from .solvers import dct_supported

        seen = {}
        with torch.cuda.device(self.device):
            for r in requests:
                a = r.args
                if r.kind == "composite":
                    if a.get("domain", "rectangle") == "mask":
                        continue
                    j = prepare_clone(a["source"], a["target"], a["mask"],
                                      a.get("offset", (0, 0)), a.get("mode", "import"),
                                      "rectangle", margin=int(a.get("margin", 24)))
                else:
                    j = prepare_tone_map(a["radiance"], a.get("config"))
                    if self.tonemap_method == "dct" and dct_supported(*j.rhs.shape[1:]):
                        continue
                if j.rhs is None:
                    continue
                key = (BOUNDARY_OF[r.kind], tuple(int(x) for x in j.rhs.shape[1:]))
                if key not in self.tiles and key not in seen:
                    seen[key] = int(j.rhs.shape[0])
            for (bc, shape), count in seen.items():
                plan = Plan([shape] * count, bc, 1.0, self._plan_options(bc, False),
                            device=self.device)
                self.tiles[(bc, shape)] = autotune(plan, solves=solves)
        return dict(self.tiles)

    def stats(self) -> dict:
        """Batching and plan-residency counters of the requests served so far."""
        nb = len(self.batches)
        return {
            "batches": nb,
            "mean_requests_per_batch": round(sum(n for _, n, _ in self.batches) / max(nb, 1), 2),
            "max_requests_per_batch": max((n for _, n, _ in self.batches), default=0),
            "plan_hits": self.plan_hits,
            "plan_builds": self.plan_builds,
            "unpooled_builds": self.unpooled_builds,
            "graphs": self.graphs,
            "tuned_shapes": len(self.tiles),
            "resident_plan_peak_GB": round(self.pool_peak / 1e9, 3),
        }

    def _device_batch(self, batch: list):
        import torch

        from .hdr import finish_tone_map, prepare_tone_map
        from .poisson import finish_clone, prepare_clone
        from .solvers import SolverReport, dct_supported, relative_residual, solve_neumann

        stream = self._stream()
        results: list = [None] * len(batch)
        with torch.cuda.device(self.device), torch.cuda.stream(stream):
            jobs: list = [None] * len(batch)
            for i, r in enumerate(batch):
                try:
                    a = r.args
                    if r.kind == "composite":
                        jobs[i] = prepare_clone(a["source"], a["target"], a["mask"],
                                                a.get("offset", (0, 0)), a.get("mode", "import"),
                                                a.get("domain", "rectangle"),
                                                margin=int(a.get("margin", 24)))
                    else:
                        jobs[i] = prepare_tone_map(a["radiance"], a.get("config"))
                except (ValueError, NotImplementedError) as e:
                    results[i] = (None, [], str(e))
            if self.tonemap_method == "dct":
                # exact integration of each tone map on this stream (hdr.py:175)
                for i, (r, j) in enumerate(zip(batch, jobs)):
                    if r.kind != "tonemap" or j is None or not dct_supported(*j.rhs.shape[1:]):
                        continue
                    t0 = time.perf_counter()
                    u = solve_neumann(j.rhs[0], method="dct")
                    rr = relative_residual(u, j.rhs[0], 1.0, "neumann")
                    img, _ = finish_tone_map(j, u)
                    results[i] = (img, [SolverReport(
                        method="dct", iterations=1, relative_residual=rr,
                        seconds=time.perf_counter() - t0, converged=True)], None)
                    jobs[i] = None
            planes, owner, masks = [], [], []
            for i, j in enumerate(jobs):
                if j is not None and j.rhs is not None:
                    for k in range(j.rhs.shape[0]):
                        planes.append(j.rhs[k])
                        owner.append(i)
                        # the unknowns of a mask-route composite (None: every point)
                        masks.append(getattr(j, "unknown", None))
            sols: dict = {}
            reps: dict = {}
            # one plan per plane shape: a request's planes share its window's shape, and a
            # plan of one shape takes that shape's tiles (alone or batched alike) and is
            # the one the pool keeps for it
            groups: dict = {}
            for k, pl in enumerate(planes):
                groups.setdefault(tuple(pl.shape), []).append(k)
            for idx in groups.values():
                t_start = time.perf_counter()
                gm = [masks[k] for k in idx]
                go = [owner[k] for k in idx]
                try:
                    pooled = self._plan([tuple(planes[k].shape) for k in idx],
                                        BOUNDARY_OF[batch[0].kind],
                                        gm if any(m is not None for m in gm) else None)
                    try:
                        us, rs = pooled.solve([planes[k] for k in idx])
                    finally:
                        self._release(pooled)
                except Exception as e:  # noqa: BLE001 -- the plan's requests fail alone
                    msg = f"{type(e).__name__}: {e}"
                    for i in set(go):
                        results[i] = (None, [], msg)
                        jobs[i] = None
                    continue
                dt = time.perf_counter() - t_start
                for i, u, rep in zip(go, us, rs):
                    sols.setdefault(i, []).append(u)
                    reps.setdefault(i, []).append(SolverReport(
                        method=self.method, iterations=int(rep.iterations),
                        relative_residual=float(rep.relative_residual), seconds=dt,
                        converged=bool(rep.converged)))
            for i, (r, j) in enumerate(zip(batch, jobs)):
                if j is None:
                    continue
                if r.kind == "composite":
                    sol = torch.stack(sols[i]) if i in sols else None
                    results[i] = (finish_clone(j, sol), reps.get(i, []), None)
                else:
                    img, _ = finish_tone_map(j, sols[i][0])
                    results[i] = (img, reps.get(i, []), None)
            stream.synchronize()
        return results


def latency_stats(completions: Sequence[Completion]) -> dict:
    """p50 / p99 / max latency (ms) of the served requests, throughput and sizes."""
    ok = [c for c in completions if c.error is None]
    if not ok:
        return {"served": 0, "rejected": len(completions)}
    lat = np.array([c.latency for c in ok]) * 1e3
    span = max(c.finished for c in ok) - min(c.arrival for c in ok)
    return {
        "served": len(ok),
        "rejected": len(completions) - len(ok),
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "max_ms": float(lat.max()),
        "mean_ms": float(lat.mean()),
        "requests_per_s": len(ok) / span if span > 0 else float("nan"),
        "mpix_per_s": sum(c.pixels for c in ok) / 1e6 / span if span > 0 else float("nan"),
    }


def poisson_arrivals(count: int, qps: float, seed: int) -> np.ndarray:
    """Arrival times (s) of a Poisson process of rate ``qps``."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / qps, count))
