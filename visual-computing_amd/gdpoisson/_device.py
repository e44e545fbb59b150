"""numpy <-> device tensor plumbing and the plan cache."""

from __future__ import annotations

import threading
from collections import OrderedDict

import numpy as np
import torch

from .plan import Plan, PlanOptions


def to_device(x, device=None) -> tuple[torch.Tensor, bool]:
    """Return (contiguous float64 device tensor, input_was_host)."""
    if isinstance(x, torch.Tensor):
        if x.device.type == "cuda":
            t = x.to(torch.float64)
            return (t if t.is_contiguous() else t.contiguous()), False
        return x.to(torch.float64).contiguous().cuda(device), True
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    return torch.from_numpy(a).cuda(device), True


def to_host_like(t: torch.Tensor, was_host: bool):
    return t.cpu().numpy() if was_host else t


_CACHE: "OrderedDict[tuple, Plan]" = OrderedDict()
_CACHE_MAX = 24
_CACHE_LOCK = threading.Lock()


def cached_plan(shapes, boundary, spacing, options: PlanOptions) -> Plan:
    """A plan for this configuration, kept for reuse.  Plans are kept per host thread: a
    plan's workspace and host state serve one solve at a time."""
    dev = torch.cuda.current_device()
    key = (tuple(tuple(s) for s in shapes), boundary if isinstance(boundary, str) else tuple(boundary),
           spacing if isinstance(spacing, (int, float)) else tuple(spacing), options, dev,
           threading.get_ident())
    with _CACHE_LOCK:
        p = _CACHE.get(key)
        if p is not None:
            _CACHE.move_to_end(key)
            return p
    p = Plan(shapes, boundary, spacing, options)
    with _CACHE_LOCK:
        _CACHE[key] = p
        while len(_CACHE) > _CACHE_MAX:
            _CACHE.popitem(last=False)
    return p


def clear_plan_cache():
    with _CACHE_LOCK:
        _CACHE.clear()
