"""ctypes binding of libgdpoisson.so (declarations in include/gdpoisson.h).

The library is required: if it is missing or cannot be loaded every entry point
raises, there is no CPU fallback.  PyTorch supplies device memory (tensors) and the
HIP stream (``torch.cuda.current_stream().cuda_stream``).
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int, c_int32, c_size_t, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libgdpoisson.so"

GDP_OK, GDP_EINVAL, GDP_ENOMEM, GDP_EHIP, GDP_ENOTSUP = 0, 1, 2, 3, 4
DIRICHLET, NEUMANN = 0, 1
FP64, MIXED = 0, 1

# every symbol declared in include/gdpoisson.h
EXPORTS = (
    "gdp_last_error",
    "gdp_version",
    "gdp_laplacian",
    "gdp_gradient",
    "gdp_divergence",
    "gdp_fold_boundary",
    "gdp_clone_rhs",
    "gdp_guidance",
    "gdp_illumination_field",
    "gdp_clone_rhs_masked",
    "gdp_masked_rhs",
    "gdp_plan_set_mask",
    "gdp_tm_work_bytes",
    "gdp_tm_log_luminance",
    "gdp_tm_downsample",
    "gdp_tm_level_factor",
    "gdp_tm_rhs",
    "gdp_tm_compose",
    "gdp_percentile_work_bytes",
    "gdp_percentile",
    "gdp_tm_normalise",
    "gdp_dct_supported",
    "gdp_dct_tables_bytes",
    "gdp_dct_tables",
    "gdp_dct_work_bytes",
    "gdp_neumann_dct_solve",
    "gdp_neumann_dct_solve_static",
    "gdp_dst_supported",
    "gdp_dst_tables_bytes",
    "gdp_dst_tables",
    "gdp_dirichlet_dst_solve",
    "gdp_smooth",
    "gdp_restrict",
    "gdp_prolong",
    "gdp_relative_residual",
    "gdp_plan_create",
    "gdp_plan_workspace_bytes",
    "gdp_plan_bind",
    "gdp_plan_destroy",
    "gdp_plan_levels",
    "gdp_plan_level_shape",
    "gdp_plan_set_tile",
    "gdp_tile_menu",
    "gdp_vcycle",
    "gdp_mg_solve",
    "gdp_mgcg_solve",
    "gdp_cg_solve",
    "gdp_plan_profile",
    "gdp_plan_profile_count",
    "gdp_plan_profile_entry",
)


class Problem(ctypes.Structure):
    _fields_ = [
        ("rows", c_int32),
        ("cols", c_int32),
        ("boundary", c_int32),
        ("flags", c_int32),
        ("spacing", c_double),
    ]


class Options(ctypes.Structure):
    _fields_ = [
        ("precision", c_int32),
        ("pre_smoothing", c_int32),
        ("post_smoothing", c_int32),
        ("coarsest_size", c_int32),
        ("coarsest_sweeps", c_int32),
        ("max_iterations", c_int32),
        ("flexible", c_int32),
        ("use_graph", c_int32),
        ("tolerance", c_double),
    ]


class Report(ctypes.Structure):
    _fields_ = [
        ("iterations", c_int32),
        ("converged", c_int32),
        ("relative_residual", c_double),
        ("recursive_residual", c_double),
    ]


class GdpError(RuntimeError):
    """A libgdpoisson call failed (HIP error or internal limit)."""


_lib = None

P = c_void_p
PP = POINTER(c_void_p)


def _declare(lib):
    sig = {
        "gdp_last_error": (ctypes.c_char_p, []),
        "gdp_version": (c_int, []),
        "gdp_laplacian": (c_int, [P, P, c_int, c_int, c_int, c_double, c_int, P]),
        "gdp_gradient": (c_int, [P, P, P, c_int, c_int, c_int, P]),
        "gdp_divergence": (c_int, [P, P, P, c_int, c_int, c_int, P]),
        "gdp_fold_boundary": (c_int, [P, P, P, c_int, c_int, c_int, c_double, P]),
        "gdp_clone_rhs": (c_int, [P, c_int, c_int, c_int, c_int, P, P, c_int, c_int, c_int,
                                  c_int, c_int, c_int, c_int, c_int, P, P]),
        "gdp_guidance": (c_int, [P, P, P, P, P, P, c_double, c_int, c_int, c_int, c_int, P, P,
                                 P]),
        "gdp_illumination_field": (c_int, [P, P, c_size_t, c_double, c_double, c_double, P, P,
                                           P]),
        "gdp_clone_rhs_masked": (c_int, [P, c_int, c_int, c_int, c_int, P, P, c_int, c_int,
                                         c_int, c_int, c_int, c_int, c_int, c_int, P, P, P]),
        "gdp_masked_rhs": (c_int, [P, P, P, P, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                   P, P]),
        "gdp_plan_set_mask": (c_int, [P, c_int, P]),
        "gdp_tm_work_bytes": (c_int, [c_int, c_int, POINTER(c_size_t)]),
        "gdp_tm_log_luminance": (c_int, [P, c_int, c_int, P, P]),
        "gdp_tm_downsample": (c_int, [P, c_int, c_int, P, P]),
        "gdp_tm_level_factor": (c_int, [P, c_int, c_int, c_double, c_double, c_double, c_double,
                                        P, c_int, c_int, P, P, P]),
        "gdp_tm_rhs": (c_int, [P, P, c_int, c_int, P, P]),
        "gdp_tm_compose": (c_int, [P, P, c_int, c_int, c_double, P, P, P]),
        "gdp_percentile_work_bytes": (c_int, [POINTER(c_size_t)]),
        "gdp_percentile": (c_int, [P, c_size_t, c_double, P, P, P]),
        "gdp_tm_normalise": (c_int, [P, c_size_t, P, P]),
        "gdp_dct_supported": (c_int, [c_int, c_int]),
        "gdp_dct_tables_bytes": (c_int, [c_int, c_int, POINTER(c_size_t)]),
        "gdp_dct_tables": (c_int, [c_int, c_int, P, P]),
        "gdp_dct_work_bytes": (c_int, [c_int, c_int, POINTER(c_size_t)]),
        "gdp_neumann_dct_solve": (c_int, [P, P, c_int, c_int, P, P, P]),
        "gdp_neumann_dct_solve_static": (c_int, [P, P, c_int, c_int, P, P]),
        "gdp_dst_supported": (c_int, [c_int, c_int]),
        "gdp_dst_tables_bytes": (c_int, [c_int, c_int, POINTER(c_size_t)]),
        "gdp_dst_tables": (c_int, [c_int, c_int, P, P]),
        "gdp_dirichlet_dst_solve": (c_int, [P, P, c_int, c_int, c_double, P, P, P]),
        "gdp_smooth": (c_int, [P, P, c_int, c_int, c_double, c_int, c_int, c_int, P]),
        "gdp_restrict": (c_int, [P, P, c_int, c_int, c_int, c_int, c_int, P]),
        "gdp_prolong": (c_int, [P, P, c_int, c_int, c_int, c_int, c_int, P]),
        "gdp_relative_residual": (
            c_int,
            [P, P, c_int, c_int, c_int, c_double, c_int, POINTER(c_double), P],
        ),
        "gdp_plan_create": (c_int, [POINTER(Problem), c_int, POINTER(Options), POINTER(P)]),
        "gdp_plan_workspace_bytes": (c_int, [P, POINTER(c_size_t)]),
        "gdp_plan_bind": (c_int, [P, P, c_size_t]),
        "gdp_plan_destroy": (c_int, [P]),
        "gdp_plan_levels": (c_int, [P, c_int, POINTER(c_int)]),
        "gdp_plan_level_shape": (c_int, [P, c_int, c_int, POINTER(c_int), POINTER(c_int)]),
        "gdp_plan_set_tile": (c_int, [P, c_int, c_int, c_int]),
        "gdp_tile_menu": (c_int, [c_int, c_int, c_int, POINTER(c_int), POINTER(c_int)]),
        "gdp_vcycle": (c_int, [P, PP, PP, PP, P]),
        "gdp_mg_solve": (c_int, [P, PP, PP, PP, POINTER(Report), POINTER(c_double), P]),
        "gdp_mgcg_solve": (c_int, [P, PP, PP, POINTER(Report), P]),
        "gdp_cg_solve": (c_int, [P, PP, PP, POINTER(Report), P]),
        "gdp_plan_profile": (c_int, [P, c_int]),
        "gdp_plan_profile_count": (c_int, [P, POINTER(c_int)]),
        "gdp_plan_profile_entry": (
            c_int,
            [P, c_int, ctypes.c_char_p, c_int, POINTER(c_double), POINTER(ctypes.c_longlong),
             POINTER(c_double)],
        ),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def lib():
    """Load libgdpoisson.so once; raise if it is absent (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise GdpError(
                f"{LIB_PATH} is missing: build it with gdpoisson._build.build() "
                "(hipcc, gfx950) before using the solver"
            )
        lb = ctypes.CDLL(str(LIB_PATH))
        _declare(lb)
        _lib = lb
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc == GDP_OK:
        return
    msg = lib().gdp_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == GDP_EINVAL:
        raise ValueError(text)
    if rc == GDP_ENOMEM:
        raise MemoryError(text)
    if rc == GDP_ENOTSUP:
        raise NotImplementedError(text)
    raise GdpError(text)


def ptr_array(ptrs):
    arr = (c_void_p * len(ptrs))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr


def current_stream():
    import torch

    return c_void_p(torch.cuda.current_stream().cuda_stream)
