"""C-ABI checks that need no GPU: the library loads, exports every declared symbol,
the ctypes structures match the C layout, and the host-side planner (level
hierarchy, validation, workspace sizing) agrees with the oracle."""

import ctypes
import re
import subprocess
from pathlib import Path

import pytest

from oracle import multigrid as M

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "gdpoisson.h"


@pytest.fixture(scope="module")
def L():
    from gdpoisson import _lib

    if not _lib.LIB_PATH.exists():
        import __graft_entry__

        __graft_entry__.build()
    return _lib


def declared():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(gdp_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(L):
    lib = L.lib()
    names = declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(L.EXPORTS) == names


def test_struct_layout_matches_c(L, tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "gdpoisson.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu\\n\", sizeof(gdp_problem),"
        " sizeof(gdp_options), sizeof(gdp_report), offsetof(gdp_problem, spacing),"
        " offsetof(gdp_options, tolerance), offsetof(gdp_report, recursive_residual));"
        "return 0;}\n"
    )
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True,
                                          check=True).stdout.split()]
    want = [ctypes.sizeof(L.Problem), ctypes.sizeof(L.Options), ctypes.sizeof(L.Report),
            L.Problem.spacing.offset, L.Options.tolerance.offset,
            L.Report.recursive_residual.offset]
    assert got == want


def make_plan(L, shapes, bc, precision=0, **kw):
    n = len(shapes)
    probs = (L.Problem * n)()
    for i, (m, nn) in enumerate(shapes):
        probs[i] = L.Problem(m, nn, bc, kw.get("flags", 0), 1.0)
    cs = kw.get("coarsest_size", 16 if bc == L.NEUMANN else 8)
    opt = L.Options(precision, kw.get("pre", 1), kw.get("post", 1), cs, 60,
                    kw.get("max_iterations", 100), 0, 1, 1e-8)
    plan = ctypes.c_void_p()
    rc = L.lib().gdp_plan_create(probs, n, ctypes.byref(opt), ctypes.byref(plan))
    return rc, plan


def plan_levels(L, plan, i):
    lib = L.lib()
    nl = ctypes.c_int()
    assert lib.gdp_plan_levels(plan, i, ctypes.byref(nl)) == 0
    out = []
    for lv in range(nl.value):
        m, n = ctypes.c_int(), ctypes.c_int()
        assert lib.gdp_plan_level_shape(plan, i, lv, ctypes.byref(m), ctypes.byref(n)) == 0
        out.append((m.value, n.value))
    return out


@pytest.mark.parametrize("bcname", ["dirichlet", "neumann"])
def test_planner_level_chain_matches_oracle(L, bcname):
    bc = L.NEUMANN if bcname == "neumann" else L.DIRICHLET
    shapes = [(1022, 1022), (1024, 1024), (47, 61), (2160, 3840), (255, 257), (9, 300)]
    rc, plan = make_plan(L, shapes, bc)
    assert rc == 0
    opts = M.MultigridOptions(coarsest_size=16 if bcname == "neumann" else 8)
    for i, s in enumerate(shapes):
        assert plan_levels(L, plan, i) == [tuple(x) for x in M.level_shapes(s, opts, bcname)]
    ws = ctypes.c_size_t()
    assert L.lib().gdp_plan_workspace_bytes(plan, ctypes.byref(ws)) == 0
    # the workspace holds at least x, r, d0, d1, q in FP64 for every fine grid
    assert ws.value >= sum(5 * 8 * m * n for m, n in shapes)
    assert L.lib().gdp_plan_destroy(plan) == 0


def test_planner_rejects_bad_input(L):
    rc, _ = make_plan(L, [(0, 10)], L.DIRICHLET)
    assert rc == L.GDP_EINVAL
    rc, _ = make_plan(L, [(64, 64)], L.DIRICHLET, pre=7)
    assert rc == L.GDP_ENOTSUP
    assert b"smoothing" in L.lib().gdp_last_error()
    rc, _ = make_plan(L, [(64, 64)], 5)
    assert rc == L.GDP_EINVAL
    # a coarsest level far too large for the LDS tail is refused with a message
    rc, _ = make_plan(L, [(3, 100000)], L.NEUMANN)
    assert rc == L.GDP_ENOTSUP and b"LDS" in L.lib().gdp_last_error()


def test_python_errors_map_status_codes(L):
    with pytest.raises(ValueError):
        L.check(L.GDP_EINVAL)
    with pytest.raises(MemoryError):
        L.check(L.GDP_ENOMEM)
    with pytest.raises(NotImplementedError):
        L.check(L.GDP_ENOTSUP)
    with pytest.raises(L.GdpError):
        L.check(L.GDP_EHIP)


def test_tile_menu(L):
    th, tw = ctypes.c_int(), ctypes.c_int()
    for kind, size in ((0, 16), (1, 6), (2, 6)):
        for idx in range(size):
            for nu in range(4):
                assert L.lib().gdp_tile_menu(kind, idx, nu, ctypes.byref(th),
                                             ctypes.byref(tw)) == 0
                assert th.value > 0 and th.value % 4 == 0 and tw.value > 0
    # strips are 128 fine columns wide including their margins
    L.lib().gdp_tile_menu(0, 0, 2, ctypes.byref(th), ctypes.byref(tw))
    assert tw.value == 58
    L.lib().gdp_tile_menu(1, 0, 2, ctypes.byref(th), ctypes.byref(tw))
    assert tw.value == 116
    # the down menu's entries 4, 5, 6 are streamed strips of 8, 16 and 32 coarse rows
    # (32 for nu = 1 and 2; other sweep counts get the 16-row strips)
    L.lib().gdp_tile_menu(0, 5, 2, ctypes.byref(th), ctypes.byref(tw))
    assert (th.value, tw.value) == (64, 58)
    L.lib().gdp_tile_menu(0, 6, 2, ctypes.byref(th), ctypes.byref(tw))
    assert th.value == 128
    L.lib().gdp_tile_menu(0, 6, 3, ctypes.byref(th), ctypes.byref(tw))
    assert th.value == 64
    # entries 7, 8, 9: block-tiled down kernels of 8, 16 and 32 coarse rows per block
    for idx, rows in ((7, 8), (8, 16), (9, 32)):
        for nu in range(4):
            L.lib().gdp_tile_menu(0, idx, nu, ctypes.byref(th), ctypes.byref(tw))
            assert th.value == rows
    # up entries 4, 5: block-tiled up kernels of 16 and 32 fine rows per block
    for idx, rows in ((4, 16), (5, 32)):
        L.lib().gdp_tile_menu(1, idx, 2, ctypes.byref(th), ctypes.byref(tw))
        assert th.value == rows
    # entries 10 .. 15: the streamed strips of entries 4, 5, 6 and the blocks of entries
    # 7, 8, 9 carrying the PCG update (the same tiles)
    th2, tw2 = ctypes.c_int(), ctypes.c_int()
    for idx in range(10, 16):
        for nu in range(4):
            L.lib().gdp_tile_menu(0, idx, nu, ctypes.byref(th), ctypes.byref(tw))
            L.lib().gdp_tile_menu(0, idx - 6, nu, ctypes.byref(th2), ctypes.byref(tw2))
            assert (th.value, tw.value) == (th2.value, tw2.value)
    for bad in ((3, 0, 1), (0, 16, 1), (1, 6, 1), (2, 6, 1), (0, 0, 4)):
        assert L.lib().gdp_tile_menu(*bad, ctypes.byref(th), ctypes.byref(tw)) == L.GDP_EINVAL


def test_fused_update_entries_are_finest_level_only(L):
    """The down entries that carry the PCG update are refused on the coarser levels."""
    rc, plan = make_plan(L, [(256, 256)], L.NEUMANN)
    assert rc == 0
    assert L.lib().gdp_plan_set_tile(plan, 0, 0, 11) == 0
    assert L.lib().gdp_plan_set_tile(plan, 0, 1, 11) == L.GDP_EINVAL
    assert b"finest" in L.lib().gdp_last_error()
    assert L.lib().gdp_plan_set_tile(plan, 0, 1, 5) == 0
    assert L.lib().gdp_plan_destroy(plan) == 0


def test_planner_problem_flags(L):
    """GDP_PROBLEM_MASKED (the problem will get a masked domain) needs the Dirichlet
    boundary; other flag bits are rejected; the level chain is the same."""
    rc, plan = make_plan(L, [(255, 257)], L.DIRICHLET, flags=1)
    assert rc == 0
    rc0, plan0 = make_plan(L, [(255, 257)], L.DIRICHLET)
    assert rc0 == 0 and plan_levels(L, plan, 0) == plan_levels(L, plan0, 0)
    ws, ws0 = ctypes.c_size_t(), ctypes.c_size_t()
    assert L.lib().gdp_plan_workspace_bytes(plan, ctypes.byref(ws)) == 0
    assert L.lib().gdp_plan_workspace_bytes(plan0, ctypes.byref(ws0)) == 0
    assert ws.value >= ws0.value  # its coarse levels are strip levels with tile tables
    assert L.lib().gdp_plan_destroy(plan) == 0 and L.lib().gdp_plan_destroy(plan0) == 0
    rc, _ = make_plan(L, [(64, 64)], L.NEUMANN, flags=1)
    assert rc == L.GDP_EINVAL
    rc, _ = make_plan(L, [(64, 64)], L.DIRICHLET, flags=2)
    assert rc == L.GDP_EINVAL and b"flags" in L.lib().gdp_last_error()
