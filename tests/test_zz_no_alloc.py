"""The allocation-free entry point gdp_neumann_dct_solve_static (no counterpart in the
reference): the exact Neumann solve (solvers.py:224-262) with its work buffer in a static
device array.  Called through the C ABI and compared, bit for bit, with
gdp_neumann_dct_solve given a work buffer, on two calls in a row (the second smaller, so
it runs on what the first left in the array), and with the oracle's DCT solve."""

import ctypes

import numpy as np
import pytest

from oracle import generators as G
from oracle import solvers as OS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def L():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gdpoisson import _lib

    return _lib


def _tables(L, m, n):
    import torch

    nbytes = ctypes.c_size_t()
    assert L.lib().gdp_dct_tables_bytes(m, n, ctypes.byref(nbytes)) == 0
    tab = torch.empty(nbytes.value // 8, dtype=torch.float64, device="cuda")
    assert L.lib().gdp_dct_tables(m, n, ctypes.c_void_p(tab.data_ptr()), None) == 0
    return tab


def test_static_work_solve_equals_the_work_buffer_solve(L):
    import torch

    lib = L.lib()
    for k, (m, n) in enumerate([(512, 480), (96, 60)]):
        b_np, _ = G.neumann_problem((m, n), seed=500 + k)
        b = torch.from_numpy(b_np).cuda()
        tab = _tables(L, m, n)
        work = torch.empty(2 * m * n, dtype=torch.float64, device="cuda")
        want = torch.empty_like(b)
        got = torch.full_like(b, float("nan"))
        assert lib.gdp_neumann_dct_solve(ctypes.c_void_p(b.data_ptr()),
                                         ctypes.c_void_p(want.data_ptr()), m, n,
                                         ctypes.c_void_p(tab.data_ptr()),
                                         ctypes.c_void_p(work.data_ptr()), None) == 0
        assert lib.gdp_neumann_dct_solve_static(ctypes.c_void_p(b.data_ptr()),
                                                ctypes.c_void_p(got.data_ptr()), m, n,
                                                ctypes.c_void_p(tab.data_ptr()), None) == 0
        torch.cuda.synchronize()
        assert torch.equal(got, want), (m, n)
        assert np.abs(got.cpu().numpy() - OS.solve_neumann(b_np)).max() < 1e-9


def test_static_work_solve_rejects_sizes_beyond_its_array(L):
    import torch

    lib = L.lib()
    b = torch.zeros(4, dtype=torch.float64, device="cuda")
    # 4096 x 1200 needs 2 * 4,915,200 doubles of work, more than the 64 MiB array; the
    # call is refused before anything is launched
    rc = lib.gdp_neumann_dct_solve_static(ctypes.c_void_p(b.data_ptr()),
                                          ctypes.c_void_p(b.data_ptr()), 4096, 1200,
                                          ctypes.c_void_p(b.data_ptr()), None)
    assert rc == L.GDP_ENOMEM
    assert b"static" in lib.gdp_last_error()
    rc = lib.gdp_neumann_dct_solve_static(ctypes.c_void_p(b.data_ptr()),
                                          ctypes.c_void_p(b.data_ptr()), 7, 11,
                                          ctypes.c_void_p(b.data_ptr()), None)
    assert rc == L.GDP_ENOTSUP
