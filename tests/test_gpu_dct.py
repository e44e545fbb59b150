"""The exact device Neumann solve (gdp_neumann_dct_solve, solve_neumann of
gradient-domain/src/gradient_domain/solvers.py:224-262) against the oracle's scipy DCT
and against the closed form of gradient-domain/tests/test_solvers.py:90-108."""

import ctypes

import numpy as np
import pytest

from oracle import generators as G
from oracle import operators as OO
from oracle import solvers as OS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gd():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gdpoisson

    return gdpoisson


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("shape", [(16, 16), (45, 75), (96, 128), (120, 90), (1, 8), (8, 1),
                                   (64, 200), (243, 125), (1024, 1024), (4096, 12)])
def test_dct_solve_matches_the_oracle(gd, shape):
    b = np.random.default_rng(sum(shape)).standard_normal(shape)
    assert gd.solvers.dct_supported(*shape)
    u = gd.solve_neumann(b)
    want = OS.solve_neumann(b)
    assert rel(u, want) < 1e-11
    assert abs(u.mean()) < 1e-12


def test_dct_reproduces_an_image(gd):
    # test_solvers.py:90-108: the solve of div(grad t) returns t - mean(t)
    t = G.blob_image((90, 120), seed=11)
    b = OO.divergence(*OO.gradient(t))
    u = gd.solve_neumann(b)
    assert np.abs(u - (t - t.mean())).max() < 1e-10


def test_unsupported_sides_fall_back_to_multigrid(gd):
    import torch

    from gdpoisson import _lib

    shape = (7, 11)
    assert not gd.solvers.dct_supported(*shape)
    nbytes = ctypes.c_size_t()
    assert _lib.lib().gdp_dct_tables_bytes(7, 11, ctypes.byref(nbytes)) == _lib.GDP_ENOTSUP
    d = torch.zeros(shape, dtype=torch.float64, device="cuda")
    assert _lib.lib().gdp_neumann_dct_solve(ctypes.c_void_p(d.data_ptr()),
                                            ctypes.c_void_p(d.data_ptr()), 7, 11,
                                            ctypes.c_void_p(d.data_ptr()),
                                            ctypes.c_void_p(d.data_ptr()), None) == _lib.GDP_ENOTSUP
    b = np.random.default_rng(3).standard_normal(shape)
    assert rel(gd.solve_neumann(b), OS.solve_neumann(b)) < 1e-9


def test_dct_solve_at_the_tone_map_size(gd):
    import torch

    b = np.random.default_rng(5).standard_normal((2160, 3840))
    u = gd.solve_neumann(torch.from_numpy(b).cuda())
    assert rel(u.cpu().numpy(), OS.solve_neumann(b)) < 1e-10
