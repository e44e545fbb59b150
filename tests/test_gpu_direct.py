"""GPU parity of the direct method (SOLVERS["direct"], solvers.py:85-90).

The reference factorises sparse_laplacian (operators.py:110-136) with SuperLU; the device
solves the same system exactly by DST-I diagonalisation (sides m with 2 (m + 1) a
product of 2, 3, 5), else by FP64 MG-PCG to 1e-13.  Checked against the oracle's SuperLU
solve, the reference's published direct-method accuracy (docs/benchmark.json) and the
reference's own tests that call method="direct" (test_applications.py:49-55, 149-153).
"""

import json
from pathlib import Path

import numpy as np
import pytest

from oracle import generators as G
from oracle import operators as OO
from oracle import solvers as OS

pytestmark = pytest.mark.gpu

GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_benchmark.json").read_text())


@pytest.fixture(scope="module")
def gd():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gdpoisson

    return gdpoisson


def test_supported_sides(gd):
    from gdpoisson.solvers import dst_supported

    assert dst_supported(63, 63) and dst_supported(1023, 1023) and dst_supported(2047, 1)
    assert dst_supported(47, 95)            # 96 = 2^5 3, 192 = 2^6 3
    assert not dst_supported(47, 61)        # 124 = 4 x 31
    assert not dst_supported(2048, 2048)    # 4098 > 4096


@pytest.mark.parametrize("shape,spacing", [((63, 63), 1.0), ((47, 61), 1.0), ((95, 127), 0.5),
                                           ((1, 9), 1.0), ((511, 511), 1.0),
                                           ((2, 2047), 2.0)])
def test_direct_matches_superlu(gd, shape, spacing):
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    b = rng.standard_normal(shape)
    u, rep = gd.solve_system(b, method="direct", spacing=spacing)
    want, _ = OS.solve_system(b, method="direct", spacing=spacing)
    scale = np.abs(want).max()
    err = float(np.abs(u - want).max())
    print(f"direct {shape} h={spacing}: |u - superlu| {err:.2e} (|u| {scale:.2e}), "
          f"{rep.iterations} it, res {rep.relative_residual:.2e}")
    assert rep.converged and rep.relative_residual < 1e-12
    assert err <= 1e-10 * max(1.0, scale)
    # the operator applied to the answer gives b back (the defining property)
    assert np.abs(OO.laplacian(u, spacing) - b).max() < 1e-9 * np.abs(b).max()


@pytest.mark.parametrize("size", [63, 127, 255, 511, 1023])
def test_direct_reproduces_the_published_accuracy(gd, size):
    """docs/benchmark.json records the direct method's error against the manufactured
    solution at 63..511 (4.4e-15 .. 1.2e-13, about 4x per doubling of the side); the DST
    solve stays within 10x of it, and at 1023 within 10x of the 511 record times 4."""
    b, x = G.manufactured_problem((size, size), seed=0)
    u, rep = gd.solve_system(b, method="direct")
    err = float(np.abs(u - x).max())
    rec = {r["shape"][0]: r["max_error"] for r in GOLDEN["records"]["direct"]}
    bound = 10.0 * (rec[size] if size in rec else 4.0 * rec[511])
    print(f"direct {size}: err {err:.3e} (bound {bound:.3e}), res {rep.relative_residual:.2e}")
    assert rep.iterations == 0 and rep.converged
    assert err < max(bound, 1e-13)
    assert rep.relative_residual < 1e-13


def test_direct_is_in_the_registry_and_keeps_the_contract(gd):
    assert "direct" in gd.SOLVERS
    b, x = G.manufactured_problem((63, 63), seed=3)
    u, it, conv = gd.SOLVERS["direct"](b, 1.0, 1e-8, 2000)
    assert isinstance(u, np.ndarray) and it == 0 and conv
    assert np.abs(u - x).max() < 1e-12
    b0 = b.copy()
    gd.solve_system(b, method="direct")
    assert np.array_equal(b, b0)  # b is never modified
    z, rep = gd.solve_system(np.zeros((31, 31)), method="direct")
    assert np.all(z == 0.0) and rep.relative_residual == 0.0
    with pytest.raises(ValueError):
        gd.solve_system(b, method="direct", boundary="neumann")


def test_solve_dirichlet_direct_reproduces_an_image(gd):
    """test_applications.py:49-55 with method="direct"."""
    image = np.random.default_rng(0).random((40, 52, 3))
    result, reports = gd.solve_dirichlet(gd.GuidanceField.of(image), image, method="direct")
    assert np.abs(result - image).max() < 1e-9
    assert len(reports) == 3


@pytest.mark.parametrize("shape", [(96, 96), (97, 129)])
def test_texture_flatten_direct_with_every_edge_is_the_identity(gd, shape):
    """test_applications.py:149-153 with method="direct" (97 x 129: the 95 x 127 interior
    is on the DST path; 96 x 96: the MG-PCG fallback)."""
    image = np.random.default_rng(1).random(shape + (3,))
    result, _ = gd.texture_flatten(image, np.ones(shape), method="direct")
    assert np.abs(result - image).max() < 1e-9


def test_solver_scaling_on_the_device(gd):
    """benchmark.solver_scaling (benchmark.py:98-149) through the device solvers: every
    method meets the tolerance and recovers the manufactured solution; the iteration
    counts of the multigrid methods are the published ones."""
    res = gd.solver_scaling([63, 127], methods=["direct", "cg", "multigrid", "mgcg"])
    for method, records in res.items():
        assert [r.shape for r in records] == [(63, 63), (127, 127)]
        for r in records:
            assert r.relative_residual < 1e-8 and r.max_error < 1e-6
            pub = [p for p in GOLDEN["records"][method] if p["shape"] == list(r.shape)][0]
            if method in ("multigrid", "mgcg"):
                assert r.iterations == pub["iterations"]
    table = gd.BenchmarkTable(res).to_markdown()
    assert table.count("\n") == 2 + 8 - 1
