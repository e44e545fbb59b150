"""Pins of the FP64 oracle (CPU).

Each function of ``oracle/`` is checked against something that does not restate its
own formula: the reference's published solver-scaling numbers
(tests/golden/reference_benchmark.json, from gradient-domain/docs/benchmark.json),
the closed forms and thresholds of the reference tests (gradient-domain/tests/),
dense brute-force matrices, adjoint/symmetry identities and the exact DCT solve.
"""

import json
from pathlib import Path

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import generators as G
from oracle import multigrid as M
from oracle import operators as O
from oracle import poisson as OP
from oracle import solvers as S

GOLDEN = json.loads((Path(__file__).parent / "golden" / "reference_benchmark.json").read_text())


def published(method, size):
    for r in GOLDEN["records"][method]:
        if r["shape"] == [size, size]:
            return r
    raise KeyError((method, size))


# ----------------------------------------------------------------- published numbers
@pytest.mark.parametrize("size", [63, 127, 255, 511])
def test_mgcg_reproduces_published_iterations_and_residuals(size):
    """docs/benchmark.json: 7 iterations; residual and max error to 6 digits."""
    b, x = G.manufactured_problem((size, size), seed=0)
    u, rep = S.solve_system(b, "mgcg", tolerance=1e-8, max_iterations=5000)
    ref = published("mgcg", size)
    assert rep.iterations == ref["iterations"] == 7
    assert rep.relative_residual == pytest.approx(ref["relative_residual"], rel=1e-6)
    assert float(np.abs(u - x).max()) == pytest.approx(ref["max_error"], rel=1e-6)


@pytest.mark.parametrize("size", [63, 127, 255])
def test_multigrid_reproduces_published_cycles(size):
    b, x = G.manufactured_problem((size, size), seed=0)
    u, rep = S.solve_system(b, "multigrid", tolerance=1e-8, max_iterations=5000)
    ref = published("multigrid", size)
    assert rep.iterations == ref["iterations"]
    assert rep.relative_residual == pytest.approx(ref["relative_residual"], rel=1e-6)
    assert float(np.abs(u - x).max()) == pytest.approx(ref["max_error"], rel=1e-6)


@pytest.mark.parametrize("size", [63, 127])
def test_cg_reproduces_published_iterations(size):
    b, x = G.manufactured_problem((size, size), seed=0)
    u, rep = S.solve_system(b, "cg", tolerance=1e-8, max_iterations=5000)
    ref = published("cg", size)
    assert abs(rep.iterations - ref["iterations"]) <= 1
    assert rep.relative_residual == pytest.approx(ref["relative_residual"], rel=1e-3)


@pytest.mark.parametrize("size", [63, 127, 255])
def test_direct_reproduces_published_accuracy(size):
    b, x = G.manufactured_problem((size, size), seed=0)
    u, rep = S.solve_system(b, "direct")
    assert rep.relative_residual < 1e-14
    assert float(np.abs(u - x).max()) < 1e-12


# ----------------------------------------------------------------- reference tests
def test_divergence_is_the_negative_adjoint_of_the_gradient():
    rng = np.random.default_rng(11)
    u = rng.standard_normal((23, 31))
    vx, vy = rng.standard_normal((23, 31)), rng.standard_normal((23, 31))
    gx, gy = O.gradient(u)
    left = float(np.sum(gx * vx) + np.sum(gy * vy))
    right = -float(np.sum(u * O.divergence(vx, vy)))
    assert left == pytest.approx(right, rel=1e-12)


def test_gradient_closed_forms():
    gx, gy = O.gradient(np.full((9, 12), 3.5))
    assert np.allclose(gx, 0.0) and np.allclose(gy, 0.0)
    rows, cols = np.mgrid[0:8, 0:10]
    gx, gy = O.gradient(2.0 * cols + 3.0 * rows)
    assert np.allclose(gx[:, :-1], 2.0) and np.allclose(gy[:-1, :], 3.0)
    assert np.all(gx[:, -1] == 0.0) and np.all(gy[-1, :] == 0.0)


def test_composition_is_the_neumann_laplacian():
    rng = np.random.default_rng(11)
    u = rng.standard_normal((17, 21))
    p = np.pad(u, 1, mode="edge")
    expected = p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:] - 4.0 * u
    assert np.allclose(O.divergence(*O.gradient(u)), expected)
    assert np.allclose(O.neumann_laplacian(u), expected)
    assert np.allclose(O.neumann_laplacian(u, 0.5), 4.0 * expected)


def test_dense_and_sparse_laplacians_agree():
    rng = np.random.default_rng(11)
    u = rng.standard_normal((13, 19))
    for h in (1.0, 0.5, 2.0):
        dense = O.laplacian(u, h)
        sparse = (O.sparse_laplacian(u.shape, h) @ u.reshape(-1)).reshape(u.shape)
        assert np.allclose(dense, sparse)


def test_laplacian_is_symmetric_negative_definite():
    a = O.sparse_laplacian((7, 9)).toarray()
    assert np.allclose(a, a.T)
    assert np.linalg.eigvalsh(a).max() < 0.0


def test_folding_the_boundary():
    rng = np.random.default_rng(11)
    full = rng.standard_normal((14, 18))
    interior = full[1:-1, 1:-1]
    padded = O.laplacian(full)[1:-1, 1:-1]
    folded = O.laplacian(interior) - O.fold_boundary(np.zeros_like(interior), full)
    assert np.allclose(folded, padded)
    f = O.fold_boundary(np.zeros((3, 3)), np.ones((5, 5)), spacing=0.5)
    assert f[0, 0] == pytest.approx(-8.0)
    assert f[0, 1] == pytest.approx(-4.0)
    assert f[1, 1] == pytest.approx(0.0)


@pytest.mark.parametrize("bc", ["dirichlet", "neumann"])
def test_restriction_is_the_scaled_adjoint_of_prolongation(bc):
    rng = np.random.default_rng(6)
    for fine_shape in ((31, 25), (30, 24), (17, 40)):
        cs = M.coarse_shape(fine_shape, bc)
        f = rng.standard_normal(fine_shape)
        c = rng.standard_normal(cs)
        left = float(np.sum(M.restrict(f, bc) * c))
        right = 0.25 * float(np.sum(f * M.prolong(c, fine_shape, bc)))
        assert left == pytest.approx(right, rel=1e-12)


def test_prolongation_reproduces_a_constant():
    fine = M.prolong(np.ones((7, 7)), (15, 15))
    assert np.allclose(fine[1:-1, 1:-1], 1.0)
    # the cell-centred Neumann interpolation reproduces constants everywhere
    for fine_shape in ((16, 16), (15, 17), (9, 4)):
        cs = M.coarse_shape(fine_shape, "neumann")
        assert np.allclose(M.prolong(np.ones(cs), fine_shape, "neumann"), 1.0)


def brute_prolongation_1d(fine, coarse, bc):
    """Dense P from the geometric definition (not from the oracle's code)."""
    P = np.zeros((fine, coarse))
    if bc == "dirichlet":
        xs_f = np.arange(1, fine + 1) / (fine + 1.0)  # fine vertices of (0, 1)
        xs_c = np.arange(1, coarse + 1) * 2.0 / (fine + 1.0)  # coarse vertex I = fine 2I+1
        for i, x in enumerate(xs_f):
            for I, xc in enumerate(xs_c):
                w = 1.0 - abs(x - xc) * (fine + 1.0) / 2.0
                if w > 0:
                    P[i, I] = w
    else:
        # cell centres (i + 1/2) h and (I + 1/2) 2h; reflect coarse values at the ends
        for i in range(fine):
            x = (i + 0.5) / 2.0 - 0.5  # position in coarse index units
            lo = int(np.floor(x))
            t = x - lo
            for I, w in ((lo, 1.0 - t), (lo + 1, t)):
                I = min(max(I, 0), coarse - 1)
                P[i, I] += w
    return P


@pytest.mark.parametrize("bc", ["dirichlet", "neumann"])
@pytest.mark.parametrize("fine", [7, 8, 15, 16, 9])
def test_prolongation_matches_brute_force(bc, fine):
    coarse = M.coarse_shape((fine, fine), bc)[0]
    pfun = M.neumann_prolongation_1d if bc == "neumann" else M.prolongation_1d
    assert np.allclose(pfun(fine, coarse).toarray(), brute_prolongation_1d(fine, coarse, bc))


@pytest.mark.parametrize("bc", ["dirichlet", "neumann"])
def test_red_black_half_sweep_solves_its_colour_exactly(bc):
    """After colour c is updated each of its points satisfies its own equation."""
    rng = np.random.default_rng(21)
    u = rng.standard_normal((19, 23))
    b = rng.standard_normal((19, 23))
    h = 0.5
    out = M.smooth(u.copy(), b, h, 1, bc, (0, 1))
    mask0 = M.parity_mask(u.shape, 0)
    # after the sweep colour 1 (updated last) satisfies (A out) = b exactly
    r = b - M.apply_operator(out, h, bc)
    assert np.abs(r[~mask0]).max() < 1e-12
    # colour 0 was exact before colour 1 moved: use the colour-1 values of u
    half = np.where(mask0, out, u)
    r0 = b - M.apply_operator(half, h, bc)
    assert np.abs(r0[mask0]).max() < 1e-12


@pytest.mark.parametrize("size", [31, 63, 127, 255])
def test_multigrid_convergence_does_not_depend_on_the_grid(size):
    b, _ = G.manufactured_problem((size, size), seed=5)
    _, rep = M.solve(b, options=M.MultigridOptions(tolerance=1e-10))
    assert rep.converged and rep.cycles <= 10 and rep.convergence_factor < 0.1


def test_one_cycle_reduces_the_residual_by_an_order_of_magnitude():
    b, _ = G.manufactured_problem((127, 127), seed=7)
    _, rep = M.solve(b, options=M.MultigridOptions(max_cycles=1, tolerance=0.0))
    assert rep.residuals[1] < 0.1 * rep.residuals[0]


def test_dst_coarsest_solve_is_exact():
    rng = np.random.default_rng(31)
    for shape, h in (((13, 17), 0.5), ((30, 30), 4.0), ((1, 5), 1.0)):
        u = rng.standard_normal(shape)
        assert np.abs(M.dst_dirichlet_solve(O.laplacian(u, h), h) - u).max() < 1e-11
    # the exact-coarsest Dirichlet cycle: one cycle on a coarsest-only grid is exact
    u = rng.standard_normal((20, 24))
    b = O.laplacian(u)
    opts = M.MultigridOptions(coarsest_size=32, coarsest_sweeps=0)
    assert np.abs(M.v_cycle(np.zeros_like(b), b, 1.0, opts) - u).max() < 1e-11


def test_neumann_solver_is_exact():
    rng = np.random.default_rng(8)
    for shape in ((64, 64), (97, 131)):
        expected = rng.standard_normal(shape)
        expected -= expected.mean()
        p = np.pad(expected, 1, mode="edge")
        b = p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:] - 4.0 * expected
        assert np.abs(S.solve_neumann(b) - expected).max() < 1e-10
        assert np.abs(M.dct_neumann_solve(b) - expected).max() < 1e-10


def test_neumann_solution_has_zero_mean():
    rng = np.random.default_rng(9)
    b = rng.standard_normal((32, 40))
    b -= b.mean()
    assert abs(float(S.solve_neumann(b).mean())) < 1e-12


def test_residual_matches_the_operator():
    b, _ = G.manufactured_problem((33, 33), seed=10)
    u, rep = S.solve_system(b, "mgcg", tolerance=1e-9)
    res = float(np.linalg.norm(b - O.laplacian(u))) / float(np.linalg.norm(b))
    assert res == pytest.approx(rep.relative_residual, rel=1e-9)


def test_unknown_method_and_zero_rhs():
    with pytest.raises(KeyError):
        S.solve_system(np.zeros((8, 8)), method="jacobi")
    u, rep = S.solve_system(np.zeros((16, 16)))
    assert np.allclose(u, 0.0) and rep.relative_residual == 0.0


# ----------------------------------------------------------------- Neumann MG (new)
def test_neumann_vcycle_is_symmetric():
    """<V r1, r2> = <r1, V r2> for the zero-guess cycle (symmetric preconditioner)."""
    rng = np.random.default_rng(30)
    opts = M.MultigridOptions(pre_smoothing=1, post_smoothing=1, coarsest_size=8)
    for shape in ((32, 32), (33, 29)):
        r1 = rng.standard_normal(shape)
        r2 = rng.standard_normal(shape)
        r1 -= r1.mean()
        r2 -= r2.mean()
        v1 = M.v_cycle(np.zeros(shape), r1, 1.0, opts, "neumann")
        v2 = M.v_cycle(np.zeros(shape), r2, 1.0, opts, "neumann")
        assert float(np.sum(v1 * r2)) == pytest.approx(float(np.sum(r1 * v2)), rel=1e-10)


def test_neumann_level_chain():
    opts = S.default_pcg_options("neumann")
    assert M.level_shapes((1024, 1024), opts, "neumann")[-1] == (16, 16)
    assert len(M.level_shapes((1024, 1024), opts, "neumann")) == 7
    assert M.level_shapes((2160, 3840), opts, "neumann")[-1] == (9, 15)
    o32 = M.MultigridOptions(coarsest_size=32)
    assert M.level_shapes((1024, 1024), o32, "neumann")[-1] == (32, 32)
    d = M.level_shapes((1022, 1022), M.MultigridOptions(), "dirichlet")
    assert [s[0] for s in d] == [1022, 510, 254, 126, 62, 30, 14, 6]


@pytest.mark.parametrize("shape", [(96, 96), (81, 130)])
def test_neumann_mgcg_matches_dct_and_closed_form(shape):
    b, exact = G.neumann_problem(shape, seed=1001)
    u, it, conv = S.mgcg(b, tolerance=1e-10, bc="neumann")
    assert conv
    assert np.abs(u - exact).max() < 1e-7
    assert np.abs(u - S.solve_neumann(b)).max() < 1e-7
    assert S.relative_residual(u, b, 1.0, "neumann") <= 1e-10


def test_mixed_precision_oracle_tracks_fp64():
    """The same cycle in FP32 (FP64 outer PCG) needs at most one more iteration."""
    b, exact = G.neumann_problem((128, 160), seed=1002)
    opts = S.default_pcg_options("neumann", "mixed")
    u64, it64, _ = S.mgcg(b, bc="neumann", options=opts)
    u32, it32, c32 = S.mgcg(b, bc="neumann", precision="mixed", flexible=True, options=opts)
    assert c32 and abs(it32 - it64) <= 1
    assert np.abs(u32 - exact).max() < 1e-6


def test_generators_are_seeded_and_shaped():
    a = G.blob_image((64, 80), 3)
    assert a.shape == (64, 80) and np.array_equal(a, G.blob_image((64, 80), 3))
    m = G.blob_mask((128, 128), 5)
    assert 0.25 < m.mean() < 0.35
    shapes = G.batch_shapes()
    assert len(shapes) == 64
    for h, w in shapes:
        s = np.sqrt(h * w)
        assert 128 * 0.99 <= s <= 2048 * 1.01 and 0.45 <= h / w <= 2.1
    rhs, tgt, mask = G.compositing_problem((64, 72))
    assert rhs.shape == (3, 62, 70) and tgt.shape == (3, 64, 72)
    b = G.tonemap_problem((40, 50), 7)
    assert abs(b.sum()) < 1e-8 * np.abs(b).sum()


def test_integrating_an_image_reproduces_it():
    """test_applications.py:49-55: integrating an image's own gradient returns it."""
    from oracle import poisson as OP

    image = np.random.default_rng(0).random((40, 52, 3))
    result, reports = OP.solve_dirichlet(OP.GuidanceField.of(image), image, method="direct")
    assert np.abs(result - image).max() < 1e-9
    assert len(reports) == 3


@pytest.mark.parametrize("mode", ["import", "mixed", "average"])
def test_cloning_the_target_into_itself_changes_nothing(mode):
    from oracle import poisson as OP

    rng = np.random.default_rng(1)
    target = rng.random((48, 60, 3))
    mask = np.zeros((30, 20), dtype=bool)
    mask[5:25, 3:17] = True
    source = target[10:40, 20:40]
    result, _ = OP.seamless_clone(source, target, mask, offset=(10, 20), mode=mode, margin=6,
                                  domain="rectangle")
    assert np.abs(result - target).max() < 1e-9


def test_clone_rhs_matches_the_compositing_generator():
    from oracle import poisson as OP

    source, target, mask = G.compositing_images((64, 80))
    field, _ = OP.clone_window(source, target, mask, margin=10**6)
    rhs = OP.clone_rhs(field, target)
    want, tgt, m2 = G.compositing_problem((64, 80))
    assert np.array_equal(m2, mask) and np.array_equal(np.moveaxis(tgt, 0, 2), target)
    assert np.array_equal(np.moveaxis(rhs, 2, 0), want)


def _disc(shape, radius, centre=None):
    r, c = np.mgrid[0:shape[0], 0:shape[1]]
    cy, cx = centre if centre else (shape[0] / 2, shape[1] / 2)
    return (r - cy) ** 2 + (c - cx) ** 2 < radius ** 2


def test_masked_solve_over_the_whole_interior_is_the_rectangle_solve():
    from oracle import poisson as OP

    rng = np.random.default_rng(21)
    img = rng.random((30, 41, 3))
    field = OP.GuidanceField(rng.standard_normal((30, 41, 3)), rng.standard_normal((30, 41, 3)))
    masked, _ = OP.solve_masked(field, img, np.ones((30, 41), dtype=bool))
    rect, _ = OP.solve_dirichlet(field, img, method="direct")
    assert np.abs(masked - rect).max() < 1e-10


def test_masked_solve_satisfies_the_poisson_equation_on_the_mask():
    # L(result) = div(field) at every unknown; every other pixel keeps its value
    from oracle import poisson as OP
    from oracle.operators import divergence

    rng = np.random.default_rng(22)
    img = rng.random((40, 52))
    field = OP.GuidanceField(rng.standard_normal((40, 52)), rng.standard_normal((40, 52)))
    mask = _disc((40, 52), 15.0) | _disc((40, 52), 6.0, (3, 4))  # touches the border
    out, _ = OP.solve_masked(field, img, mask)
    inner = np.zeros_like(mask)
    inner[1:-1, 1:-1] = mask[1:-1, 1:-1]
    p = np.pad(out, 1)
    lap = p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:] - 4 * out
    assert np.abs(lap - divergence(field.gx, field.gy))[inner].max() < 1e-10
    assert np.array_equal(out[~inner], img[~inner])


@pytest.mark.parametrize("mode", ["import", "mixed", "average"])
def test_mask_route_cloning_the_target_into_itself_changes_nothing(mode):
    from oracle import poisson as OP

    tgt = np.random.default_rng(23).random((32, 36, 3))
    mask = _disc((20, 24), 8.0)
    out, _ = OP.seamless_clone(tgt[5:25, 6:30], tgt, mask, (5, 6), mode=mode, domain="mask")
    assert np.abs(out - tgt).max() < 1e-10


# ---------------------------------------------------------------- texture_flatten /
# illumination_change (poisson.py:351-437)


def _texture_example(shape=(96, 128), seed=1):
    """Piecewise-constant RGB regions plus a fine texture; edges marks the region borders."""
    rng = np.random.default_rng(seed)
    rows, cols = np.mgrid[0:shape[0], 0:shape[1]]
    regions = (rows > shape[0] // 2).astype(int) + 2 * (cols > shape[1] // 3).astype(int)
    colours = rng.random((4, 3))
    image = colours[regions] + 0.05 * np.sin(1.7 * cols + 0.9 * rows)[..., None]
    edges = np.zeros(shape, dtype=bool)
    edges[:, :-1] |= regions[:, 1:] != regions[:, :-1]
    edges[:-1, :] |= regions[1:, :] != regions[:-1, :]
    return image, edges.astype(float)


def test_texture_flatten_with_unit_edges_returns_the_image():
    """test_applications.py:150-152: weights of one keep the image's own field."""
    image, _ = _texture_example((40, 52))
    out, _ = OP.texture_flatten(image, np.ones(image.shape[:2]), method="direct")
    assert np.abs(out - image).max() < 1e-9
    mask = np.zeros(image.shape[:2], dtype=bool)
    mask[5:30, 8:40] = True
    out, _ = OP.texture_flatten(image, np.ones(image.shape[:2]), mask=mask)
    assert np.abs(out - image).max() < 1e-9


def test_texture_flatten_with_zero_edges_is_harmonic():
    """A zero field leaves the interior harmonic: the discrete Laplacian of the result
    vanishes at every pixel off the fixed ring."""
    image, _ = _texture_example((30, 44))
    out, _ = OP.texture_flatten(image, np.zeros(image.shape[:2]), method="direct")
    lap = O.laplacian(out)
    assert np.abs(lap[1:-1, 1:-1]).max() < 1e-9
    assert np.array_equal(out[0], image[0]) and np.array_equal(out[:, -1], image[:, -1])


def test_texture_flatten_reduces_texture_energy():
    """test_applications.py:155-164 on this build's texture example."""
    image, edges = _texture_example()
    flat, _ = OP.texture_flatten(image, edges, method="direct")

    def energy(x):
        interior = ~edges.astype(bool)
        variation = np.abs(np.diff(x, axis=1)).mean(axis=2)
        return float(variation[interior[:, :-1]].mean())

    assert energy(flat) < 0.2 * energy(image)


def test_illumination_change_with_beta_zero_returns_the_image():
    """beta = 0 makes the factor alpha ** 0 * g ** 0 = 1: the image's own field."""
    rng = np.random.default_rng(3)
    image = rng.random((32, 40, 3))
    mask = np.zeros((32, 40), dtype=bool)
    mask[4:28, 6:30] = True
    out, _ = OP.illumination_change(image, mask, alpha=0.3, beta=0.0)
    assert np.abs(out - image).max() < 1e-9


def test_illumination_change_lifts_a_dark_selection():
    """test_applications.py:167-175."""
    rows, cols = np.mgrid[0:96, 0:96]
    image = np.repeat((0.05 + 0.02 * np.sin(cols * 0.4))[..., None], 3, axis=2)
    mask = (rows - 48) ** 2 + (cols - 48) ** 2 < 30.0 ** 2
    inner = (rows - 48) ** 2 + (cols - 48) ** 2 < 24.0 ** 2
    out, _ = OP.illumination_change(image, mask, alpha=0.4, beta=0.6)
    ring = mask & ~inner
    assert out[ring].std() > image[ring].std()
    assert np.array_equal(out[~mask], image[~mask])
