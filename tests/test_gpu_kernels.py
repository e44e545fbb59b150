"""GPU parity: HIP kernels (through the C ABI) vs the FP64 oracle on seeded inputs.

Tolerances: operators and one V-cycle agree with the oracle to round-off (they do
the same FP64 arithmetic in a different summation order); solves agree to the
solver tolerance.  Shapes are the reference tests' (gradient-domain/tests/).
"""

import numpy as np
import pytest

from oracle import generators as G
from oracle import multigrid as OM
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


@pytest.mark.parametrize("shape", [(23, 31), (64, 64), (1, 7), (130, 257)])
@pytest.mark.parametrize("spacing", [1.0, 0.5, 2.0])
def test_laplacians_match_oracle(gd, shape, spacing):
    rng = np.random.default_rng(11)
    u = rng.standard_normal(shape)
    assert rel(gd.laplacian(u, spacing), OO.laplacian(u, spacing)) < 1e-14
    assert rel(gd.neumann_laplacian(u, spacing), OO.neumann_laplacian(u, spacing)) < 1e-14
    u3 = rng.standard_normal(shape + (3,))
    assert rel(gd.laplacian(u3, spacing), OO.laplacian(u3, spacing)) < 1e-14


@pytest.mark.parametrize("shape", [(23, 31), (17, 21, 3), (1, 5), (40, 1)])
def test_gradient_divergence_match_oracle(gd, shape):
    rng = np.random.default_rng(12)
    u = rng.standard_normal(shape)
    gx, gy = gd.gradient(u)
    ox, oy = OO.gradient(u)
    assert np.array_equal(gx, ox) and np.array_equal(gy, oy)
    vx, vy = rng.standard_normal(shape), rng.standard_normal(shape)
    assert np.array_equal(gd.divergence(vx, vy), OO.divergence(vx, vy))


def test_fold_boundary_matches_oracle(gd):
    rng = np.random.default_rng(13)
    full = rng.standard_normal((14, 18, 3))
    rhs = rng.standard_normal((12, 16, 3))
    for h in (1.0, 0.5):
        assert np.array_equal(gd.fold_boundary(rhs, full, h), OO.fold_boundary(rhs, full, h))
    f = gd.fold_boundary(np.zeros((3, 3)), np.ones((5, 5)), spacing=0.5)
    assert f[0, 0] == pytest.approx(-8.0) and f[0, 1] == pytest.approx(-4.0)
    assert f[1, 1] == pytest.approx(0.0)


@pytest.mark.parametrize("bc", ["dirichlet", "neumann"])
@pytest.mark.parametrize("shape", [(31, 25), (64, 64), (33, 40)])
def test_smoother_matches_oracle(gd, bc, shape):
    from gdpoisson.multigrid import _smooth

    rng = np.random.default_rng(14)
    u = rng.standard_normal(shape)
    b = rng.standard_normal(shape)
    for sweeps in (1, 3):
        for order in (0, 1):
            ord_ = (0, 1) if order == 0 else (1, 0)
            got = _smooth(u, b, 0.5, sweeps, bc, order)
            want = OM.smooth(u.copy(), b, 0.5, sweeps, bc, ord_)
            assert rel(got, want) < 1e-13


@pytest.mark.parametrize("bc", ["dirichlet", "neumann"])
@pytest.mark.parametrize("shape", [(31, 25), (30, 24), (64, 65), (7, 9)])
def test_transfers_match_oracle(gd, bc, shape):
    from gdpoisson.multigrid import _prolong, _restrict

    rng = np.random.default_rng(6)
    f = rng.standard_normal(shape)
    cs = OM.coarse_shape(shape, bc)
    c = rng.standard_normal(cs)
    assert rel(_restrict(f, bc), OM.restrict(f, bc)) < 1e-14
    assert rel(_prolong(c, shape, bc), OM.prolong(c, shape, bc)) < 1e-14


@pytest.mark.parametrize("shape", [(63, 63), (47, 61), (127, 127), (255, 255), (31, 25)])
@pytest.mark.parametrize("nu", [1, 2])
def test_dirichlet_vcycle_matches_oracle(gd, shape, nu):
    rng = np.random.default_rng(15)
    b, _ = G.manufactured_problem(shape, seed=5)
    u = rng.standard_normal(shape)
    opts = gd.MultigridOptions(pre_smoothing=nu, post_smoothing=nu)
    got = gd.v_cycle(u, b, 1.0, opts)
    want = OM.v_cycle(u, b, 1.0, OM.MultigridOptions(pre_smoothing=nu, post_smoothing=nu))
    assert rel(got, want) < 1e-11


@pytest.mark.parametrize("shape", [(63, 63), (254, 256), (127, 200)])
@pytest.mark.parametrize("nu", [1, 2])
def test_dirichlet_exact_coarsest_vcycle_matches_oracle(gd, shape, nu):
    """coarsest_sweeps = 0: exact DST-I coarsest solve (the mixed solver's cycle)."""
    rng = np.random.default_rng(16)
    b, _ = G.manufactured_problem(shape, seed=6)
    u = rng.standard_normal(shape)
    kw = dict(pre_smoothing=nu, post_smoothing=nu, coarsest_size=32, coarsest_sweeps=0)
    got = gd.v_cycle(u, b, 1.0, gd.MultigridOptions(**kw))
    want = OM.v_cycle(u, b, 1.0, OM.MultigridOptions(**kw))
    assert rel(got, want) < 1e-11
    got32 = gd.v_cycle(u, b, 1.0, gd.MultigridOptions(**kw), precision="mixed")
    assert rel(got32, want) < 1e-4


@pytest.mark.parametrize("bc", ["dirichlet", "neumann"])
@pytest.mark.parametrize("nu", [0, 1, 2, 3])
def test_vcycle_across_many_strips_matches_oracle(gd, bc, nu):
    """A grid several strips wide and tall on every strip level, for each sweep count."""
    shape = (263, 391)
    rng = np.random.default_rng(40 + nu)
    b = rng.standard_normal(shape)
    u = rng.standard_normal(shape)
    kw = dict(pre_smoothing=nu, post_smoothing=nu, coarsest_size=16, coarsest_sweeps=0)
    if bc == "dirichlet":
        kw = dict(pre_smoothing=nu, post_smoothing=nu)
    got = gd.v_cycle(u, b, 1.0, gd.MultigridOptions(**kw), boundary=bc)
    want = OM.v_cycle(u, b, 1.0, OM.MultigridOptions(**kw), bc)
    assert rel(got, want) < 1e-11


@pytest.mark.parametrize("shape", [(64, 64), (255, 257), (135, 240), (300, 200)])
def test_neumann_vcycle_matches_oracle(gd, shape):
    b, _ = G.neumann_problem(shape, seed=17)
    u = np.zeros(shape)
    opts = gd.MultigridOptions(pre_smoothing=1, post_smoothing=1, coarsest_size=16)
    got = gd.v_cycle(u, b, 1.0, opts, boundary="neumann")
    want = OM.v_cycle(u, b, 1.0, OM.MultigridOptions(pre_smoothing=1, post_smoothing=1,
                                                    coarsest_size=16), "neumann")
    assert rel(got, want) < 1e-11


@pytest.mark.parametrize("size", [63, 127, 255, 511])
def test_mgcg_reproduces_published_benchmark(gd, size):
    """docs/benchmark.json (seed 0, tol 1e-8): mgcg takes 7 iterations at 63..1023."""
    b, x = G.manufactured_problem((size, size), seed=0)
    u, rep = gd.solve_system(b, method="mgcg", tolerance=1e-8)
    ou, orep = OS.solve_system(b, method="mgcg", tolerance=1e-8)
    assert rep.iterations == orep.iterations == 7
    assert rep.converged
    assert rep.relative_residual == pytest.approx(orep.relative_residual, rel=1e-6)
    assert np.abs(u - ou).max() < 1e-12
    assert np.abs(u - x).max() < 1e-6


@pytest.mark.parametrize("method", ["cg", "multigrid", "mgcg", "mgcg_mixed"])
def test_every_solver_recovers_the_manufactured_solution(gd, method):
    b, expected = G.manufactured_problem((63, 63), seed=3)
    u, rep = gd.solve_system(b, method=method, tolerance=1e-10)
    assert rep.converged
    assert rep.relative_residual < 1e-8
    assert np.abs(u - expected).max() < 1e-6


def test_solvers_agree_with_direct(gd):
    b, _ = G.manufactured_problem((47, 61), seed=4)
    ref, _ = OS.solve_system(b, method="direct")
    for method in ("cg", "multigrid", "mgcg", "mgcg_mixed"):
        u, _ = gd.solve_system(b, method=method, tolerance=1e-11)
        assert np.abs(u - ref).max() < 1e-6


def test_unknown_method_and_zero_rhs(gd):
    with pytest.raises(KeyError):
        gd.solve_system(np.zeros((8, 8)), method="jacobi")
    u, rep = gd.solve_system(np.zeros((16, 16)), method="mgcg")
    assert np.all(u == 0.0) and rep.relative_residual == 0.0


@pytest.mark.parametrize("size", [31, 63, 127, 255])
def test_multigrid_cycles_match_oracle(gd, size):
    b, _ = G.manufactured_problem((size, size), seed=5)
    _, rep = gd.solve(b, options=gd.MultigridOptions(tolerance=1e-10))
    _, orep = OM.solve(b, options=OM.MultigridOptions(tolerance=1e-10))
    assert rep.converged and rep.cycles == orep.cycles <= 10
    assert rep.convergence_factor < 0.1
    assert np.allclose(rep.residuals, orep.residuals, rtol=1e-6, atol=0)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("shape", [(256, 256), (255, 257), (135, 240)])
def test_neumann_mgcg_matches_closed_form(gd, precision, shape):
    b, exact = G.neumann_problem(shape, seed=1001)
    method = "mgcg_mixed" if precision == "mixed" else "mgcg"
    u, rep = gd.solve_system(b, method=method, tolerance=1e-8, boundary="neumann")
    ou, oit, _ = OS.mgcg(b, bc="neumann", precision=precision, flexible=precision == "mixed")
    assert rep.converged and rep.relative_residual <= 1e-8
    assert abs(rep.iterations - oit) <= 1
    assert abs(u.mean()) < 1e-10
    assert np.abs(u - exact).max() < 1e-6
    assert np.abs(u - ou).max() < 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("bc", ["dirichlet", "neumann"])
@pytest.mark.parametrize("width", [1, 2, 3, 126, 127, 128, 129, 253, 254])
def test_pcg_at_strip_boundary_widths(gd, width, bc, precision):
    """Widths at the edges of the Krylov apply's strips (126 owned columns, the first strip
    also owning column 0, neighbours overlapping by two) and a height that is no multiple
    of its 4-row strips: the solve recovers a random discrete solution."""
    from oracle.operators import divergence, gradient, laplacian

    shape = (9, width)
    x = np.random.default_rng(width).standard_normal(shape)
    b = laplacian(x) if bc == "dirichlet" else divergence(*gradient(x))
    method = "mgcg_mixed" if precision == "mixed" else "mgcg"
    u, rep = gd.solve_system(b, method=method, tolerance=1e-11, boundary=bc)
    assert rep.converged and rep.relative_residual <= 1e-11
    if bc == "neumann":
        u, x = u - u.mean(), x - x.mean()
    assert np.abs(u - x).max() < 1e-6


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("bc", ["dirichlet", "neumann"])
@pytest.mark.parametrize("shape", [(2, 2), (3, 5), (40, 128), (128, 40), (64, 128), (65, 129),
                                   (63, 127), (33, 128), (200, 9), (9, 200), (1, 300),
                                   (300, 1), (3, 1000)])
def test_pcg_on_ragged_shapes(gd, shape, bc, precision):
    """Shapes at the edges of the coarse tail's placement (register tail: at most 64 rows
    in FP32, 32 in FP64, 128 columns; the LDS tail otherwise), thin grids both ways and
    thin long grids whose exact coarsest solve does not fit (coarsest sweeps instead):
    the solve recovers a random discrete solution."""
    from oracle.operators import divergence, gradient, laplacian

    x = np.random.default_rng(shape[0] * 7919 + shape[1]).standard_normal(shape)
    b = laplacian(x) if bc == "dirichlet" else divergence(*gradient(x))
    method = "mgcg_mixed" if precision == "mixed" else "mgcg"
    u, rep = gd.solve_system(b, method=method, tolerance=1e-11, boundary=bc)
    assert rep.converged and rep.relative_residual <= 1e-11
    if bc == "neumann":
        u, x = u - u.mean(), x - x.mean()
    assert np.abs(u - x).max() < 1e-6


@pytest.mark.parametrize("shape", [(96, 130), (64, 64)])
def test_solve_neumann_matches_the_dct_oracle(gd, shape):
    """solvers.py:224-262 contract: zero-mean solution, constant part of b discarded."""
    b, _ = G.neumann_problem(shape, seed=31)
    b = b + 0.5  # a constant component, which the solve must discard
    u = gd.solve_neumann(b)
    assert abs(float(u.mean())) < 1e-12
    assert np.abs(u - OS.solve_neumann(b)).max() < 1e-9


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_uneven_problems_match_each_alone(gd, precision):
    """Uneven shapes in one plan (the batched configuration at test sizes): every problem
    gives bit-for-bit the result it gives alone, and meets the closed form."""
    import torch

    shapes = G.batch_shapes(count=9, seed=4000, low=24, high=300)
    bs, exact = [], []
    for k, s in enumerate(shapes):
        b, x = G.neumann_problem(s, seed=4100 + k)
        bs.append(torch.from_numpy(b).cuda())
        exact.append(x)
    opts = gd.PlanOptions(precision=precision)
    together, reps = gd.Plan(shapes, boundary="neumann", options=opts).mgcg(bs)
    for k, s in enumerate(shapes):
        alone, rep1 = gd.Plan([s], boundary="neumann", options=opts).mgcg([bs[k]])
        assert reps[k].converged and reps[k].relative_residual <= 1e-8
        assert reps[k].iterations == rep1[0].iterations
        assert torch.equal(together[k], alone[0]), (k, s)
        assert np.abs(together[k].cpu().numpy() - exact[k]).max() < 1e-6


@pytest.mark.parametrize("count,low,high", [(300, 24, 96), (200, 200, 300)])
def test_more_problems_than_free_cus_match_each_alone(gd, count, low, high):
    """More strip-tail problems than CUs (300 small grids), and 200 grids whose Krylov
    points are too many for the CUs the tail leaves free: the x half of the update is
    then not deferred to the tail launch, while each problem alone defers it.  Results
    are bit for bit the same either way, and every problem converges."""
    import torch

    shapes = G.batch_shapes(count=count, seed=4200 + count, low=low, high=high)
    bs, exact = [], []
    for k, s in enumerate(shapes):
        b, x = G.neumann_problem(s, seed=4300 + k)
        bs.append(torch.from_numpy(b).cuda())
        exact.append(x)
    opts = gd.PlanOptions(precision="mixed")
    together, reps = gd.Plan(shapes, boundary="neumann", options=opts).mgcg(bs)
    assert all(r.converged and r.relative_residual <= 1e-8 for r in reps)
    for k in (0, 1, count // 2, count - 1):
        alone, rep1 = gd.Plan([shapes[k]], boundary="neumann", options=opts).mgcg([bs[k]])
        assert reps[k].iterations == rep1[0].iterations
        assert torch.equal(together[k], alone[0]), (k, shapes[k])
        assert np.abs(together[k].cpu().numpy() - exact[k]).max() < 1e-6


@pytest.mark.parametrize("precision,flexible", [("fp64", None), ("fp64", True), ("mixed", False)])
def test_apply_carried_x_update_at_other_options(gd, precision, flexible):
    """The x update carried by the applies (a batch whose Krylov points the CUs left free
    by the tails cannot absorb), which add it to x every other iteration and hold it back
    otherwise, in FP64 (Fletcher-Reeves and flexible) and in mixed Fletcher-Reeves PCG:
    problems that finish after an odd and after an even number of iterations get bit for
    bit their solve alone (where the tail launch adds x every iteration)."""
    import torch

    shapes = G.batch_shapes(count=200, seed=4400, low=200, high=300)
    bs = [torch.from_numpy(G.neumann_problem(s, seed=4500 + k)[0]).cuda()
          for k, s in enumerate(shapes)]
    opts = gd.PlanOptions(precision=precision, flexible=flexible)
    together, reps = gd.Plan(shapes, boundary="neumann", options=opts).mgcg(bs)
    assert all(r.converged and r.relative_residual <= 1e-8 for r in reps)
    its = {r.iterations for r in reps}
    picks = [next(k for k, r in enumerate(reps) if r.iterations == i) for i in sorted(its)]
    for k in picks + [len(shapes) - 1]:
        alone, rep1 = gd.Plan([shapes[k]], boundary="neumann", options=opts).mgcg([bs[k]])
        assert reps[k].iterations == rep1[0].iterations
        assert torch.equal(together[k], alone[0]), (k, shapes[k])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_residual_replacement_below_the_rounding_floor(gd, precision):
    """Tolerances near and below the FP64 rounding floor, where the recursive residual can
    meet the target while the true one does not: the device check then reopens the problem
    from its true residual (at most 4 times per problem) and afterwards accepts it
    unconverged.  Every report carries the true residual of the solution it returns (the
    oracle's b - A u), converged says whether that met the tolerance, and each problem
    gets the same iterations and bits batched and alone."""
    import torch

    shapes = [(130, 70), (64, 64), (257, 129)]
    bs = [G.neumann_problem(s, seed=77 + k)[0] for k, s in enumerate(shapes)]
    for tol in (1e-13, 1e-15):
        opts = gd.PlanOptions(precision=precision, tolerance=tol, max_iterations=80)
        u, reps = gd.Plan(shapes, boundary="neumann", options=opts).mgcg(
            [torch.from_numpy(b).cuda() for b in bs])
        for k, b in enumerate(bs):
            true = OS.relative_residual(u[k].cpu().numpy(), b, 1.0, bc=OS.NEUMANN)
            # near the floor the two residual computations differ by their own rounding
            assert reps[k].relative_residual == pytest.approx(true, rel=0.05, abs=2e-15)
            assert reps[k].converged == (reps[k].relative_residual <= tol)
            assert reps[k].iterations <= 80
        assert any(not r.converged for r in reps) or tol > 1e-14
        for k, b in enumerate(bs):
            alone, rep1 = gd.Plan([shapes[k]], boundary="neumann", options=opts).mgcg(
                [torch.from_numpy(b).cuda()])
            assert rep1[0].iterations == reps[k].iterations, (tol, k)
            assert torch.equal(alone[0], u[k]), (tol, k)


@pytest.mark.parametrize("bc", ["dirichlet", "neumann"])
@pytest.mark.parametrize("menu", [0, 1, 2, 3])
def test_every_tile_menu_entry_matches_oracle(gd, bc, menu):
    """Each compiled strip height (down and up, every tiled level) gives the oracle's cycle."""
    import torch

    shape = (300, 260)
    rng = np.random.default_rng(50 + menu)
    b = rng.standard_normal(shape)
    u = rng.standard_normal(shape)
    kw = dict(pre_smoothing=2, post_smoothing=2)
    if bc == "neumann":
        kw.update(coarsest_size=16)
    tiles = {(k, l): menu for k in (0, 1) for l in range(4)}
    tiles[(2, 0)] = menu
    plan = gd.Plan([shape], boundary=bc, options=gd.PlanOptions(**kw), tiles=tiles)
    (got,) = plan.vcycle([torch.from_numpy(u).cuda()], [torch.from_numpy(b).cuda()])
    want = OM.v_cycle(u, b, 1.0, OM.MultigridOptions(**kw), bc)
    assert rel(got.cpu().numpy(), want) < 1e-11


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("bc", ["dirichlet", "neumann"])
@pytest.mark.parametrize("nu", [1, 2, 3])
@pytest.mark.parametrize("stream", [4, 5, 6, 7, 8, 9])
@pytest.mark.parametrize("shape", [(300, 260), (97, 131), (1030, 250)])
def test_streamed_down_kernel_is_bit_identical(gd, shape, stream, nu, bc, precision):
    """The wavefront-order down kernel (menu entries 4, 5, 6: 8, 16 and 32 coarse rows per
    strip; 32 falls back to 16 for nu = 3) and the block-tiled one (entries 7, 8, 9: 8, 16
    and 32 coarse rows per block, rows shared by its waves through LDS) give the same bits
    as the sweep-order kernel (entry 3) on every tiled level, and the oracle's V-cycle to
    round-off; ragged and tall shapes put strips and blocks across the grid's edges and
    past its last rows."""
    import torch

    rng = np.random.default_rng(60 + nu + stream)
    b = rng.standard_normal(shape)
    u = rng.standard_normal(shape)
    kw = dict(pre_smoothing=nu, post_smoothing=nu, precision=precision)
    if bc == "neumann":
        kw.update(coarsest_size=16)
    outs = {}
    for menu in (3, stream):
        tiles = {(0, l): menu for l in range(6)}
        plan = gd.Plan([shape], boundary=bc, options=gd.PlanOptions(**kw), tiles=tiles)
        (got,) = plan.vcycle([torch.from_numpy(u).cuda()], [torch.from_numpy(b).cuda()])
        outs[menu] = got
    assert torch.equal(outs[3], outs[stream])
    if precision == "fp64":
        okw = {k: v for k, v in kw.items() if k != "precision"}
        want = OM.v_cycle(u, b, 1.0, OM.MultigridOptions(**okw), bc)
        assert rel(outs[stream].cpu().numpy(), want) < 1e-11


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("bc", ["dirichlet", "neumann"])
@pytest.mark.parametrize("nu", [1, 2, 3])
@pytest.mark.parametrize("up", [4, 5])
@pytest.mark.parametrize("shape", [(300, 260), (97, 131), (1030, 250)])
def test_block_up_kernel_is_bit_identical(gd, shape, up, nu, bc, precision):
    """The block-tiled up kernel (menu entries 4, 5: 16 and 32 fine rows per block, rows
    shared by its waves through LDS) gives the same bits as the strip kernel (entry 3) on
    every tiled level, and the oracle's V-cycle to round-off; a whole solve with it (the
    finest level's dots summed over its blocks, in another order) agrees to round-off."""
    import torch

    rng = np.random.default_rng(80 + nu + up)
    b = rng.standard_normal(shape)
    u = rng.standard_normal(shape)
    kw = dict(pre_smoothing=nu, post_smoothing=nu, precision=precision)
    if bc == "neumann":
        kw.update(coarsest_size=16)
        b = b - b.mean()
    outs, sols = {}, {}
    for menu in (3, up):
        tiles = {(1, l): menu for l in range(6)}
        plan = gd.Plan([shape], boundary=bc, options=gd.PlanOptions(**kw), tiles=tiles)
        bt = torch.from_numpy(b).cuda()
        (outs[menu],) = plan.vcycle([torch.from_numpy(u).cuda()], [bt])
        (sols[menu],), (rep,) = plan.mgcg([bt])
        assert rep.converged
    assert torch.equal(outs[3], outs[up])
    assert rel(sols[up].cpu().numpy(), sols[3].cpu().numpy()) < (1e-10 if precision == "fp64"
                                                                   else 1e-6)
    if precision == "fp64":
        okw = {k: v for k, v in kw.items() if k != "precision"}
        want = OM.v_cycle(u, b, 1.0, OM.MultigridOptions(**okw), bc)
        assert rel(outs[up].cpu().numpy(), want) < 1e-11


@pytest.mark.parametrize("down", [5, 8])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_streamed_down_kernel_in_batched_solves(gd, precision, down):
    """Batched uneven problems solved with the streamed (entry 5) or block-tiled (entry 8)
    down kernel on every level: the same bits as with the sweep-order kernel, each problem
    alone and batched."""
    import torch

    shapes = G.batch_shapes(count=7, seed=4400, low=40, high=700)
    bs = []
    for k, s in enumerate(shapes):
        b, _ = G.neumann_problem(s, seed=4500 + k)
        bs.append(torch.from_numpy(b).cuda())
    opts = gd.PlanOptions(precision=precision)
    ref, _ = gd.Plan(shapes, boundary="neumann", options=opts,
                     tiles={(0, l): 3 for l in range(6)}).mgcg(bs)
    got, reps = gd.Plan(shapes, boundary="neumann", options=opts,
                        tiles={(0, l): down for l in range(6)}).mgcg(bs)
    assert all(r.converged and r.relative_residual <= 1e-8 for r in reps)
    for k in range(len(shapes)):
        assert torch.equal(ref[k], got[k]), (k, shapes[k])
    alone, _ = gd.Plan([shapes[3]], boundary="neumann", options=opts,
                       tiles={(0, l): down for l in range(6)}).mgcg([bs[3]])
    assert torch.equal(alone[0], got[3])


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("bc", ["dirichlet", "neumann"])
@pytest.mark.parametrize("fused", [10, 11, 12, 13, 14, 15])
@pytest.mark.parametrize("shapes", [[(1024, 1024)] * 3, [(300, 260)],
                                    [(300, 260), (1030, 250), (257, 255)]])
def test_fused_update_down_kernel_gives_the_same_solve(gd, shapes, fused, bc, precision):
    """Down entries 10 .. 15 carry the PCG update in the finest down kernel (the update's
    right-hand side formed from r and q, its convergence test at the tail kernels' entry,
    r stored by the finest up kernel): the same iterations and the same solution
    bits as the streamed strips and blocks of entries 4 .. 9 with k_update, alone and
    batched; the profile shows the fused kernel ran."""
    import torch

    bs = []
    for k, s in enumerate(shapes):
        if bc == "neumann":
            b, _ = G.neumann_problem(s, seed=4700 + k)
        else:
            b = np.random.default_rng(4800 + k).standard_normal(s)
        bs.append(torch.from_numpy(b).cuda())
    opts = gd.PlanOptions(precision=precision)
    outs, reps = {}, {}
    for menu in (fused - 6, fused):
        plan = gd.Plan(shapes, boundary=bc, options=opts, tiles={(0, 0): menu})
        outs[menu], reps[menu] = plan.mgcg(bs)
        plan.profile(True)
        plan.mgcg(bs)
        kinds = set(plan.profile_report())
        plan.profile(False)
        assert ("down L0+update" in kinds) == (menu == fused), kinds
        assert ("update (x deferred)" in kinds) == (menu != fused), kinds
    for k in range(len(shapes)):
        a, f = reps[fused - 6][k], reps[fused][k]
        assert f.converged and f.relative_residual <= 1e-8
        assert a.iterations == f.iterations and a.relative_residual == f.relative_residual
        assert torch.equal(outs[fused - 6][k], outs[fused][k]), (k, shapes[k])
    if len(shapes) > 1:
        alone, _ = gd.Plan([shapes[1]], boundary=bc, options=opts,
                           tiles={(0, 0): fused}).mgcg([bs[1]])
        assert torch.equal(alone[0], outs[fused][1])


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
@pytest.mark.parametrize("nu", [1, 3])
@pytest.mark.parametrize("fused", [10, 12, 13, 15])
@pytest.mark.parametrize("bc", ["dirichlet", "neumann"])
def test_fused_update_down_kernel_at_other_sweep_counts(gd, bc, fused, nu, precision):
    """The fused update-down entries at nu = 1 and 3 (entry 12's 32-row strips run as
    16-row strips at nu = 3): the bits of the plain entries with k_update."""
    import torch

    shapes = [(300, 260), (513, 190)]
    bs = []
    for k, s in enumerate(shapes):
        if bc == "neumann":
            b, _ = G.neumann_problem(s, seed=5100 + k)
        else:
            b = np.random.default_rng(5200 + k).standard_normal(s)
        bs.append(torch.from_numpy(b).cuda())
    opts = gd.PlanOptions(precision=precision, pre_smoothing=nu, post_smoothing=nu)
    outs, reps = {}, {}
    for menu in (fused - 6, fused):
        outs[menu], reps[menu] = gd.Plan(shapes, boundary=bc, options=opts,
                                         tiles={(0, 0): menu}).mgcg(bs)
    for k in range(len(shapes)):
        assert reps[fused][k].converged
        assert reps[fused - 6][k].iterations == reps[fused][k].iterations
        assert torch.equal(outs[fused - 6][k], outs[fused][k]), (k, shapes[k])


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_fused_update_entry_without_two_tiled_levels(gd, precision):
    """A plan with a problem whose finest level is in the coarse tail keeps k_update
    (that problem has no finest down kernel to carry it): a fused entry then solves as
    the plain streamed strips."""
    import torch

    shapes = [(300, 260), (30, 100)]  # in the tail from its finest level at both precisions
    bs = [torch.from_numpy(G.neumann_problem(s, seed=4900 + k)[0]).cuda()
          for k, s in enumerate(shapes)]
    opts = gd.PlanOptions(precision=precision)
    ref, _ = gd.Plan(shapes, boundary="neumann", options=opts, tiles={(0, 0): 5}).mgcg(bs)
    plan = gd.Plan(shapes, boundary="neumann", options=opts, tiles={(0, 0): 11})
    got, reps = plan.mgcg(bs)
    plan.profile(True)
    plan.mgcg(bs)
    assert "down L0+update" not in plan.profile_report()
    plan.profile(False)
    assert all(r.converged for r in reps)
    for k in range(len(shapes)):
        assert torch.equal(ref[k], got[k])


@pytest.mark.parametrize("kry", range(6))
def test_every_krylov_tile_solves_like_the_oracle(gd, kry):
    """Each compiled Krylov tile shape (apply, update, set-up and check kernels) gives the
    oracle's FP64 MG-PCG: same iteration count, same solution to round-off."""
    import torch

    b, _ = G.manufactured_problem((300, 260), seed=70 + kry)
    plan = gd.Plan([b.shape], boundary="dirichlet", options=gd.PlanOptions(precision="fp64"),
                   tiles={(2, 0): kry})
    (u,), (rep,) = plan.mgcg([torch.from_numpy(b).cuda()])
    want, wrep = OS.solve_system(b, "mgcg")
    assert rep.iterations == wrep.iterations
    assert rel(u.cpu().numpy(), want) < 1e-10


def test_autotuned_plan_solves_the_same_problem(gd):
    import torch

    shape = (512, 384)
    bs, exact = [], []
    for c in range(3):
        b, x = G.neumann_problem(shape, seed=60 + c)
        bs.append(torch.from_numpy(b).cuda())
        exact.append(x)
    opts = gd.PlanOptions(precision="mixed")
    ref, rep0 = gd.Plan([shape] * 3, boundary="neumann", options=opts).mgcg(bs)
    plan = gd.Plan([shape] * 3, boundary="neumann", options=opts)
    choice = gd.autotune(plan, use_cache=False)
    assert set(k for k, _ in choice) <= {0, 1, 2}
    from gdpoisson.autotune import menu_size

    sizes = {k: menu_size(k) for k in (0, 1, 2)}
    assert sizes == {0: 16, 1: 6, 2: 6}
    assert all(0 <= v < sizes[k] for (k, _), v in choice.items())
    got, rep = plan.mgcg(bs)
    for c in range(3):
        assert rep[c].converged and rep[c].relative_residual <= 1e-8
        assert abs(rep[c].iterations - rep0[c].iterations) <= 1
        assert np.abs(got[c].cpu().numpy() - exact[c]).max() < 1e-6
        assert torch.allclose(got[c], ref[c], rtol=0, atol=1e-7)


# ---------------------------------------------------------------- compositing front end
def _clone_example(shape=(120, 150)):
    rng = np.random.default_rng(70)
    target = rng.random(shape + (3,))
    source = rng.random((50, 64, 3)) * 0.5 + 0.25
    mask = np.zeros((50, 64), dtype=bool)
    yy, xx = np.mgrid[0:50, 0:64]
    mask[(yy - 25) ** 2 + (xx - 32) ** 2 < 20 ** 2] = True
    return source, target, mask, (30, 40)


@pytest.mark.parametrize("mode", ["import", "mixed", "average"])
def test_clone_rhs_is_bit_exact(gd, mode):
    import ctypes

    import torch

    from gdpoisson import _lib
    from oracle import poisson as OP

    source, target, mask, off = _clone_example()
    field, win = OP.clone_window(source, target, mask, off, mode, margin=9)
    want = OP.clone_rhs(OP.GuidanceField(field.gx[win], field.gy[win]), target[win])
    ps = torch.from_numpy(np.moveaxis(source, 2, 0).copy()).cuda()
    pt = torch.from_numpy(np.moveaxis(target, 2, 0).copy()).cuda()
    m = torch.from_numpy(mask.astype(np.uint8)).cuda()
    top, left = win[0].start, win[1].start
    hh, ww = win[0].stop - top, win[1].stop - left
    out = torch.empty((3, hh - 2, ww - 2), dtype=torch.float64, device="cuda")
    rc = _lib.lib().gdp_clone_rhs(
        ctypes.c_void_p(ps.data_ptr()), 50, 64, off[0], off[1], ctypes.c_void_p(m.data_ptr()),
        ctypes.c_void_p(pt.data_ptr()), 120, 150, 3, top, left, hh, ww,
        {"import": 0, "mixed": 1, "average": 2}[mode], ctypes.c_void_p(out.data_ptr()), None)
    assert rc == 0
    assert np.array_equal(np.moveaxis(out.cpu().numpy(), 0, 2), want)


def test_guidance_field_operations_match_the_oracle(gd):
    from oracle import poisson as OP

    rng = np.random.default_rng(71)
    a, b = rng.standard_normal((30, 40, 3)), rng.standard_normal((30, 40, 3))
    mask = rng.random((30, 40)) > 0.5
    factor = rng.random((30, 40))
    ga, gb = gd.GuidanceField.of(a), gd.GuidanceField.of(b)
    oa, ob = OP.GuidanceField.of(a), OP.GuidanceField.of(b)
    for got, want in ((ga.select(mask, gb), oa.select(mask, ob)),
                      (ga.mix_by_magnitude(gb), oa.mix_by_magnitude(ob)),
                      (ga.scaled(factor), oa.scaled(factor)), (ga.scaled(0.5), oa.scaled(0.5))):
        assert np.array_equal(got.gx, want.gx) and np.array_equal(got.gy, want.gy)


def test_integrating_an_image_reproduces_it_on_the_gpu(gd):
    """test_applications.py:49-55 on the GPU path."""
    image = np.random.default_rng(0).random((40, 52, 3))
    result, reports = gd.solve_dirichlet(gd.GuidanceField.of(image), image, tolerance=1e-12)
    assert np.abs(result - image).max() < 1e-9
    assert len(reports) == 3 and all(r.converged for r in reports)


@pytest.mark.parametrize("method", ["mgcg", "mgcg_mixed"])
@pytest.mark.parametrize("mode", ["import", "mixed"])
def test_seamless_clone_matches_the_oracle(gd, method, mode):
    from oracle import poisson as OP

    source, target, mask, off = _clone_example()
    got, reps = gd.seamless_clone(source, target, mask, off, mode=mode, method=method, margin=9,
                                  domain="rectangle")
    want, _ = OP.seamless_clone(source, target, mask, off, mode=mode, margin=9, domain="rectangle")
    assert all(r.converged and r.relative_residual <= 1e-8 for r in reps)
    assert np.abs(got - want).max() < 1e-6
    # outside the window the target is untouched
    outside = np.ones(target.shape[:2], dtype=bool)
    outside[20:91, 31:113] = False
    assert np.array_equal(got[outside], target[outside])
    with pytest.raises(ValueError):
        gd.seamless_clone(source, target, mask, off, domain="lasso")
    with pytest.raises(ValueError):
        gd.seamless_clone(source, target, mask, off, mode="unknown")


def test_compositing_config_at_full_size(gd):
    """2048 x 2048 RGB compositing (blob mask, FP32 MG + FP64 refinement) agrees with the
    FP64 GPU solve of the same window."""
    import torch

    source, target, mask = G.compositing_images((2048, 2048))
    s, t = torch.from_numpy(source).cuda(), torch.from_numpy(target).cuda()
    mixed, reps = gd.seamless_clone(s, t, mask, (0, 0), method="mgcg_mixed", domain="rectangle")
    fp64, _ = gd.seamless_clone(s, t, mask, (0, 0), method="mgcg", domain="rectangle")
    assert all(r.converged and r.relative_residual <= 1e-8 for r in reps)
    assert torch.abs(mixed - fp64).max().item() < 1e-6


def test_batched_64_problem_workload(gd):
    """The batched configuration at full size: 64 RGB Neumann problems with sides
    log-uniform in [128, 2048] and aspect U[0.5, 2] (192 channels, about 145 M unknowns)
    in one mixed-precision plan.  Every channel reaches a true relative residual of 1e-8
    and its closed form to 1e-6; sampled problems are bit for bit the same solved alone."""
    import time

    import torch

    from gdpoisson import workloads as W

    shapes = W.batch_shapes()
    assert len(shapes) == 64
    bs, targets, plane_shapes = [], [], []
    for p, s in enumerate(shapes):
        for c in range(3):
            t = W.blob_image(s, 4000 + 10 * p + c)
            bs.append(W.neumann_rhs(t))
            targets.append(t - t.mean())
            plane_shapes.append(s)
    opts = gd.PlanOptions(precision="mixed")
    plan = gd.Plan(plane_shapes, boundary="neumann", options=opts)
    plan.mgcg(bs)  # first use: graph capture, tables
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    us, reps = plan.mgcg(bs)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    mpix = sum(a * b for a, b in shapes) / 1e6
    print(f"batched 64 RGB problems: {mpix:.1f} MPix in {dt * 1e3:.1f} ms = "
          f"{mpix / dt:.0f} MPix/s, iterations {min(r.iterations for r in reps)}.."
          f"{max(r.iterations for r in reps)}")
    assert all(r.converged and r.relative_residual <= 1e-8 for r in reps)
    err = max(float(np.abs(u.cpu().numpy() - t).max()) for u, t in zip(us, targets))
    assert err < 1e-6
    for p in (0, 17, 40, 63):
        alone = gd.Plan([shapes[p]] * 3, boundary="neumann", options=opts)
        ua, _ = alone.mgcg(bs[3 * p:3 * p + 3])
        for c in range(3):
            assert torch.equal(ua[c], us[3 * p + c])
