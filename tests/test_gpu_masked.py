"""Masked-domain solves on the device (solve_masked and seamless_clone(domain="mask"),
gradient-domain/src/gradient_domain/poisson.py:154-244, 337-338) against the oracle's
SuperLU solve of the same system."""

import numpy as np
import pytest

from oracle import generators as G
from oracle import poisson as OP

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gd():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gdpoisson

    return gdpoisson


def _disc(shape, radius, centre=None):
    r, c = np.mgrid[0:shape[0], 0:shape[1]]
    cy, cx = centre if centre else (shape[0] / 2, shape[1] / 2)
    return (r - cy) ** 2 + (c - cx) ** 2 < radius ** 2


MASKS = {
    "disc": lambda s: _disc(s, min(s) / 3),
    "border": lambda s: _disc(s, min(s) / 3) | _disc(s, 9.0, (2, 3)),
    "blobs": lambda s: G.blob_mask(s, seed=31, sigma=6.0),
    "full": lambda s: np.ones(s, dtype=bool),
}


@pytest.mark.parametrize("kind", sorted(MASKS))
@pytest.mark.parametrize("method", ["mgcg", "mgcg_mixed"])
def test_solve_masked_matches_superlu(gd, kind, method):
    rng = np.random.default_rng(41)
    shape = (70, 90)
    img = rng.random(shape + (3,))
    gx, gy = rng.standard_normal(shape + (3,)) * 0.1, rng.standard_normal(shape + (3,)) * 0.1
    mask = MASKS[kind](shape)
    tol = 1e-12 if method == "mgcg" else 1e-8
    got, reps = gd.solve_masked(gd.GuidanceField(gx, gy), img, mask, method=method,
                                tolerance=tol)
    want, _ = OP.solve_masked(OP.GuidanceField(gx, gy), img, mask)
    assert len(reps) == 3 and all(r.converged and r.relative_residual <= tol for r in reps)
    assert np.abs(got - want).max() < (1e-9 if method == "mgcg" else 1e-6)
    inner = np.zeros(shape, dtype=bool)
    inner[1:-1, 1:-1] = mask[1:-1, 1:-1]
    assert np.array_equal(got[~inner], img[~inner])  # test_applications.py:65-75


@pytest.mark.parametrize("mode", ["import", "mixed", "average"])
def test_mask_route_clone_matches_the_oracle(gd, mode):
    src, tgt, _ = G.compositing_images((96, 112), config=3)
    mask = G.blob_mask((96, 112), seed=32, sigma=8.0)
    got, reps = gd.seamless_clone(src, tgt, mask, (0, 0), mode=mode, domain="mask",
                                  method="mgcg_mixed")
    want, _ = OP.seamless_clone(src, tgt, mask, (0, 0), mode=mode, domain="mask")
    assert all(r.converged and r.relative_residual <= 1e-8 for r in reps)
    assert np.abs(got - want).max() < 1e-6


def test_mask_route_with_an_offset_source(gd):
    rng = np.random.default_rng(33)
    tgt = rng.random((80, 100, 3))
    src = rng.random((30, 40, 3))
    mask = _disc((30, 40), 12.0)
    got, _ = gd.seamless_clone(src, tgt, mask, (25, 31), domain="mask", method="mgcg",
                               tolerance=1e-12)
    want, _ = OP.seamless_clone(src, tgt, mask, (25, 31), domain="mask")
    assert np.abs(got - want).max() < 1e-9


def test_masked_compositing_at_512(gd):
    """A blob-masked 512 x 512 RGB composite (about 30% of the pixels unknown) against
    SuperLU; the iteration count is reported."""
    src, tgt, mask = G.compositing_images((512, 512), config=3)
    got, reps = gd.seamless_clone(src, tgt, mask, (0, 0), domain="mask", method="mgcg_mixed")
    want, _ = OP.seamless_clone(src, tgt, mask, (0, 0), domain="mask")
    print("masked 512x512 iterations", [r.iterations for r in reps])
    assert all(r.converged and r.relative_residual <= 1e-8 for r in reps)
    assert np.abs(got - want).max() < 1e-6


def test_masked_compositing_at_full_size(gd):
    """2048 x 2048 RGB, blob mask: mixed and FP64 masked solves agree."""
    import torch

    src, tgt, mask = G.compositing_images((2048, 2048))
    s, t = torch.from_numpy(src).cuda(), torch.from_numpy(tgt).cuda()
    mixed, reps = gd.seamless_clone(s, t, mask, (0, 0), domain="mask", method="mgcg_mixed")
    fp64, reps64 = gd.seamless_clone(s, t, mask, (0, 0), domain="mask", method="mgcg")
    print("masked 2048x2048 iterations mixed", [r.iterations for r in reps],
          "fp64", [r.iterations for r in reps64])
    assert all(r.converged and r.relative_residual <= 1e-8 for r in reps + reps64)
    assert torch.abs(mixed - fp64).max().item() < 1e-6


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_masked_and_unmasked_problems_in_one_plan(gd, precision):
    """A plan holding masked and unmasked Dirichlet problems of different shapes: every
    masked problem is bit for bit its solve alone, and the unmasked one (run by the masked
    level kernels with every grid point an unknown) agrees with its solve by the unmasked
    kernels."""
    import torch

    from gdpoisson.plan import Plan, PlanOptions

    rng = np.random.default_rng(7)
    shapes = [(70, 90), (64, 64), (50, 130)]
    masks = [torch.from_numpy(MASKS["blobs"](shapes[0]).astype(np.uint8)).cuda(), None,
             torch.from_numpy(_disc(shapes[2], 20.0).astype(np.uint8)).cuda()]
    bs = []
    for s, mk in zip(shapes, masks):
        b = torch.from_numpy(rng.standard_normal(s)).cuda()
        bs.append(b * mk if mk is not None else b)  # no equation at the fixed points
    opts = PlanOptions(precision=precision)
    tol = opts.tolerance
    us, reps = Plan(shapes, "dirichlet", 1.0, opts, masks=masks).mgcg(bs)
    for k, (s, mk) in enumerate(zip(shapes, masks)):
        alone = Plan([s], "dirichlet", 1.0, opts, masks=[mk] if mk is not None else None)
        (ua,), (ra,) = alone.mgcg([bs[k]])
        assert reps[k].converged and reps[k].relative_residual <= tol
        assert ra.converged and ra.relative_residual <= tol
        if mk is not None:
            assert torch.equal(us[k], ua), k
            assert torch.count_nonzero(us[k][mk == 0]).item() == 0  # fixed points stay zero
        else:
            scale = ua.abs().max().item()
            assert (us[k] - ua).abs().max().item() <= 1e-6 * scale


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_masked_batch_with_the_apply_carried_x_update(gd, precision):
    """200 masked Dirichlet problems in one plan: too many Krylov points for the CUs the
    tails leave free, so the masked applies carry the x update (every other iteration,
    two directions at a time), where a problem alone has it added by its tail launch.
    Sampled problems are bit for bit their solve alone, and converge."""
    import torch

    from gdpoisson.plan import Plan, PlanOptions

    shapes = G.batch_shapes(count=200, seed=4600, low=200, high=300)
    rng = np.random.default_rng(47)
    masks, bs = [], []
    for k, s in enumerate(shapes):
        mk = G.blob_mask(s, seed=4700 + k, sigma=6.0) if k % 2 else _disc(s, min(s) / 3)
        masks.append(torch.from_numpy(mk.astype(np.uint8)).cuda())
        bs.append(torch.from_numpy(rng.standard_normal(s) * mk).cuda())
    opts = PlanOptions(precision=precision)
    us, reps = Plan(shapes, "dirichlet", 1.0, opts, masks=masks).mgcg(bs)
    assert all(r.converged and r.relative_residual <= opts.tolerance for r in reps)
    for k in (0, 1, 101, len(shapes) - 1):
        (ua,), (ra,) = Plan([shapes[k]], "dirichlet", 1.0, opts, masks=[masks[k]]).mgcg([bs[k]])
        assert ra.iterations == reps[k].iterations, k
        assert torch.equal(us[k], ua), k
        assert torch.count_nonzero(us[k][masks[k] == 0]).item() == 0


def test_masked_solve_of_an_empty_selection_keeps_the_image(gd):
    """No selected pixel off the border (poisson.py:183-184): the image comes back, no
    reports."""
    rng = np.random.default_rng(5)
    img = rng.random((20, 30, 3))
    mask = np.zeros((20, 30), dtype=bool)
    mask[0, :] = mask[:, -1] = True  # border pixels only
    field = gd.GuidanceField(np.zeros_like(img), np.zeros_like(img))
    got, reps = gd.solve_masked(field, img, mask)
    assert reps == [] and np.array_equal(got, img)


@pytest.mark.parametrize("method", ["mgcg", "mgcg_mixed"])
@pytest.mark.parametrize("where", [(1, 1), (10, 15), (18, 28)])
def test_masked_solve_of_a_single_pixel(gd, method, where):
    """One unknown (at the border's inner corner, inside, at the far inner corner): the
    value that satisfies its equation, as SuperLU gives it."""
    rng = np.random.default_rng(6)
    img = rng.random((20, 30))
    gx, gy = rng.standard_normal((20, 30)) * 0.1, rng.standard_normal((20, 30)) * 0.1
    mask = np.zeros((20, 30), dtype=bool)
    mask[where] = True
    got, reps = gd.solve_masked(gd.GuidanceField(gx, gy), img, mask, method=method)
    want, _ = OP.solve_masked(OP.GuidanceField(gx, gy), img, mask)
    assert len(reps) == 1 and reps[0].converged
    assert np.abs(got - want).max() < 1e-9
    assert np.array_equal(got[~mask], img[~mask])


@pytest.mark.parametrize("entry", [11, 14])
@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_fused_update_entries_fall_back_on_masked_plans(gd, entry, precision):
    """A masked plan runs the finest level with the masked strip kernels and a separate
    update whatever down entry it is given: the down entries that carry the update
    (streamed 11, block-tiled 14) give the bits of the tallest masked strips (entry 3)."""
    import torch

    shape = (300, 260)
    mask = torch.from_numpy(G.blob_mask(shape, seed=77, sigma=8.0).astype(np.uint8)).cuda()
    rng = np.random.default_rng(78)
    # a masked right-hand side is zero at the fixed points (their values live in it)
    bs = [torch.from_numpy(rng.standard_normal(shape)).cuda() * mask for _ in range(2)]
    opts = gd.PlanOptions(precision=precision)
    outs = {}
    for menu in (3, entry):
        plan = gd.Plan([shape] * 2, boundary="dirichlet", options=opts,
                       tiles={(0, 0): menu}, masks=[mask, mask])
        outs[menu], reps = plan.mgcg(bs)
        assert all(r.converged for r in reps)
        plan.profile(True)
        plan.mgcg(bs)
        assert "down L0+update" not in plan.profile_report()
        plan.profile(False)
    for k in range(2):
        assert torch.equal(outs[3][k], outs[entry][k])
