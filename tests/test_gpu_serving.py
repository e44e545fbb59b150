"""The serving loop on the device: every request's result is bit-for-bit the same alone
and in the mix (batched with others in one plan, overlapped with other batches on other
streams), served results agree with the oracle, and a timed stream reports latency."""

import numpy as np
import pytest

from oracle import generators as G
from oracle import hdr as H
from oracle import poisson as OP

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gd():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import gdpoisson

    return gdpoisson


def _requests():
    """Six compositing requests (squares of two sizes cut from 256 x 256 blob images and a
    blob mask; two by the mask route) and four small tone-mapping requests (two of one
    shape), all arriving at once: requests of one kind and shape batch together."""
    import torch

    from gdpoisson.serving import Request

    src, tgt, mask = G.compositing_images((256, 256), config=5)
    s, t, m = (torch.from_numpy(x).cuda() for x in (src, tgt, mask))
    rng = np.random.default_rng(3)
    reqs = []
    for k, side in enumerate([96, 96, 128, 96, 128, 128]):
        r0, c0 = (int(v) for v in rng.integers(0, 256 - side + 1, 2))
        win = (slice(r0, r0 + side), slice(c0, c0 + side))
        reqs.append(Request("composite", {"source": s[win], "target": t[win], "mask": m[win],
                                          "offset": (0, 0), "mode": ["import", "mixed"][k % 2],
                                          "domain": "mask" if k % 3 == 2 else "rectangle"}))
    for k, shape in enumerate([(96, 128), (120, 90), (96, 128), (64, 200)]):
        reqs.append(Request("tonemap", {
            "radiance": torch.from_numpy(G.radiance_map(shape, seed=40 + k)).cuda()}))
    return reqs


def test_results_are_the_same_alone_and_in_the_mix(gd):
    import torch

    from gdpoisson.serving import Request, Server

    reqs = _requests()
    srv = Server(budget_bytes=4 << 30, slots=3, max_batch=12)
    mixed = srv.run(reqs, realtime=False)
    assert any(n > 1 for _, n, _ in srv.batches), "nothing was batched"
    assert all(c.error is None for c in mixed)
    for r, c in zip(reqs, mixed):
        assert all(rep.converged and rep.relative_residual <= 1e-8 for rep in c.reports)
        alone = Server(budget_bytes=4 << 30, slots=1).run(
            [Request(r.kind, r.args)], realtime=False)[0]
        assert torch.equal(alone.output, c.output), (r.kind, c.rid)


def test_served_results_match_the_oracle(gd):
    from gdpoisson.serving import Server

    reqs = _requests()
    done = Server(budget_bytes=4 << 30, slots=2, max_batch=12).run(reqs, realtime=False)
    for r, c in zip(reqs, done):
        a = r.args
        if r.kind == "composite":
            want, _ = OP.seamless_clone(a["source"].cpu().numpy(), a["target"].cpu().numpy(),
                                        a["mask"].cpu().numpy(), (0, 0), mode=a["mode"],
                                        domain=a["domain"])
        else:
            want, _ = H.tone_map(a["radiance"].cpu().numpy())
        assert np.abs(c.output.cpu().numpy() - want).max() < 1e-6, (r.kind, c.rid)


def test_a_timed_stream_reports_latency(gd):
    from gdpoisson import workloads as W
    from gdpoisson.serving import Server, latency_stats

    reqs = W.serving_stream(count=16, qps=200.0, seed=7, sides=(128, 384),
                            tonemap_shape=(270, 480), base_side=512)
    done = Server(slots=3).run(reqs)
    st = latency_stats(done)
    assert st["served"] == 16 and st["rejected"] == 0
    assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert all(rep.converged for c in done for rep in c.reports)
    assert {c.kind for c in done} == {"composite", "tonemap"}


def test_resident_plans_are_reused_and_keep_the_bits(gd):
    """A second stream on the same server reuses the plans (and their captured graphs) the
    first built: no new plan for shapes seen before, and the same bits as the first time
    and as a fresh server; tuned tiles (chosen per shape) keep alone == in the mix."""
    import torch

    from gdpoisson.serving import Request, Server

    reqs = _requests()
    # one worker, so the second stream's batches meet the first stream's plans
    srv = Server(budget_bytes=4 << 30, slots=1, max_batch=12)
    first = srv.run(reqs, realtime=False)
    pooled1 = srv.plan_builds - srv.unpooled_builds
    unpooled1 = srv.unpooled_builds
    assert pooled1 > 0 and srv.plan_hits == 0
    again = srv.run([Request(r.kind, r.args) for r in reqs], realtime=False)
    st = srv.stats()
    # every rectangle plan of the second stream was resident (plans are per plane shape);
    # only the mask-route plans (their masks change per request) were built again
    assert st["graphs"] and st["plan_hits"] == pooled1
    assert srv.plan_builds - srv.unpooled_builds == pooled1
    assert srv.unpooled_builds == 2 * unpooled1
    for a, b in zip(first, again):
        assert torch.equal(a.output, b.output)
    tuned = Server(budget_bytes=4 << 30, slots=2, max_batch=12)
    table = tuned.tune(reqs, solves=1)
    assert table, "no plane shape was tuned"
    mixed = tuned.run(reqs, realtime=False)
    for r, c in zip(reqs, mixed):
        alone = Server(budget_bytes=4 << 30, slots=1, tiles=table).run(
            [Request(r.kind, r.args)], realtime=False)[0]
        assert torch.equal(alone.output, c.output), (r.kind, c.rid)
        assert all(rep.converged and rep.relative_residual <= 1e-8 for rep in c.reports)
