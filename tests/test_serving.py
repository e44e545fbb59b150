"""Scheduling of the serving loop (gdpoisson.serving) with a stand-in executor on the CPU:
admission by budget, slot limit, batching rules, continuous admission and rejection.
The device pipeline itself is covered by tests/test_gpu_serving.py."""

import threading
import time

import numpy as np
import pytest

from gdpoisson.serving import Request, Server, latency_stats, poisson_arrivals


class FakeDevice:
    """Sleeps each batch's longest service time; records concurrency and bytes."""

    def __init__(self):
        self.lock = threading.Lock()
        self.active = 0
        self.active_bytes = 0
        self.peak = 0
        self.peak_bytes = 0
        self.batches = []

    def __call__(self, batch):
        nbytes = sum(r.args["nbytes"] for r in batch)
        with self.lock:
            self.active += 1
            self.active_bytes += nbytes
            self.peak = max(self.peak, self.active)
            self.peak_bytes = max(self.peak_bytes, self.active_bytes)
            self.batches.append([r.rid for r in batch])
        time.sleep(max(r.args["service"] for r in batch))
        with self.lock:
            self.active -= 1
            self.active_bytes -= nbytes
        return [(r.rid, [], None) for r in batch]


def _req(kind, nbytes, service, arrival=0.0, pixels=100):
    return Request(kind, {"nbytes": nbytes, "pixels": pixels, "service": service}, arrival)


def test_requests_complete_within_budget_and_slots():
    rng = np.random.default_rng(0)
    arr = poisson_arrivals(40, qps=400.0, seed=1)
    reqs = [_req(["composite", "tonemap"][rng.integers(2)], int(rng.integers(5, 40)),
                 float(rng.uniform(0.002, 0.01)), float(a)) for a in arr]
    dev = FakeDevice()
    srv = Server(budget_bytes=100, slots=3, max_batch=6, executor=dev)
    done = srv.run(reqs)
    assert [c.rid for c in done] == list(range(40))
    assert all(c.error is None and c.output == c.rid for c in done)
    assert dev.peak <= 3 and dev.peak_bytes <= 100 and srv.peak_bytes <= 100
    assert all(c.admitted >= c.arrival - 1e-9 and c.finished >= c.admitted for c in done)
    for c, r in zip(done, reqs):
        assert c.latency >= r.args["service"] * 0.9


def test_batches_hold_one_kind_and_respect_max_batch():
    reqs = [_req("composite", 1, 0.005) for _ in range(5)] + \
        [_req("tonemap", 1, 0.005) for _ in range(9)]
    dev = FakeDevice()
    srv = Server(budget_bytes=1000, slots=1, max_batch=4, executor=dev)
    srv.run(reqs, realtime=False)
    kinds = {i: r.kind for i, r in enumerate(reqs)}
    for b in dev.batches:
        assert len({kinds[i] for i in b}) == 1
        probs = sum(3 if kinds[i] == "composite" else 1 for i in b)
        assert probs <= 4 or len(b) == 1
    # RGB compositing (3 problems) cannot pair under max_batch 4; tone maps group by 4
    assert sorted(len(b) for b in dev.batches if kinds[b[0]] == "tonemap") == [1, 4, 4]
    # the oldest waiting request always leads: batches start in request order
    assert [b[0] for b in dev.batches] == sorted(b[0] for b in dev.batches)


def test_requests_arriving_during_work_are_admitted_at_once():
    long = _req("composite", 10, 0.30, 0.0)
    short = _req("tonemap", 10, 0.02, 0.05)
    srv = Server(budget_bytes=100, slots=2, executor=FakeDevice())
    a, b = srv.run([long, short])
    assert b.admitted < 0.1 < a.finished
    assert b.finished < a.finished


def test_admission_waits_for_budget_not_for_the_stream_to_drain():
    # two requests that do not fit together: the second starts when the first ends
    first = _req("tonemap", 80, 0.05, 0.0)
    second = _req("tonemap", 80, 0.01, 0.0)
    third = _req("composite", 10, 0.01, 0.0)
    srv = Server(budget_bytes=100, slots=3, executor=FakeDevice())
    a, b, c = srv.run([first, second, third])
    assert b.admitted >= a.finished - 1e-3
    # the head of the queue (second) blocks younger requests until it fits
    assert c.admitted >= b.admitted - 1e-3


def test_a_request_larger_than_the_budget_is_rejected():
    reqs = [_req("tonemap", 500, 0.001), _req("tonemap", 10, 0.001)]
    done = Server(budget_bytes=100, slots=2, executor=FakeDevice()).run(reqs)
    assert done[0].error and done[1].error is None
    st = latency_stats(done)
    assert st["served"] == 1 and st["rejected"] == 1


def test_a_failing_batch_fails_its_requests_only():
    """An executor that raises (a plan that cannot be built, a shape the kernels do not
    support) gives error completions for that batch; the loop keeps serving the rest and
    hands the batch's budget back."""
    inner = FakeDevice()

    def flaky(batch):
        if any(r.args.get("poison") for r in batch):
            raise NotImplementedError("coarsest level does not fit in LDS")
        return inner(batch)

    reqs = [_req("tonemap", 60, 0.002, 0.0), _req("composite", 60, 0.002, 0.001),
            _req("tonemap", 60, 0.002, 0.002), _req("composite", 60, 0.002, 0.003)]
    reqs[1].args["poison"] = True
    srv = Server(budget_bytes=100, slots=1, max_batch=1, executor=flaky)
    done = srv.run(reqs)
    assert [c.rid for c in done] == [0, 1, 2, 3]
    assert "NotImplementedError" in done[1].error and done[1].output is None
    assert all(done[i].error is None and done[i].output == i for i in (0, 2, 3))
    # every request used the whole budget alone: the failed batch released its share
    assert srv.peak_bytes == 60


def test_latency_stats():
    from gdpoisson.serving import Completion

    cs = [Completion(i, "tonemap", 0.0, 0.0, (i + 1) * 1e-3, 0, 1, 10) for i in range(100)]
    st = latency_stats(cs)
    lat = (np.arange(100) + 1.0)
    assert st["p50_ms"] == pytest.approx(np.percentile(lat, 50))
    assert st["p99_ms"] == pytest.approx(np.percentile(lat, 99))
    assert st["requests_per_s"] == pytest.approx(100 / 0.1)


def test_unknown_kind_is_rejected():
    with pytest.raises(ValueError):
        Server(budget_bytes=10, executor=FakeDevice()).run([Request("blur", {"nbytes": 1})])


def test_workload_generators_match_the_oracle_generators():
    from gdpoisson import workloads as W
    from oracle import generators as G

    assert np.array_equal(W.blob_image((40, 56), 7), G.blob_image((40, 56), 7))
    assert np.array_equal(W.blob_mask((64, 48), 9), G.blob_mask((64, 48), 9))
    assert W.batch_shapes() == G.batch_shapes()
    assert np.array_equal(W.lognormal_radiance((6, 5), 3), G.lognormal_radiance((6, 5), 3))
    s1, t1, m1 = W.compositing_images((40, 30))
    s2, t2, m2 = G.compositing_images((40, 30))
    assert np.array_equal(s1, s2) and np.array_equal(t1, t2) and np.array_equal(m1, m2)


def test_batches_hold_one_image_shape():
    """Requests batch only with requests of the same kind and image shape (so the batch's
    planes repeat shapes a resident plan was built for); the oldest request still leads."""
    import numpy as np

    def comp(shape, arrival):
        return Request("composite", {"nbytes": 10, "pixels": 1, "service": 0.002,
                                     "target": np.zeros(shape)}, arrival)

    reqs = [comp((64, 64, 3), 0.0), comp((32, 32, 3), 0.0), comp((64, 64, 3), 0.0),
            comp((64, 64, 3), 0.0), comp((32, 32, 3), 0.0)]
    dev = FakeDevice()
    done = Server(budget_bytes=1000, slots=1, max_batch=12, executor=dev).run(
        reqs, realtime=False)
    assert all(c.error is None for c in done)
    assert dev.batches == [[0, 2, 3], [1, 4]]


def test_plan_capacity_buckets():
    """Pooled plans come in capacities 1, 2, 3, 4, 6, 8, 12, 16, 24, ...: a batch's plane
    count is rounded up to the next one, so at most a third of a plan's planes idle."""
    from gdpoisson.serving import plan_capacity

    assert [plan_capacity(n) for n in range(1, 14)] == [1, 2, 3, 4, 6, 6, 8, 8, 12, 12, 12,
                                                        12, 16]
    for n in range(1, 200):
        c = plan_capacity(n)
        assert c >= n and (c - n) * 3 <= c
