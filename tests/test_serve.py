"""Serving loop logic on the CPU: admission by memory budget, continuous batching of steps,
image residency, and the failure paths (invalid requests, requests that can never fit, an
executor error, stopping with a queue).  The device executor is replaced by a fake that records
what each step contained."""
import threading
import time

import numpy as np
import pytest

from py_ortho_amd import serve
from py_ortho_amd.serve import Admission, DepthRequest, DepthServer, ImageStore, RequestRejected, plan_step


class _Ev:
    def query(self):
        return True

    def synchronize(self):
        pass


class FakeExecutor:
    """Step = per request (i0, nh); results: depth filled with the request's plane count, score
    with the number of steps it took.  Checks the budget on every admission."""

    def __init__(self, server_ref, fail_on=None, step_delay=0.0):
        self.server_ref = server_ref
        self.fail_on = fail_on
        self.step_delay = step_delay
        self.steps = []
        self.prepared = []

    @staticmethod
    def working_bytes(H, W, planes):
        return 12 * 4 * H * W + 4 * planes

    def prepare(self, a, store):
        if self.fail_on is not None and a.req.tag == self.fail_on:
            raise RuntimeError("prepare failed")
        for k in a.keys:
            store.pin(k)
        srv = self.server_ref[0]
        assert srv.adm.used + store.pinned_bytes_with([]) <= srv.adm.budget
        H, W = np.shape(store.loader(a.req.ref)[0])
        a.state = {"H": H, "W": W, "steps": 0}
        self.prepared.append(a.req.tag)

    on_step = None

    def step(self, items):
        if self.step_delay:
            time.sleep(self.step_delay)
        if self.on_step is not None:
            self.on_step(len(self.steps))
        self.steps.append([(a.req.tag, i0, nh, fin) for a, i0, nh, fin in items])
        done = []
        for a, i0, nh, fin in items:
            a.state["steps"] += 1
            if fin:
                H, W = a.state["H"], a.state["W"]
                done.append((a, _Arr(np.full((H, W), a.req.planes, np.float32)),
                             _Arr(np.full((H, W), a.state["steps"], np.float32))))
        return _Ev(), done

    def release(self, a, store):
        for k in a.keys:
            store.unpin(k)
        a.state = None


class _Arr:
    def __init__(self, a):
        self.a = a

    def numpy(self):
        return self.a


def _loader(sizes):
    def load(key):
        H, W = sizes[key]
        return np.zeros((H, W), np.float32), None
    return load


def _server(sizes, budget, step_planes=16, **kw):
    ref = [None]
    store = ImageStore(_loader(sizes), budget)
    ex = FakeExecutor(ref, **{k: v for k, v in kw.items() if k in ("fail_on", "step_delay")})
    srv = DepthServer(store, ex, budget_bytes=budget, step_planes=step_planes,
                      **{k: v for k, v in kw.items() if k in ("max_active", "lookahead", "max_active_pixels")})
    ref[0] = srv
    return srv, store, ex


def test_plan_step_ranges():
    assert plan_step([(0, 40), (32, 40), (0, 5)], 16) == [(0, 16, False), (32, 8, True), (0, 5, True)]
    assert plan_step([(16, 32)], 16) == [(16, 16, True)]


def test_admission_fifo_and_budget():
    adm = Admission(100)
    for name, ws in (("a", 40), ("b", 50), ("c", 30), ("d", 5)):
        adm.offer(name, ws, 0)
    ws = {"a": 40, "b": 50, "c": 30, "d": 5}
    got = adm.take(ws.get, lambda it: 0, 10)
    assert got == ["a", "b"] and adm.used == 90        # c does not fit; d waits behind it (FIFO)
    adm.release(40)
    assert adm.take(ws.get, lambda it: 0, 10) == ["c", "d"]
    with pytest.raises(RequestRejected):
        adm.offer("huge", 90, 20)


def test_image_store_lru_and_pins():
    sizes = {k: (10, 10) for k in "abcd"}           # 400 bytes each
    store = ImageStore(_loader(sizes), 1000)
    store.pin("a"), store.pin("b")
    store.unpin("b")
    store.pin("c")                                   # 1200 > 1000: b (unpinned, oldest) goes
    assert store.resident("a") and not store.resident("b") and store.resident("c")
    store.pin("a")
    assert store.uploads == 3                        # one upload per key while resident
    store.unpin("a"), store.unpin("a"), store.unpin("c")
    store.set_budget(400)
    assert store.resident_bytes <= 400
    assert store.pinned_bytes_with(["d"]) == 400


def test_continuous_batching_and_results():
    sizes = {f"v{i}": (20 + i, 30) for i in range(8)}
    budget = 10 ** 9
    srv, store, ex = _server(sizes, budget, step_planes=16)
    futs = [srv.submit(DepthRequest("v0", ["v1", "v2"], 10.0, 20.0, 64, tag="r0"))]
    late = threading.Event()

    def arrive(n):                                    # the others arrive while r0 is mid-sweep
        if n == 1:
            for i in range(1, 5):
                futs.append(srv.submit(DepthRequest(f"v{i}", [f"v{(i + 1) % 8}"], 10.0, 20.0, 16 * i,
                                                    tag=f"r{i}")))
            late.set()
    ex.on_step = arrive
    srv.start()
    assert late.wait(10)
    res = [f.result(timeout=10) for f in futs]
    srv.stop()
    for i, r in enumerate(res):
        assert r.tag == f"r{i}"
        assert r.depth.shape == sizes[f"v{i}"] and (r.depth == (64 if i == 0 else 16 * i)).all()
        assert r.latency >= 0 and r.t_admit >= r.t_arrival
    # late arrivals joined steps in which r0 was still sweeping
    joint = [s for s in ex.steps if any(t == "r0" for t, *_ in s) and any(t != "r0" for t, *_ in s)]
    assert joint and srv.stats["max_active"] > 1
    # every request's steps cover its planes in order, once
    for i in range(5):
        rng = [(i0, nh) for s in ex.steps for t, i0, nh, fin in s if t == f"r{i}"]
        assert [x[0] for x in rng] == sorted(x[0] for x in rng)
        assert sum(nh for _, nh in rng) == (64 if i == 0 else 16 * i)
    assert srv.idle() and srv.adm.used == 0


def test_budget_limits_concurrency():
    sizes = {f"v{i}": (100, 100) for i in range(6)}  # image 40 kB, working set 480 kB
    ws = FakeExecutor.working_bytes(100, 100, 32)
    budget = 2 * ws + 4 * 40000 + 1000                # room for two requests with their images
    srv, store, ex = _server(sizes, budget, step_planes=8, step_delay=0.001)
    srv.start()
    futs = [srv.submit(DepthRequest(f"v{i}", [f"v{(i + 1) % 6}"], 10.0, 20.0, 32)) for i in range(6)]
    for f in futs:
        f.result(timeout=10)
    srv.stop()
    assert max(len(s) for s in ex.steps) <= 2
    assert srv.stats["finished"] == 6


def test_invalid_and_oversize_requests_fail_alone():
    sizes = {f"v{i}": (10, 10) for i in range(20)}
    sizes["big"] = (4000, 4000)
    srv, store, ex = _server(sizes, 10 ** 6)
    srv.start()
    bad = [srv.submit(DepthRequest("v0", [], 10.0, 20.0, 8)),
           srv.submit(DepthRequest("v0", [f"v{i}" for i in range(1, 18)], 10.0, 20.0, 8)),
           srv.submit(DepthRequest("v0", ["v1"], 10.0, 20.0, 0)),
           srv.submit(DepthRequest("v0", ["v1"], 20.0, 10.0, 8)),
           srv.submit(DepthRequest("v0", ["v0"], 10.0, 20.0, 8)),
           srv.submit(DepthRequest("big", ["v1"], 10.0, 20.0, 8))]
    good = srv.submit(DepthRequest("v2", ["v3"], 10.0, 20.0, 8))
    for f in bad:
        with pytest.raises(RequestRejected):
            f.result(timeout=5)
    assert good.result(timeout=5).depth.shape == (10, 10)
    srv.stop()


def test_executor_error_fails_only_that_request():
    sizes = {f"v{i}": (10, 10) for i in range(4)}
    srv, store, ex = _server(sizes, 10 ** 6, fail_on="boom")
    srv.start()
    f1 = srv.submit(DepthRequest("v0", ["v1"], 10.0, 20.0, 8, tag="boom"))
    f2 = srv.submit(DepthRequest("v2", ["v3"], 10.0, 20.0, 8, tag="ok"))
    with pytest.raises(RuntimeError):
        f1.result(timeout=5)
    assert f2.result(timeout=5).tag == "ok"
    srv.stop()
    assert srv.adm.used == 0 and store.pinned_bytes_with([]) == 0


def test_stop_without_drain_cancels_queue():
    sizes = {f"v{i}": (10, 10) for i in range(4)}
    srv, store, ex = _server(sizes, 10 ** 6)
    futs = [srv.submit(DepthRequest("v0", ["v1"], 10.0, 20.0, 8)) for _ in range(3)]
    srv.stop(drain=False)                             # never started: everything queued is cancelled
    assert all(f.cancelled() for f in futs)
    late = srv.submit(DepthRequest("v0", ["v1"], 10.0, 20.0, 8))
    with pytest.raises(RuntimeError):
        late.result(timeout=1)


def test_pixel_limit_bounds_the_step_but_never_starves():
    sizes = {f"v{i}": (30, 40) for i in range(8)}     # 1200 px each
    sizes["big"] = (100, 100)                         # alone above the limit: still admitted when idle
    srv, store, ex = _server(sizes, 10 ** 9, step_planes=8, max_active_pixels=2500)
    futs = [srv.submit(DepthRequest(f"v{i}", [f"v{(i + 1) % 8}"], 10.0, 20.0, 32)) for i in range(6)]
    futs.append(srv.submit(DepthRequest("big", ["v0"], 10.0, 20.0, 16)))
    srv.start()
    for f in futs:
        f.result(timeout=10)
    srv.stop()
    assert max(len(s) for s in ex.steps) <= 2
    assert srv.stats["finished"] == 7



def test_images_of_requests_admitted_together_count_against_the_budget():
    """Image bytes dominate: two requests whose distinct sources do not fit together must not be
    admitted in the same pass (the FakeExecutor asserts the budget when it pins)."""
    sizes = {f"v{i}": (100, 100) for i in range(40)}  # image 40 kB
    ws = FakeExecutor.working_bytes(100, 100, 8)       # 480 kB working set
    srcs_a = [f"v{i}" for i in range(1, 17)]
    srcs_b = [f"v{i}" for i in range(21, 37)]
    one = ws + 17 * 40000                              # one request with its 17 images
    budget = one + ws + 10 * 40000                     # the second's working set fits, its images do not
    srv, store, ex = _server(sizes, budget, step_planes=8)
    fa = srv.submit(DepthRequest("v0", srcs_a, 10.0, 20.0, 8, tag="a"))
    fb = srv.submit(DepthRequest("v20", srcs_b, 10.0, 20.0, 8, tag="b"))
    srv.start()
    assert fa.result(timeout=10).tag == "a" and fb.result(timeout=10).tag == "b"
    srv.stop()
    assert max(len(s) for s in ex.steps) == 1            # never in flight together
    assert srv.stats["finished"] == 2 and srv.stats["failed"] == 0
