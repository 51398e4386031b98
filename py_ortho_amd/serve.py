"""Depth-map serving on one GPU: requests of uneven image sizes and neighbour counts arrive over
time; images stay resident in HBM, requests are admitted by a memory budget, and all admitted
requests advance together in continuous-batching steps.

The reference computes depth maps from a thread pool over views (densify.py:596) behind an LRU
image cache (render.py:19-64: byte-bounded, one load per key).  Here:

- `ImageStore`  device-resident images by key, bounded in bytes, least-recently-used eviction of
  images no admitted request uses, one upload per key.
- `Admission`   FIFO admission by device bytes: a request enters when the working sets of the
  admitted requests (reference statistics, rays, winner state, output) plus every image they pin
  fit under the budget; cached images no request pins are evicted to make room; a request that
  can never fit is rejected at once.
- `plan_step`   every admitted request's next hypothesis range (at most `step_planes`), and which
  requests finish in this step.
- `DepthServer` the loop: admit, plan, launch one om_sweep_multi over all admitted requests (a
  persistent launch over every (request, tile) work item), copy finished depth maps to pinned host
  memory, resolve their futures.  A request that arrives while others are on the GPU joins at the
  next step; the host stays at most `lookahead` steps ahead of the device.

Each request's result equals the single-view sweep + parabola of its own inputs (dense.sweep,
dense.sweep_finalize) bit for bit, whatever else is in flight: the winner state of a view depends
only on its hypotheses in order, and steps split that order without changing it.
"""
from __future__ import annotations

import collections
import ctypes
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import Callable, Dict, Hashable, List, Optional, Sequence, Tuple

import numpy as np

__all__ = ["DepthRequest", "DepthResult", "ImageStore", "Admission", "plan_step", "DepthServer", "GpuExecutor",
           "RequestRejected"]


class RequestRejected(Exception):
    """The request can never be served (invalid, or larger than the device budget)."""


@dataclass
class DepthRequest:
    ref: Hashable                   # image key of the reference view
    srcs: Sequence[Hashable]        # 1..16 source image keys
    dmin: float                     # depth range (m); planes are uniform in inverse depth
    dmax: float
    planes: int                     # number of fronto-parallel planes
    tag: object = None              # caller's label, returned with the result


@dataclass
class DepthResult:
    depth: np.ndarray               # (H, W) float32 parabola depth
    score: np.ndarray               # (H, W) float32 best top-k mean NCC
    t_arrival: float
    t_admit: float
    t_done: float
    tag: object = None

    @property
    def latency(self) -> float:
        return self.t_done - self.t_arrival


@dataclass
class _Entry:
    tensor: object
    nbytes: int
    pins: int = 0


class ImageStore:
    """Device-resident images (float32 (h, w), NaN = masked) with their cameras, by key.

    `loader(key) -> (image, camera)` supplies missing images (host arrays or tensors); `pin`
    uploads if needed and holds an image for an admitted request, `unpin` releases it.  Unpinned
    images are evicted least recently used first when the resident bytes exceed `budget_bytes`."""

    def __init__(self, loader: Callable, budget_bytes: float, device="cuda", upload: Optional[Callable] = None):
        self.loader = loader
        self.budget = int(budget_bytes)
        self.device = device
        self.upload = upload
        self.cams: Dict[Hashable, object] = {}
        self._lru: "collections.OrderedDict[Hashable, _Entry]" = collections.OrderedDict()
        self.resident_bytes = 0
        self.uploads = 0
        self._lock = threading.Lock()

    def image_bytes(self, key) -> int:
        """Device bytes of an image (float32), resident or not."""
        with self._lock:
            e = self._lru.get(key)
            if e is not None:
                return e.nbytes
        return int(np.prod(np.shape(self.loader(key)[0]))) * 4

    def pinned_bytes_with(self, keys) -> int:
        """Bytes of the pinned images once `keys` are pinned as well."""
        with self._lock:
            tot = sum(e.nbytes for e in self._lru.values() if e.pins > 0)
            extra = [k for k in set(keys) if k not in self._lru or self._lru[k].pins == 0]
        return tot + sum(self.image_bytes(k) for k in extra)

    def set_budget(self, budget_bytes: float):
        with self._lock:
            self.budget = int(budget_bytes)
            self._evict()

    def pin(self, key):
        with self._lock:
            e = self._lru.get(key)
            if e is not None:
                self._lru.move_to_end(key)
                e.pins += 1
                return e.tensor
        img, cam = self.loader(key)
        t = self.upload(img) if self.upload else img
        nb = int(np.prod(np.shape(img))) * 4
        with self._lock:
            e = self._lru.get(key)
            if e is None:
                e = _Entry(t, nb)
                self._lru[key] = e
                self.cams[key] = cam
                self.resident_bytes += nb
                self.uploads += 1
            e.pins += 1
            self._lru.move_to_end(key)
            self._evict()
            return e.tensor

    def unpin(self, key):
        with self._lock:
            e = self._lru.get(key)
            if e is not None and e.pins > 0:
                e.pins -= 1
            self._evict()

    def camera(self, key):
        with self._lock:
            if key in self.cams:
                return self.cams[key]
        return self.loader(key)[1]

    def _evict(self):
        for k in list(self._lru.keys()):
            if self.resident_bytes <= self.budget:
                break
            e = self._lru[k]
            if e.pins == 0:
                self.resident_bytes -= e.nbytes
                del self._lru[k]
                self.cams.pop(k, None)

    def resident(self, key) -> bool:
        with self._lock:
            return key in self._lru


class Admission:
    """FIFO admission by device bytes.  The budget covers the working sets of admitted requests
    plus the images they pin; `offer` queues a request (rejecting one that could never fit),
    `take` admits queued requests in arrival order while the next one fits, `release` returns a
    finished request's working set."""

    def __init__(self, budget_bytes: float):
        self.budget = int(budget_bytes)
        self.used = 0                        # working-set bytes of admitted requests
        self.queue: "collections.deque[object]" = collections.deque()

    def offer(self, item, ws_bytes: int, img_bytes: int):
        if ws_bytes + img_bytes > self.budget:
            raise RequestRejected(f"request needs {(ws_bytes + img_bytes) / 1e9:.2f} GB, more than the "
                                  f"{self.budget / 1e9:.2f} GB budget")
        self.queue.append(item)

    def take(self, ws_of: Callable, pinned_with: Callable, room: int) -> list:
        """pinned_with(item) = bytes of all pinned images once item's are pinned too."""
        out = []
        while self.queue and len(out) < room:
            item = self.queue[0]
            ws = ws_of(item)
            if self.used + ws + pinned_with(item) > self.budget:
                break
            self.queue.popleft()
            self.used += ws
            out.append(item)
        return out

    def release(self, ws_bytes: int):
        self.used -= int(ws_bytes)
        assert self.used >= 0

    def __len__(self):
        return len(self.queue)


def plan_step(progress: Sequence[Tuple[int, int]], step_planes: int) -> List[Tuple[int, int, bool]]:
    """progress[j] = (next hypothesis, total) of admitted request j -> per request (i0, nh, finishes)."""
    out = []
    for nxt, total in progress:
        nh = max(0, min(step_planes, total - nxt))
        out.append((nxt, nh, nxt + nh >= total))
    return out


# ------------------------------------------------------------------------------------ execution
class _Active:
    __slots__ = ("req", "fut", "need", "t_arrival", "t_admit", "next", "total", "state", "keys")

    def __init__(self, req, fut, need, t_arrival):
        self.req, self.fut, self.need, self.t_arrival = req, fut, need, t_arrival
        self.t_admit = 0.0
        self.next = 0
        self.total = int(req.planes)
        self.state = None
        self.keys = [req.ref] + list(req.srcs)


class GpuExecutor:
    """Device side of the server: per-request working set and one om_sweep_multi per step."""

    def __init__(self, window: int = 5, top_k: int = 2, tile_h: int = 0, max_views: int = 512, ring: int = 4,
                 device="cuda"):
        import torch

        from . import _lib
        self.torch = torch
        self.lib = _lib.lib()
        self.window, self.top_k, self.tile_h = int(window), int(top_k), int(tile_h)
        self.device = device
        self.stream = torch.cuda.Stream(device=device)
        self.max_views = int(max_views)
        nb = int(self.lib.om_sweep_multi_bytes(self.max_views))
        self.ring = [(torch.empty(nb, dtype=torch.uint8, device=device),
                      torch.empty(nb, dtype=torch.uint8, pin_memory=True), None) for _ in range(ring)]
        self.ring_i = 0
        self.table_bytes = nb

    def upload(self, img):
        torch = self.torch
        with torch.cuda.stream(self.stream):
            if isinstance(img, torch.Tensor):
                t = img.to(self.device, torch.float32, non_blocking=True)
            else:
                t = torch.from_numpy(np.ascontiguousarray(img, np.float32)).pin_memory().to(self.device,
                                                                                            non_blocking=True)
            return t.contiguous()

    @staticmethod
    def working_bytes(H: int, W: int, planes: int) -> int:
        """Device bytes of one request besides its images: reference (zeroed copy), mu, sd, rays (3),
        winner state (best, prev, s_prev_best, s_next_best, idx), depth: 12 planes of 4 bytes, and
        the plane table."""
        return 12 * 4 * H * W + 4 * planes

    def prepare(self, a: _Active, store: ImageStore):
        from . import block, dense
        torch = self.torch
        req = a.req
        imgs = [store.pin(k) for k in a.keys]
        cams = [store.camera(k) for k in a.keys]
        ref = imgs[0]
        H, W = int(ref.shape[0]), int(ref.shape[1])
        c = cams[0]
        with torch.cuda.stream(self.stream):
            refz = torch.where(torch.isfinite(ref), ref, torch.zeros_like(ref)).contiguous()
            mu, sd = dense.ref_stats(refz, self.window)
            ray = dense.rays_world(H, W, c.R, c.cx, c.cy, c.f, c.k1, c.k2, c.k3, device=ref.device)
            planes = torch.from_numpy(block.inverse_depth_planes(req.dmin, req.dmax, req.planes)).pin_memory().to(
                ref.device, non_blocking=True)
            st = dense.SweepState(H, W, ref.device)
            depth = torch.empty((H, W), dtype=torch.float32, device=ref.device)
        geo = np.stack([block._geo18(cs, c) for cs in cams[1:]])
        a.state = dict(ref=refz, mu=mu, sd=sd, ray=ray, planes=planes, st=st, depth=depth, srcs=imgs[1:], geo=geo,
                       H=H, W=W)

    def step(self, items: List[Tuple[_Active, int, int, bool]]):
        """Launches one step; -> (event, [(active, host depth, host score)] of finishing requests)."""
        from . import _lib, dense
        torch = self.torch
        n = len(items)
        if n > self.max_views:
            raise ValueError("more admitted requests than max_views")
        table, staging, ev = self.ring[self.ring_i]
        if ev is not None:
            ev.synchronize()                       # the staging buffer's previous upload has been read
        descs = (_lib.OmSweepDesc * n)()
        outs = (ctypes.c_void_p * n)()
        keep = []
        for j, (a, i0, nh, fin) in enumerate(items):
            s = a.state
            # const planes: hyp_a points at hypothesis i0 of the whole table, whose entries 0 and 1
            # the parabola of a finishing step reads (hyp_a[-i0], hyp_a[1 - i0])
            hyp_a = s["planes"][i0:]
            d, k = dense._sweep_desc(s["ref"], s["mu"], s["sd"], s["srcs"], s["ray"], s["geo"], self.window,
                                     self.top_k, s["st"], _lib.OM_HYP_CONST, hyp_a, None, 0, i0, nh, 0, self.tile_h)
            descs[j] = d
            keep += [k, hyp_a]
            outs[j] = s["depth"].data_ptr() if fin else None
        done = []
        with torch.cuda.stream(self.stream):
            _lib.check(self.lib.om_sweep_multi(descs, n, outs, self.tile_h, table.data_ptr(), self.table_bytes,
                                               staging.data_ptr(), ctypes.c_void_p(self.stream.cuda_stream)),
                       "sweep step")
            for (a, i0, nh, fin) in items:
                if fin:
                    s = a.state
                    hd = torch.empty((s["H"], s["W"]), dtype=torch.float32, pin_memory=True)
                    hs = torch.empty((s["H"], s["W"]), dtype=torch.float32, pin_memory=True)
                    hd.copy_(s["depth"], non_blocking=True)
                    hs.copy_(s["st"].best, non_blocking=True)
                    done.append((a, hd, hs))
            ev = torch.cuda.Event()
            ev.record(self.stream)
        self.ring[self.ring_i] = (table, staging, ev)
        self.ring_i = (self.ring_i + 1) % len(self.ring)
        del keep
        return ev, done

    def release(self, a: _Active, store: ImageStore):
        for k in a.keys:
            store.unpin(k)
        a.state = None


class DepthServer:
    """submit() -> Future[DepthResult].  start() runs the loop in a thread; stop() ends it (queued
    requests are cancelled)."""

    def __init__(self, store: ImageStore, executor, budget_bytes: float = 200e9, step_planes: int = 32,
                 lookahead: int = 2, max_active: int = 256, max_active_pixels: float = float("inf"),
                 clock: Callable[[], float] = time.perf_counter):
        self.store, self.ex = store, executor
        self.adm = Admission(budget_bytes)
        self.step_planes = int(step_planes)
        self.lookahead = int(lookahead)
        self.max_active = int(max_active)
        # reference pixels in flight: enough to keep the device busy, few enough that admitted
        # requests finish quickly instead of all advancing together (latency under overload)
        self.max_active_pixels = float(max_active_pixels)
        self.clock = clock
        self.active: List[_Active] = []
        self._inflight: "collections.deque" = collections.deque()   # (event, finished list)
        self._cv = threading.Condition()
        self._stop = False
        self._thread: Optional[threading.Thread] = None
        self.steps = 0
        self.stats = {"admitted": 0, "finished": 0, "rejected": 0, "failed": 0, "max_active": 0}

    # ---- client side
    def submit(self, req: DepthRequest, t_arrival: Optional[float] = None) -> Future:
        fut: Future = Future()
        t = self.clock() if t_arrival is None else t_arrival
        try:
            self._validate(req)
            a = _Active(req, fut, 0, t)
            shape = np.shape(self.store.loader(req.ref)[0])
            a.need = self.ex.working_bytes(int(shape[0]), int(shape[1]), int(req.planes))
            img_bytes = sum(self.store.image_bytes(k) for k in set(a.keys))
            with self._cv:
                if self._stop:
                    raise RuntimeError("server stopped")
                self.adm.offer(a, a.need, img_bytes)
                self._cv.notify()
        except Exception as e:  # invalid or impossible: the future carries the error, the server goes on
            self.stats["rejected"] += 1
            fut.set_exception(e)
        return fut

    def _validate(self, req: DepthRequest):
        if not 1 <= len(req.srcs) <= 16:
            raise RequestRejected("a request needs 1..16 source views")
        if int(req.planes) < 1:
            raise RequestRejected("a request needs at least one plane")
        if not (0 < req.dmin < req.dmax):
            raise RequestRejected("depth range must satisfy 0 < dmin < dmax")
        if req.ref in set(req.srcs):
            raise RequestRejected("the reference view cannot be its own source")

    def start(self):
        self._thread = threading.Thread(target=self._loop, name="depth-server", daemon=True)
        self._thread.start()
        return self

    def stop(self, drain: bool = True, timeout: Optional[float] = None):
        with self._cv:
            if not drain:
                while self.adm.queue:
                    a = self.adm.queue.popleft()
                    a.fut.cancel()
            self._stop = True
            self._cv.notify()
        if self._thread is not None:
            self._thread.join(timeout)

    def idle(self) -> bool:
        with self._cv:
            return not self.active and not self.adm.queue and not self._inflight

    # ---- the loop
    def _pixels(self, a: _Active) -> int:
        return int(np.prod(np.shape(self.store.loader(a.req.ref)[0])))

    def _admit(self):
        with self._cv:
            room = self.max_active - len(self.active)
            if room <= 0:
                return
            px = sum(self._pixels(a) for a in self.active)
            got = []
            # images of the requests taken in this pass are pinned only by prepare() below, so each
            # candidate is checked against the images of every request taken before it as well
            pending = set()
            while self.adm.queue and len(got) < room:
                a = self.adm.queue[0]
                p = self._pixels(a)
                if (self.active or got) and px + p > self.max_active_pixels:
                    break
                if not self.adm.take(lambda x: x.need, lambda x: self.store.pinned_bytes_with(pending | set(x.keys)), 1):
                    break
                px += p
                pending |= set(a.keys)
                got.append(a)
        for a in got:
            try:
                a.t_admit = self.clock()
                self.ex.prepare(a, self.store)
                self.active.append(a)
                self.stats["admitted"] += 1
            except Exception as e:
                self._fail(a, e)
        # cached images that no admitted request pins give way to working sets
        self.store.set_budget(self.adm.budget - self.adm.used)
        self.stats["max_active"] = max(self.stats["max_active"], len(self.active))

    def _fail(self, a: _Active, e: Exception):
        self.stats["failed"] += 1
        try:
            self.ex.release(a, self.store)
        finally:
            with self._cv:
                self.adm.release(a.need)
            if not a.fut.done():
                a.fut.set_exception(e)

    def _complete(self, block: bool):
        while self._inflight:
            ev, done = self._inflight[0]
            if not block and not ev.query():
                return
            ev.synchronize()
            self._inflight.popleft()
            t = self.clock()
            for a, hd, hs in done:
                res = DepthResult(hd.numpy(), hs.numpy(), a.t_arrival, a.t_admit, t, a.req.tag)
                self.stats["finished"] += 1
                if not a.fut.done():
                    a.fut.set_result(res)
            block = False

    def _loop(self):
        while True:
            self._admit()
            if not self.active:
                self._complete(block=True)
                with self._cv:
                    if self._stop and not self.adm.queue:
                        break
                    if not self.adm.queue:
                        self._cv.wait(0.05)
                    continue
            plan = plan_step([(a.next, a.total) for a in self.active], self.step_planes)
            items = [(a, i0, nh, fin) for a, (i0, nh, fin) in zip(self.active, plan)]
            try:
                ev, done = self.ex.step(items)
            except Exception as e:
                for a in list(self.active):
                    self._fail(a, e)
                self.active = []
                continue
            self.steps += 1
            self._inflight.append((ev, done))
            finished = {id(a) for a, _, _ in done}
            keep = []
            for a, (i0, nh, fin) in zip(self.active, plan):
                a.next = i0 + nh
                if id(a) in finished:
                    self.ex.release(a, self.store)
                    with self._cv:
                        self.adm.release(a.need)
                else:
                    keep.append(a)
            self.active = keep
            while len(self._inflight) > self.lookahead:
                self._complete(block=True)
            self._complete(block=False)
        self._complete(block=True)
