"""Serving and batching of OS-EM reconstructions of different sizes.

No reference counterpart: the reference runs one `osem!` at a time.  Here a
request loop keeps plans (mu-map rotations, attenuation factors, PSF taps,
rotation-adjoint lists) resident on the device, pooled per configuration, and
admits requests by HBM footprint.  Every plan owns one HIP stream, so admitted
requests run concurrently on the device (each queued asynchronously on its
plan's stream from the one loop thread), and a request that arrives while
others run is admitted as soon as a plan of its configuration and the memory
for it are free.  The kernels of one plan do not depend on other plans, so a
result is the same alone and in the mix.  `Batch` runs many instances from one
library call (sr_osem_batch).

The queue and admission logic is backend-independent (`Server(backend=...)`);
`DeviceBackend` drives the library.
"""
from __future__ import annotations

import collections
import dataclasses
import time

import numpy as np

from ._lib import DimensionMismatch, SPECTError


@dataclasses.dataclass(frozen=True)
class Config:
    """What a resident plan is built from (requests with equal configs share plans)."""
    nx: int
    nz: int
    nview: int
    psf: int
    dy: float = 0.48
    fwhm_start: float = 1.5
    fwhm_end: float = 6.0

    def image_shape(self):
        return (self.nx, self.nx, self.nz)

    def views_shape(self):
        return (self.nx, self.nz, self.nview)

    def footprint(self, max_batch=64):
        """Device bytes of one plan of this config and its per-call buffers
        (attenuation factors and their products with the normaliser's PSF
        adjoint, lists, work volumes, normalisers), an upper estimate used
        for admission."""
        V = self.nx * self.nx * self.nz
        batch = min(self.nview, max_batch)
        atten = 2 * 4 * V * self.nview
        lists = 8 * 8 * self.nx * self.nx * self.nview
        work = 4 * V * (batch + 8)
        images = 4 * V * (2 + 8)
        views = 4 * self.nx * self.nz * self.nview * 4
        return atten + lists + work + images + views


@dataclasses.dataclass
class Request:
    config: Config
    x0: np.ndarray
    y: np.ndarray
    b: np.ndarray
    nblocks: int
    niter: int
    arrival: float = 0.0
    rid: int = 0
    result: np.ndarray | None = None
    start: float | None = None
    finish: float | None = None
    rejected: bool = False
    error: str | None = None  # why a rejected request was rejected

    @property
    def latency(self):
        return None if self.finish is None else self.finish - self.arrival


class _Staging:
    """A plan's request buffers, made once per plan: pinned host arrays (the
    request is copied in and the result out with plain numpy copies) and
    device arrays (the reconstruction's inputs and output)."""

    def __init__(self, cfg: Config):
        from .device import DeviceArray, PinnedArray
        img, views = cfg.image_shape(), cfg.views_shape()
        self.h = {k: PinnedArray(img if k in ("x0", "out") else views) for k in ("x0", "y", "b", "out")}
        self.d = {k: DeviceArray(img if k in ("x0", "out") else views) for k in ("x0", "y", "b", "out")}


class DeviceBackend:
    """Plans and OS-EM on the device through the C-ABI, with no host thread
    per request.  An admitted request's arrays go through its plan's pinned
    buffers and asynchronous copies on the plan's stream; the reconstruction
    is queued there with SR_ASYNC (the background was checked on the host,
    SR_NO_CHECK) and its result copied back behind it, so `start` returns at
    once and `done` polls the stream (sr_plan_query).  Plans replay their EM
    kernel sequence from a HIP graph (sr_plan_set_graph), captured on a
    plan's second request of a given subset count.  Admitted requests run on
    different plans, so their streams overlap on the device."""

    def __init__(self, autotune: bool = False, graphs: bool = True):
        from .api import SPECTplan, mumap_phantom, psf_gauss
        self._SPECTplan, self._mumap, self._psf = SPECTplan, mumap_phantom, psf_gauss
        # autotune: kernel choices are timed once per (config, subset size),
        # on one plan of the config (in Server.prepare, before concurrent
        # work, or else on first use), and set on whichever plan of the
        # config serves a request of that size, so that a request's result
        # does not depend on the plan it lands on
        self.autotune = autotune
        self.graphs = graphs
        self.tuned = {}     # (config, nsub) -> kernel choices
        self.tune_log = []  # {"config", "nsub", "choice", "ms", "us"} per tuning
        self.staging = {}   # id(plan) -> _Staging

    def tune(self, plan, cfg: Config, nsub: int):
        key = (cfg, nsub)
        if key in self.tuned:
            return
        t0 = time.perf_counter()
        rec = plan.autotune(nsub)
        self.tuned[key] = plan.choice()
        self.tune_log.append({"config": f"{cfg.nx}x{cfg.nx}x{cfg.nz}/{cfg.nview}/psf{cfg.psf}", "nsub": nsub,
                              "choice": self.tuned[key], "ms": 1e3 * (time.perf_counter() - t0),
                              "us": {k: rec[k]["choice"]["us"] for k in ("psf_method", "psf_fwd", "psf_adj", "gather")
                                     if k in rec}})

    def make_plan(self, cfg: Config):
        mumap = self._mumap(cfg.nx, cfg.nx, cfg.nz)
        psf = self._psf(ny=cfg.nx, px=cfg.psf, fwhm_start=cfg.fwhm_start, fwhm_end=cfg.fwhm_end)
        psfs = np.asfortranarray(np.repeat(psf[:, :, :, None], cfg.nview, 3))
        plan = self._SPECTplan(mumap, psfs, cfg.dy)
        if self.graphs:
            plan.set_graph(True)
        return plan

    def _stage(self, plan, cfg):
        st = self.staging.get(id(plan))
        if st is None:
            st = self.staging[id(plan)] = _Staging(cfg)
        return st

    def start(self, plan, req: Request):
        from . import _lib as L
        if self.autotune:
            nsub = -(-req.config.nview // req.nblocks)
            choice = self.tuned.get((req.config, nsub))
            if choice is None:
                self.tune(plan, req.config, nsub)
            elif plan.choice() != choice:
                plan.set_choice(**choice)
        if not np.all(req.b > 0):  # src/ml-os-em.jl:106, before any device work
            raise SPECTError("need background > 0")
        st = self._stage(plan, req.config)
        lib = L.lib()
        h = plan.handle
        for k, a in (("x0", req.x0), ("y", req.y), ("b", req.b)):
            st.h[k].array[...] = a
            L.check(lib.sr_plan_memcpy_async(h, st.d[k].ptr, st.h[k].ptr, st.d[k].nbytes, L.SR_COPY_H2D))
        L.check(lib.sr_osem_strided(h, st.d["out"].ptr, st.d["x0"].ptr, st.d["y"].ptr, st.d["b"].ptr,
                                    int(req.nblocks), int(req.niter),
                                    L.SR_MEM_DEVICE | L.SR_ASYNC | L.SR_NO_CHECK))
        L.check(lib.sr_plan_memcpy_async(h, st.h["out"].ptr, st.d["out"].ptr, st.d["out"].nbytes, L.SR_COPY_D2H))
        return st

    def done(self, plan, token) -> bool:
        import ctypes as C

        from . import _lib as L
        d = C.c_int32()
        L.check(L.lib().sr_plan_query(plan.handle, C.byref(d)))
        return bool(d.value)

    def result(self, plan, token):
        plan.synchronize()
        return np.array(token.h["out"].array, order="F")

    def release(self, plan, token):
        pass

    def close_plan(self, plan):
        st = self.staging.pop(id(plan), None)
        plan.close()
        if st is not None:
            for a in list(st.h.values()) + list(st.d.values()):
                a.free()


def batch_stats():
    """sr_osem_batch's shared-launch counters on the current device:
    instances run as shared launches, grouped launches, batch graphs captured
    and replayed."""
    import ctypes as C

    from . import _lib as L
    v = [C.c_int64() for _ in range(4)]
    L.check(L.lib().sr_batch_stats(*[C.byref(x) for x in v]))
    return dict(zip(("instances", "launches", "captures", "replays"), (x.value for x in v)))


class Batch:
    """Batched OS-EM on resident plans: request i on plans[i] (distinct
    plans), its inputs uploaded to device arrays once.  `run()` runs all
    reconstructions from one sr_osem_batch call: the subset steps of every
    instance on a z-fastest plan as shared launches (one grid per kernel
    instantiation and step), replayed from one HIP graph when the plans have
    graphs on; `run(streams=True)` each instance on its plan's stream,
    concurrently; `run(sequential=True)` one call per instance, each waited
    for.  The device buffers stay between runs, so the launch sequence is
    captured on the second run and replayed after that."""

    def __init__(self, plans, requests):
        import ctypes as C

        from .device import DeviceArray
        self.plans, self.requests = list(plans), list(requests)
        n = len(self.requests)
        self.dev = [{k: DeviceArray.from_host(getattr(r, k)) for k in ("x0", "y", "b")} for r in self.requests]
        self.outs = [DeviceArray(r.config.image_shape()) for r in self.requests]
        self._P = (C.c_void_p * n)(*[p.handle for p in self.plans])
        self._arr = {k: (C.c_void_p * n)(*[d[k].ptr for d in self.dev]) for k in ("x0", "y", "b")}
        self._out = (C.c_void_p * n)(*[o.ptr for o in self.outs])
        self._nb = (C.c_int32 * n)(*[int(r.nblocks) for r in self.requests])
        self._ni = (C.c_int32 * n)(*[int(r.niter) for r in self.requests])

    def run(self, sequential=False, streams=False):
        """Seconds for all reconstructions (returns after they finished)."""
        import ctypes as C

        from . import _lib as L
        lib = L.lib()
        mem = L.SR_MEM_DEVICE | (L.SR_BATCH_STREAMS if streams else 0)
        L.check(lib.sr_device_synchronize())
        t0 = time.perf_counter()
        if sequential:
            V = C.c_void_p
            for i in range(len(self.requests)):
                L.check(lib.sr_osem_batch(1, C.byref(V(self._P[i])), C.byref(V(self._out[i])),
                                          C.byref(V(self._arr["x0"][i])), C.byref(V(self._arr["y"][i])),
                                          C.byref(V(self._arr["b"][i])), C.byref(C.c_int32(self._nb[i])),
                                          C.byref(C.c_int32(self._ni[i])), mem))
        else:
            L.check(lib.sr_osem_batch(len(self.requests), self._P, self._out, self._arr["x0"], self._arr["y"],
                                      self._arr["b"], self._nb, self._ni, mem))
        return time.perf_counter() - t0

    def results(self):
        return [o.to_host() for o in self.outs]

    def free(self):
        for d in self.dev:
            for a in d.values():
                a.free()
        for o in self.outs:
            o.free()


class Server:
    """Continuous batching of OS-EM requests over resident, pooled plans.

    - `submit(req)` queues a request (FIFO); a request whose plan alone would
      exceed the HBM budget is rejected.
    - `poll()` retires finished requests, then admits queued ones in arrival
      order while a plan of their config is idle, or a new plan fits in the
      budget (at most `max_active` in flight).  A queued request that cannot
      be admitted does not block later ones of other configs.
    - `run(requests)` replays a timed arrival stream and returns the stats.
    """

    def __init__(self, hbm_budget: float, max_active: int = 8, backend=None, clock=time.perf_counter,
                 plans_per_config: int = 2):
        self.budget = float(hbm_budget)
        self.max_active = max_active
        # building a plan takes far longer than a reconstruction: beyond this
        # many per config, requests wait for a resident plan of their config
        self.plans_per_config = plans_per_config
        self.backend = backend if backend is not None else DeviceBackend()
        self.clock = clock
        self.pools = collections.defaultdict(list)  # config -> [plan]
        self.busy = {}                              # id(plan) -> (plan, request, token)
        self.resident = 0.0
        self.queue = collections.deque()
        self.finished = []
        self.rejected = []
        self.max_queued = 0

    def submit(self, req: Request):
        if req.config.footprint() > self.budget:
            req.rejected = True
            self.rejected.append(req)
            return
        self.queue.append(req)
        self.max_queued = max(self.max_queued, len(self.queue))

    def _idle_plan(self, cfg):
        for plan in self.pools[cfg]:
            if id(plan) not in self.busy:
                return plan
        return None

    def _evict_idle(self, need):
        """Free idle plans of other configs until `need` bytes fit."""
        for cfg, plans in list(self.pools.items()):
            for plan in list(plans):
                if self.resident + need <= self.budget:
                    return
                if id(plan) not in self.busy:
                    plans.remove(plan)
                    self.backend.close_plan(plan)
                    self.resident -= cfg.footprint()

    def poll(self):
        now = self.clock()
        for key, (plan, req, token) in list(self.busy.items()):
            if self.backend.done(plan, token):
                req.result = self.backend.result(plan, token)
                self.backend.release(plan, token)
                req.finish = self.clock()
                self.finished.append(req)
                del self.busy[key]
        pending = collections.deque()
        while self.queue:
            req = self.queue.popleft()
            if len(self.busy) >= self.max_active:
                pending.append(req)
                continue
            plan = self._idle_plan(req.config)
            if plan is None and len(self.pools[req.config]) < self.plans_per_config:
                need = req.config.footprint()
                if self.resident + need > self.budget:
                    self._evict_idle(need)
                if self.resident + need <= self.budget:
                    plan = self.backend.make_plan(req.config)
                    self.pools[req.config].append(plan)
                    self.resident += need
            if plan is None:
                pending.append(req)
                continue
            req.start = now
            try:
                token = self.backend.start(plan, req)
            except (SPECTError, DimensionMismatch) as e:  # the request's argument checks
                req.rejected, req.error = True, str(e)
                self.rejected.append(req)
                continue
            self.busy[id(plan)] = (plan, req, token)
        self.queue = pending
        return len(self.busy) + len(self.queue)

    def run(self, requests, poll_interval=2e-4, progress=None):
        """Replay requests at their `arrival` offsets (seconds from now);
        `progress(submitted)` is called after every 250th submission."""
        t0 = self.clock()
        order = sorted(requests, key=lambda r: r.arrival)
        for r in order:
            r.arrival += t0
        i = 0
        while i < len(order) or self.queue or self.busy:
            now = self.clock()
            while i < len(order) and order[i].arrival <= now:
                self.submit(order[i])
                i += 1
                if progress and i % 250 == 0:
                    progress(i)
            if not self.poll() and i < len(order):
                time.sleep(max(0.0, min(poll_interval, order[i].arrival - self.clock())))
            elif self.busy:
                time.sleep(poll_interval)
        return self.stats(self.clock() - t0)

    def prepare(self, configs, per_config=1, nsubs=None):
        """Build resident plans ahead of the requests (within the budget);
        with nsubs ({config: subset sizes}) and a tuning backend, time each
        plan's kernel choices for those subset sizes now."""
        for cfg in configs:
            while len(self.pools[cfg]) < per_config and self.resident + cfg.footprint() <= self.budget:
                self.pools[cfg].append(self.backend.make_plan(cfg))
                self.resident += cfg.footprint()
            if nsubs and getattr(self.backend, "autotune", False) and self.pools[cfg]:
                for n in sorted(nsubs.get(cfg, ())):
                    self.backend.tune(self.pools[cfg][0], cfg, n)

    def reset_stats(self):
        self.finished, self.rejected, self.max_queued = [], [], 0

    def stats(self, wall):
        lat = sorted(r.latency for r in self.finished)
        pct = (lambda q: float(np.percentile(lat, q)) * 1e3) if lat else (lambda q: None)
        return {"requests": len(self.finished) + len(self.rejected), "served": len(self.finished),
                "rejected": len(self.rejected), "max_queued": self.max_queued, "wall_s": wall,
                "recons_per_s": len(self.finished) / wall if wall > 0 else None,
                "latency_ms": {"p50": pct(50), "p99": pct(99)},
                "plans": {f"{c.nx}x{c.nx}x{c.nz}/{c.nview}": len(p) for c, p in self.pools.items()},
                "autotune": list(getattr(self.backend, "tune_log", []))}

    def close(self):
        for plans in self.pools.values():
            for plan in plans:
                self.backend.close_plan(plan)
        self.pools.clear()
        self.resident = 0.0


def batched_instances(seed: int = 7, n: int = 32, dy: float = 0.48):
    """The batched uneven workload's instances, drawn from a seed: nx in
    {64, 96, 128}, nz in {32, 64, 80}, views in {60, 120, 128}, square
    Gaussian PSF side in {15, 25, 37} (FWHM 1 -> 4 px), subsets in {4, 8};
    [(config, nblocks)]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx = int(rng.choice([64, 96, 128]))
        cfg = Config(nx, int(rng.choice([32, 64, 80])), int(rng.choice([60, 120, 128])), int(rng.choice([15, 25, 37])),
                     dy, 1.0, 4.0)
        out.append((cfg, int(rng.choice([4, 8]))))
    return out


def synthetic_request(cfg: Config, seed: int, nblocks: int, niter: int, arrival: float = 0.0, rid: int = 0,
                      scale: float = 5e6, plans=None):
    """A request of the served stream: seeded phantom, Poisson counts with mean
    total `scale` per projection set, b = 0.1 mean, x0 = 1.  The measurement
    is made by projecting through a plan of the config (kept in `plans`, a
    dict, when one is given; else a temporary one)."""
    from .api import phantom, poisson, project
    img = phantom(cfg.nx, cfg.nx, cfg.nz, seed=seed)
    plan = plans.get(cfg) if plans is not None else None
    temp = plan is None
    if temp:
        plan = DeviceBackend().make_plan(cfg)
    try:
        ytrue = project(img, plan)
    finally:
        if plans is not None:
            plans[cfg] = plan
        elif temp:
            plan.close()
    s = scale / float(ytrue.sum())
    b = np.full_like(ytrue, 0.1 * float((s * ytrue).mean()))
    y = poisson(s * ytrue + b, seed=seed)
    return Request(cfg, np.ones(img.shape, np.float32, order="F"), y, b, nblocks, niter, arrival, rid)
