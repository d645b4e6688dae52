"""Request loop and admission logic of spectrecon_amd.serve with a simulated
backend (no GPU): HBM-budget admission and rejection, plan pooling and reuse,
the in-flight cap, eviction of idle plans, and the stats."""
import numpy as np

from spectrecon_amd.serve import Config, Request, Server


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


class FakeBackend:
    """Each request takes `dur[cfg]` seconds of simulated time."""

    def __init__(self, clock, dur):
        self.clock, self.dur = clock, dur
        self.made, self.closed, self.started = [], [], []

    def make_plan(self, cfg):
        plan = ("plan", cfg, len(self.made))
        self.made.append(plan)
        return plan

    def start(self, plan, req):
        self.started.append((plan, req.rid))
        return self.clock() + self.dur[req.config]

    def done(self, plan, token):
        return self.clock() >= token

    def result(self, plan, token):
        return np.full(3, token)

    def release(self, plan, token):
        pass

    def close_plan(self, plan):
        self.closed.append(plan)


A = Config(32, 16, 16, 7)
B = Config(64, 32, 60, 15)


def req(cfg, rid, arrival=0.0):
    z = np.zeros(1, np.float32)
    return Request(cfg, z, z, z, 4, 1, arrival, rid)


def drive(server, clock, reqs, step=0.1, horizon=100.0):
    for r in reqs:
        server.submit(r)
    while (server.queue or server.busy) and clock.t < horizon:
        server.poll()
        clock.t += step
    server.poll()


def test_footprint_orders_sizes():
    assert A.footprint() < B.footprint()
    big = Config(128, 80, 128, 37)
    assert big.footprint() > 4 * 128 * 128 * 80 * 128  # at least the attenuation factors


def test_reject_over_budget_and_serve_the_rest():
    clock = Clock()
    be = FakeBackend(clock, {A: 1.0, B: 1.0})
    s = Server(hbm_budget=A.footprint() * 1.5, max_active=4, backend=be, clock=clock)
    drive(s, clock, [req(A, 0), req(B, 1), req(A, 2)])
    st = s.stats(clock.t)
    assert st["rejected"] == 1 and [r.rid for r in s.rejected] == [1]
    assert st["served"] == 2
    # one plan fits: the second A request reuses it after the first finishes
    assert len(be.made) == 1
    assert [rid for _, rid in be.started] == [0, 2]


def test_concurrency_cap_and_pool_growth():
    clock = Clock()
    be = FakeBackend(clock, {A: 1.0})
    s = Server(hbm_budget=A.footprint() * 10, max_active=3, backend=be, clock=clock, plans_per_config=3)
    for k in range(5):
        s.submit(req(A, k))
    s.poll()
    assert len(s.busy) == 3 and len(be.made) == 3 and len(s.queue) == 2
    clock.t = 1.0
    s.poll()
    assert len(s.busy) == 2 and len(be.made) == 3  # plans reused, none built
    clock.t = 2.0
    s.poll()
    assert not s.busy and not s.queue and len(s.finished) == 5


def test_idle_plans_of_other_configs_are_evicted():
    clock = Clock()
    be = FakeBackend(clock, {A: 1.0, B: 1.0})
    s = Server(hbm_budget=B.footprint() + A.footprint() / 2, max_active=2, backend=be, clock=clock)
    drive(s, clock, [req(A, 0)])
    assert len(be.made) == 1 and not be.closed
    drive(s, clock, [req(B, 1)])
    assert be.closed == [be.made[0]]  # A's idle plan made room for B
    assert len(s.finished) == 2


def test_queued_request_does_not_block_other_configs():
    clock = Clock()
    be = FakeBackend(clock, {A: 1.0, B: 5.0})
    s = Server(hbm_budget=B.footprint() + 1.5 * A.footprint(), max_active=4, backend=be, clock=clock)
    s.submit(req(B, 0))
    s.poll()
    s.submit(req(B, 1))  # waits for B's only plan (no memory for a second)
    s.submit(req(A, 2))  # admitted meanwhile
    s.poll()
    assert {r.rid for _, r, _ in s.busy.values()} == {0, 2}
    assert [r.rid for r in s.queue] == [1]


def test_stats_latency():
    clock = Clock()
    be = FakeBackend(clock, {A: 1.0})
    s = Server(hbm_budget=A.footprint() * 4, max_active=1, backend=be, clock=clock)
    drive(s, clock, [req(A, k) for k in range(3)], step=0.5)
    st = s.stats(clock.t)
    assert st["served"] == 3 and st["max_queued"] == 3
    lat = sorted(r.latency for r in s.finished)
    assert lat[0] >= 1.0 and lat[-1] >= 3.0
    assert st["latency_ms"]["p99"] >= st["latency_ms"]["p50"]


def test_plans_per_config_cap_and_prepare():
    clock = Clock()
    be = FakeBackend(clock, {A: 1.0})
    s = Server(hbm_budget=A.footprint() * 10, max_active=4, backend=be, clock=clock, plans_per_config=1)
    s.prepare([A])
    assert len(be.made) == 1
    for k in range(3):
        s.submit(req(A, k))
    s.poll()
    assert len(s.busy) == 1 and len(be.made) == 1  # the others wait for the resident plan
    drive(s, clock, [])
    assert len(s.finished) == 3 and len(be.made) == 1


def test_prepare_tunes_once_per_config_and_subset_size():
    """With a tuning backend, prepare() tunes each (config, subset size) once,
    on one plan of the config, and the stats carry the tuning log."""
    clock = Clock()

    class TuningBackend(FakeBackend):
        autotune = True

        def __init__(self, *a):
            super().__init__(*a)
            self.tune_log = []

        def tune(self, plan, cfg, nsub):
            if not any(t["key"] == (cfg, nsub) for t in self.tune_log):
                self.tune_log.append({"key": (cfg, nsub), "plan": plan})

    be = TuningBackend(clock, {A: 1.0, B: 1.0})
    srv = Server(hbm_budget=1e12, backend=be, clock=clock, plans_per_config=2)
    srv.prepare([A, B], per_config=2, nsubs={A: {4, 2}, B: {8}})
    assert len(be.made) == 4
    assert sorted((t["key"][0].nx, t["key"][1]) for t in be.tune_log) == [(32, 2), (32, 4), (64, 8)]
    assert all(t["plan"] == srv.pools[t["key"][0]][0] for t in be.tune_log)
    srv.prepare([A], per_config=2, nsubs={A: {4}})  # already tuned
    assert len(be.tune_log) == 3
    assert srv.stats(1.0)["autotune"] == be.tune_log
    plain = Server(hbm_budget=1e12, backend=FakeBackend(clock, {A: 1.0}), clock=clock)
    plain.prepare([A], nsubs={A: {4}})  # a backend without tuning: plans only
    assert plain.stats(1.0)["autotune"] == []


def test_request_failing_its_checks_is_rejected_not_fatal():
    """A request whose backend start raises the reference's argument error
    (e.g. "need background > 0") is rejected with that message; the loop
    goes on with the others."""
    from spectrecon_amd._lib import SPECTError

    clock = Clock()

    class Picky(FakeBackend):
        def start(self, plan, req):
            if req.rid == 1:
                raise SPECTError("need background > 0")
            return super().start(plan, req)

    be = Picky(clock, {A: 1.0})
    server = Server(hbm_budget=1e12, max_active=4, backend=be, clock=clock)
    reqs = [req(A, k) for k in range(3)]
    drive(server, clock, reqs)
    assert [r.rid for r in server.rejected] == [1] and reqs[1].error == "need background > 0"
    assert sorted(r.rid for r in server.finished) == [0, 2]
