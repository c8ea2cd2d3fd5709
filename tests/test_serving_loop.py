"""Host logic of the continuous-batching request loop (serving.py) on the CPU,
with stand-in contexts and batch: every request runs exactly its steps, leaves
when done, admission respects the HBM budget and the in-flight cap, and the
smallest-work-first policy serves short requests before long ones, and
aging lets a long request that has waited past the limit go first."""
import pytest

import chimera_comp_les_amd as cl
from chimera_comp_les_amd import serving as sv


class FakeSolver:
    def __init__(self, req):
        self.req = req
        self.done = 0
        self.closed = False

    def device_bytes(self):
        return footprint_of(self.req.N)

    def close(self):
        self.closed = True


def footprint_of(N):
    return N ** 3 * 8 * 40


class FakeBatch:
    log = []

    def __init__(self, budget):
        self.budget = budget
        self.members = []
        self.running = 0
        FakeBatch.log = []

    def set_tile(self, t):
        pass

    def set_mode(self, m):
        pass

    def set_graph(self, on):
        pass

    def used(self):
        return sum(s.device_bytes() for s in self.members)

    def add(self, s):
        assert not self.running
        if self.used() + s.device_bytes() > self.budget:
            return False
        self.members.append(s)
        return True

    def remove(self, s):
        assert not self.running
        self.members.remove(s)

    def launch(self, k):
        assert k > 0 and not self.running
        for s in self.members:
            s.done += k
            assert s.done <= s.req.steps
        FakeBatch.log.append([s.req for s in self.members])
        self.running = k

    def wait(self):
        self.running = 0

    def close(self):
        pass


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(cl, "Batch", FakeBatch)
    monkeypatch.setattr(cl, "footprint", lambda cfg, prec=None: footprint_of(cfg))
    monkeypatch.setattr(sv, "build", lambda payload, precision=None: FakeSolver(payload[0]))
    monkeypatch.setattr(sv, "digest", lambda s: (s.req.rid, s.done))


def requests(n=40, steps=30, seed=3):
    reqs = sv.make_requests(n, 1e6, steps, seed)  # everything arrives at once
    return reqs, {r.rid: r for r in reqs}


def payloads_of(reqs):
    # payload = (config, coordinates, state); the stand-ins keep the request in the config slot
    return {r.rid: (r, None, None) for r in reqs}


@pytest.mark.parametrize("policy", ["aging", "sjf", "all"])
def test_every_request_runs_exactly_its_steps(fake, monkeypatch, policy):
    reqs, _ = requests()
    pay = payloads_of(reqs)
    monkeypatch.setattr(cl, "footprint", lambda r, prec=None: footprint_of(r.N))
    out = sv.run_batched(reqs, pay, budget=1 << 40, tile=1, chunk=7, policy=policy, cap_cells=3 << 20)
    assert sorted(out["digests"]) == sorted(r.rid for r in reqs)
    for r in reqs:
        assert out["digests"][r.rid] == (r.rid, r.steps)
        assert out["latency_s"][r.rid] >= 0


def test_sjf_caps_work_in_flight_and_serves_short_first(fake, monkeypatch):
    reqs, by = requests(n=60)
    pay = payloads_of(reqs)
    monkeypatch.setattr(cl, "footprint", lambda r, prec=None: footprint_of(r.N))
    cap = 3 << 20
    sv.run_batched(reqs, pay, budget=1 << 40, tile=1, chunk=10, policy="sjf", cap_cells=cap)
    first_seen = {}
    for step, members in enumerate(FakeBatch.log):
        cells = sum(r.N ** 3 for r in members)
        assert cells <= cap or len(members) == 1
        for r in members:
            first_seen.setdefault(r.rid, step)
    # a request of the smallest size never starts after one of the largest size
    small = [first_seen[r.rid] for r in reqs if r.N == min(sv.SIZES)]
    large = [first_seen[r.rid] for r in reqs if r.N == max(sv.SIZES)]
    assert small and large and max(small) <= min(large)


def test_budget_limits_residency(fake, monkeypatch):
    reqs, _ = requests(n=30)
    pay = payloads_of(reqs)
    monkeypatch.setattr(cl, "footprint", lambda r, prec=None: footprint_of(r.N))
    budget = 3 * footprint_of(128)
    out = sv.run_batched(reqs, pay, budget=budget, tile=1, chunk=10, policy="all")
    assert len(out["digests"]) == len(reqs)
    for members in FakeBatch.log:
        assert sum(footprint_of(r.N) for r in members) <= budget


def test_sequential_latency_is_a_fifo_server(monkeypatch):
    """One request at a time in arrival order: a request waits for the one
    before it, an idle server starts it on arrival."""
    clock = [0.0]
    service = {0: 2.0, 1: 1.0, 2: 1.0}

    class Timed(FakeSolver):
        def advance(self, n):
            clock[0] += service[self.req.rid]
            self.done += n

    monkeypatch.setattr(sv.time, "perf_counter", lambda: clock[0])
    monkeypatch.setattr(sv, "build", lambda payload, precision=None: Timed(payload[0]))
    monkeypatch.setattr(sv, "digest", lambda s: (s.req.rid, s.done))
    reqs = [sv.Request(0, "tgv", 32, 0.3, 100.0, 5, 0.0), sv.Request(1, "tgv", 32, 0.3, 100.0, 5, 0.5),
            sv.Request(2, "tgv", 32, 0.3, 100.0, 5, 10.0)]
    out = sv.run_sequential(reqs, payloads_of(reqs))
    assert out["latency_s"] == {0: 2.0, 1: 2.5, 2: 1.0}
    assert out["digests"] == {0: (0, 5), 1: (1, 5), 2: (2, 5)}
    assert out["wall_s"] == 4.0


def test_admission_order_ages_long_requests():
    R = sv.Request
    waiting = [R(0, "tgv", 128, 0.3, 100.0, 10, 0.0), R(1, "tgv", 32, 0.3, 100.0, 10, 4.0),
               R(2, "tgv", 64, 0.3, 100.0, 10, 1.0), R(3, "tgv", 96, 0.3, 100.0, 10, 0.5)]
    ids = lambda rs: [r.rid for r in rs]
    assert ids(sv.admission_order(waiting, 4.5, "sjf", 2.0)) == [1, 2, 3, 0]
    # at t = 4.5, requests 0 (waited 4.5 s), 3 (4.0 s) and 2 (3.5 s) are past 2 s: arrival order first
    assert ids(sv.admission_order(waiting, 4.5, "aging", 2.0)) == [0, 3, 2, 1]
    assert ids(sv.admission_order(waiting, 1.2, "aging", 2.0)) == [1, 2, 3, 0]
    assert ids(sv.admission_order(waiting, 4.5, "all", 2.0)) == [0, 3, 2, 1]


def test_aging_bounds_the_wait_of_a_long_request(fake, monkeypatch):
    """A 128^3 request among a steady stream of small ones: under "sjf" with a
    cap that fits one large or many small members it waits until the small
    ones run out; under "aging" it starts once it has waited age_s."""
    clock = [0.0]

    def tick():
        clock[0] += 0.01
        return clock[0]

    monkeypatch.setattr(sv.time, "perf_counter", tick)
    monkeypatch.setattr(sv.time, "sleep", lambda d: None)
    monkeypatch.setattr(cl, "footprint", lambda r, prec=None: footprint_of(r.N))
    R = sv.Request
    reqs = [R(0, "tgv", 128, 0.3, 100.0, 20, 0.1)] + [R(q, "tgv", 32, 0.3, 100.0, 20, 0.02 * q) for q in range(1, 200)]
    pay = payloads_of(reqs)
    start = {}
    for policy in ("sjf", "aging"):
        sv.run_batched(reqs, pay, budget=1 << 40, tile=1, chunk=5, policy=policy, cap_cells=(128 ** 3) // 2,
                       age_s=0.5)
        start[policy] = next(i for i, m in enumerate(FakeBatch.log) if any(r.rid == 0 for r in m))
    assert start["aging"] < start["sjf"] / 4, start
