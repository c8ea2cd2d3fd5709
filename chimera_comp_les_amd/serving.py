"""Serving form: a request loop over independent periodic solves with
continuous batching.

Requests (isentropic vortex or viscous Taylor-Green, sizes drawn from
{32, 48, 64, 96, 128}^3, Mach uniform in [0.1, 0.5], Re log-uniform in
[100, 5000], a fixed number of SSP-RK3 steps, FP32-bulk precision) arrive as a
Poisson stream.  The loop keeps every admitted solve resident on the device,
admits waiting requests (smallest work first, while the nodes in flight stay
below a cap that keeps every CU busy, and the HBM footprint fits the budget),
also while others are in flight, advances all members together with the
batched, graph-captured RK3 step, and retires each request when its steps
are done.
The tile shape of the batched stage kernel comes from `autotune`, which times
every shape on a sample of the mix.

The same requests are also solved one at a time (`run_sequential`), which is
the baseline of the aggregate-throughput comparison; every result is hashed in
both runs so that the outputs can be checked to be identical.

    python -m chimera_comp_les_amd.serving --requests 256 --rate 64 --steps 200 --out report.md
"""
import argparse
import collections
import json
import sys
import time

import numpy as np

import chimera_comp_les_amd as cl
from chimera_comp_les_amd import cases

SIZES = (32, 48, 64, 96, 128)
TILES = {0: "32x4x2", 1: "16x4x4", 2: "16x8x2", 3: "64x2x2", 4: "6x6x6", 5: "8x5x5", 6: "4x4x3/64", 7: "8x4x3/128"}

Request = collections.namedtuple("Request", "rid kind N Ma Re steps arrival")


def make_requests(n=256, rate=64.0, steps=200, seed=20261203):
    """n requests from one seeded generator; arrival times are a Poisson
    process of `rate` requests per second."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for q in range(n):
        kind = "vortex" if rng.random() < 0.5 else "tgv"
        N = int(rng.choice(SIZES))
        Ma = float(rng.uniform(0.1, 0.5))
        Re = float(np.exp(rng.uniform(np.log(100.0), np.log(5000.0))))
        out.append(Request(q, kind, N, Ma, Re, steps, float(arrivals[q])))
    return out


def problem(req):
    if req.kind == "vortex":
        return cases.vortex_case(N=req.N, mach=req.Ma)
    return cases.tgv_case(N=req.N, Re=req.Re, Ma=req.Ma, les=0)


def payloads(reqs):
    """The initial grid and state each request brings (client data)."""
    return {r.rid: problem(r) for r in reqs}


def build(payload, precision=cl.PREC_FP32BULK):
    cfg, coor, Q = payload
    s = cl.Solver(cfg, coor, precision=precision)
    s.set_state(Q)
    return s


def work(req):
    """cell-stage updates of a request"""
    return req.N ** 3 * 3 * req.steps


def digest(s):
    """download the result and fingerprint it"""
    import xxhash
    Q, P = s.state()
    return xxhash.xxh64(Q.tobytes() + P.tobytes()).hexdigest()


def run_sequential(reqs, pay, precision=cl.PREC_FP32BULK):
    """Solve the requests one at a time: set-up, all steps, download.  The
    latency of a one-request-at-a-time server is formed from the measured
    service times: requests are served in arrival order, each starting when
    it has arrived and the previous one has finished."""
    t0 = time.perf_counter()
    res, service = {}, {}
    for r in reqs:
        a = time.perf_counter()
        s = build(pay[r.rid], precision)
        s.advance(r.steps)
        res[r.rid] = digest(s)
        s.close()
        service[r.rid] = time.perf_counter() - a
    wall = time.perf_counter() - t0
    lat, end = {}, 0.0
    for r in sorted(reqs, key=lambda q: q.arrival):
        end = max(end, r.arrival) + service[r.rid]
        lat[r.rid] = end - r.arrival
    return {"wall_s": wall, "digests": res, "latency_s": lat}


def autotune(sample, budget, nrep=5, precision=cl.PREC_FP32BULK):
    """Time one batched step of `sample` (payloads) with every tile shape;
    returns the fastest shape and the timings (ms per step)."""
    b = cl.Batch(budget)
    ms = []
    for pl in sample:
        s = build(pl, precision)
        if b.add(s):
            ms.append(s)
        else:
            s.close()
    times = {}
    for tile in TILES:
        b.set_tile(tile)
        b.time_step(1)  # graph build + warm-up
        times[tile] = b.time_step(nrep)
    b.close()
    for s in ms:
        s.close()
    best = min(times, key=times.get)
    return best, times


def run_batched(reqs, pay, budget, tile, chunk=10, precision=cl.PREC_FP32BULK, policy="sjf", cap_cells=6 << 20):
    """Continuous batching in real time.  Each round queues `chunk` batched
    steps (fewer if a member finishes earlier) and, while they run, downloads
    the requests finished in the previous round and sets up newly arrived
    ones; after the round, finished members leave and the new ones join.

    Admission: a request joins when its HBM footprint fits the budget and,
    with policy "sjf", when the nodes in flight stay below `cap_cells`
    (enough tiles to keep every CU busy); waiting requests are taken
    smallest work first, so short requests do not queue behind long ones.
    Policy "all" admits every arrived request that fits the budget, in
    arrival order."""
    pending = collections.deque(sorted(reqs, key=lambda r: r.arrival))
    waiting = []
    active = {}           # Solver -> [request, steps left]
    leaving, closing, staged = [], [], []
    b = cl.Batch(budget)
    b.set_tile(tile)
    lat, res, conc = {}, {}, []
    held = 0              # bytes of contexts outside the batch (staged / leaving / closing)
    t0 = time.perf_counter()
    while pending or waiting or active or leaving or staged:
        now = time.perf_counter() - t0
        while pending and pending[0].arrival <= now:
            waiting.append(pending.popleft())
        k = 0
        if active:
            k = min(chunk, min(v[1] for v in active.values()))
            b.launch(k)
            conc.append(len(active))
        # host work while the batch runs
        for s, r in leaving:
            res[r.rid] = digest(s)
            lat[r.rid] = time.perf_counter() - t0 - r.arrival
            closing.append(s)
        leaving = []
        if policy == "sjf":
            waiting.sort(key=work)
        inflight = sum(v[0].N ** 3 for v in active.values()) + sum(r.N ** 3 for _, r in staged)
        while waiting:
            r = waiting[0]
            need = cl.footprint(pay[r.rid][0], precision)
            if b.used() + held + need > budget:
                if not (active or staged or leaving or closing):
                    raise RuntimeError("request %d needs %d bytes, more than the budget" % (r.rid, need))
                break
            if policy == "sjf" and inflight and inflight + r.N ** 3 > cap_cells:
                break
            staged.append((build(pay[r.rid], precision), r))
            held += need
            inflight += r.N ** 3
            waiting.pop(0)
        if k:
            b.wait()
        for s in closing:   # freeing waits for an idle device: do it between rounds
            held -= s.device_bytes()
            s.close()
        closing = []
        if k:
            for s, v in list(active.items()):
                v[1] -= k
                if v[1] == 0:
                    b.remove(s)
                    leaving.append((s, v[0]))
                    held += s.device_bytes()
                    del active[s]
        for s, r in staged:
            held -= s.device_bytes()
            if not b.add(s):
                raise RuntimeError("staged request does not fit the budget")
            active[s] = [r, r.steps]
        staged = []
        if not active and not leaving and not waiting and pending:
            time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
    for s in closing:
        s.close()
    wall = time.perf_counter() - t0
    b.close()
    return {"wall_s": wall, "makespan_s": wall - min(r.arrival for r in reqs), "latency_s": lat, "digests": res,
            "mean_concurrency": float(np.mean(conc)) if conc else 0.0, "max_concurrency": max(conc) if conc else 0}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=256)
    ap.add_argument("--rate", type=float, default=64.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--chunk", type=int, default=10, help="steps between admission/retirement points")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--policy", choices=["sjf", "all"], default="sjf")
    ap.add_argument("--cap-mcells", type=float, default=6.0, help="nodes in flight (millions) under policy sjf")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    reqs = make_requests(args.requests, args.rate, args.steps, args.seed)
    budget = int(args.budget_gb * (1 << 30))
    total = sum(work(r) for r in reqs)
    pay = payloads(reqs)
    sample = [problem(Request(-1 - q, "tgv" if q % 2 else "vortex", N, 0.3, 1000.0, args.steps, 0.0))
              for q, N in enumerate(SIZES)]
    tile, ttimes = autotune(sample, budget)
    print(json.dumps({"autotune_ms_per_step": {TILES[t]: v for t, v in ttimes.items()}, "tile": TILES[tile]}),
          flush=True)
    seq = run_sequential(reqs, pay)
    print(json.dumps({"sequential_s": seq["wall_s"]}), flush=True)
    cap = int(args.cap_mcells * 1e6)
    bat = run_batched(reqs, pay, budget, tile, args.chunk, policy=args.policy, cap_cells=cap)
    lat = np.array([bat["latency_s"][r.rid] for r in reqs])
    slat = np.array([seq["latency_s"][r.rid] for r in reqs])
    same = sum(seq["digests"][r.rid] == bat["digests"][r.rid] for r in reqs)
    out = {
        "requests": len(reqs), "steps": args.steps, "rate_per_s": args.rate, "precision": "fp32bulk",
        "tile": TILES[tile], "budget_gb": args.budget_gb, "cell_stage_updates": total,
        "sequential_wall_s": seq["wall_s"], "sequential_cell_stage_per_s": total / seq["wall_s"],
        "batched_makespan_s": bat["makespan_s"], "batched_cell_stage_per_s": total / bat["makespan_s"],
        "speedup": seq["wall_s"] / bat["makespan_s"],
        "policy": args.policy, "cap_cells": cap if args.policy == "sjf" else None,
        "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
        "sequential_latency_p50_s": float(np.percentile(slat, 50)),
        "sequential_latency_p99_s": float(np.percentile(slat, 99)),
        "mean_concurrency": bat["mean_concurrency"], "max_concurrency": bat["max_concurrency"],
        "identical_results": int(same),
    }
    print(json.dumps(out), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write("# Serving: %d requests, %d steps, Poisson %.1f/s, FP32-bulk\n\n" % (len(reqs), args.steps, args.rate))
            f.write("Command: `python -m chimera_comp_les_amd.serving --requests %d --rate %g --steps %d "
                    "--budget-gb %g --chunk %d --policy %s --cap-mcells %g`\n\n"
                    % (len(reqs), args.rate, args.steps, args.budget_gb, args.chunk, args.policy, args.cap_mcells))
            f.write("Autotuner (ms per batched step of one request per size): %s -> %s\n\n"
                    % (json.dumps({TILES[t]: round(v, 4) for t, v in ttimes.items()}), TILES[tile]))
            f.write("```\n%s\n```\n" % json.dumps(out, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
