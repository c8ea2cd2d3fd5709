"""Serving form: a request loop over independent periodic solves with
continuous batching.

Requests (isentropic vortex or viscous Taylor-Green, the latter with the
dynamic Smagorinsky model in a share of them, sizes drawn from
{32, 48, 64, 96, 128}^3, Mach uniform in [0.1, 0.5], Re log-uniform in
[100, 5000], a fixed number of SSP-RK3 steps, FP32-bulk precision) arrive as a
Poisson stream.  The loop keeps every admitted solve resident on the device,
admits waiting requests (smallest work first, except that a request waiting
longer than an aging limit goes ahead of every younger one, while the nodes
in flight stay below a cap that keeps every CU busy, and the HBM footprint
fits the budget), also while others are in flight, advances all members
together with the batched, graph-captured RK3 step (batched stage
launches, or one persistent stage kernel per group of members with --mode
persistent), and retires each request when its steps are done.
The tile shape of the batched stage kernel comes from `autotune`, which times
every shape on a sample of the mix.

The same requests are also solved one at a time (`run_sequential`), which is
the baseline of the aggregate-throughput comparison; every result is hashed in
both runs so that the outputs can be checked to be identical.

    python -m chimera_comp_les_amd.serving --requests 256 --rate 64 --steps 200 --out report.md
"""
import argparse
import collections
import concurrent.futures
import json
import sys
import time

import numpy as np

import chimera_comp_les_amd as cl
from chimera_comp_les_amd import cases

SIZES = (32, 48, 64, 96, 128)
TILES = {0: "32x4x2", 1: "16x4x4", 2: "16x8x2", 3: "64x2x2", 4: "6x6x6", 5: "8x5x5", 6: "4x4x3/64", 7: "8x4x3/128"}

Request = collections.namedtuple("Request", "rid kind N Ma Re steps arrival")


def make_requests(n=256, rate=64.0, steps=200, seed=20261203, sizes=SIZES, les_share=0.25):
    """n requests from one seeded generator; arrival times are a Poisson
    process of `rate` requests per second.  Kinds: vortex (half), Taylor-Green
    with the dynamic Smagorinsky model (les_share) and without."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for q in range(n):
        u = rng.random()
        kind = "vortex" if u < 0.5 else ("tgv-les" if u < 0.5 + les_share else "tgv")
        N = int(rng.choice(sizes))
        Ma = float(rng.uniform(0.1, 0.5))
        Re = float(np.exp(rng.uniform(np.log(100.0), np.log(5000.0))))
        out.append(Request(q, kind, N, Ma, Re, steps, float(arrivals[q])))
    return out


def problem(req):
    if req.kind == "vortex":
        return cases.vortex_case(N=req.N, mach=req.Ma)
    return cases.tgv_case(N=req.N, Re=req.Re, Ma=req.Ma, les=2 if req.kind == "tgv-les" else 0)


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


def admission_order(waiting, now, policy, age_s):
    """Waiting requests in the order they are offered a slot.  "sjf": smallest
    work first; "aging": smallest work first among the requests that have
    waited less than age_s, after every older one in arrival order (a long
    request cannot be passed over for ever); "all": arrival order."""
    if policy == "sjf":
        return sorted(waiting, key=work)
    if policy == "aging":
        return sorted(waiting, key=lambda r: (0, r.arrival, 0) if now - r.arrival >= age_s else (1, work(r), r.arrival))
    return sorted(waiting, key=lambda r: r.arrival)


def run_batched(reqs, pay, budget, tile, chunk=10, precision=cl.PREC_FP32BULK, policy="aging", cap_cells=6 << 20,
                age_s=2.0, mode=cl.BATCH_LAUNCHES, workers=2, graph=True):
    """Continuous batching in real time.  Each round queues `chunk` batched
    steps (fewer if a member finishes earlier); while they run, host workers
    set up newly admitted requests (context, metrics, state upload) and
    retire finished ones (download, fingerprint, free), so that the device
    only waits for the membership change between rounds.  After the round,
    finished members leave and the set-up ones join.

    Admission: a request joins when its HBM footprint fits the budget and,
    with policies "sjf" and "aging", when the nodes in flight stay below
    `cap_cells` (enough tiles to keep every CU busy); the order in which
    waiting requests are offered a slot is `admission_order`'s.  Policy
    "all" admits every arrived request that fits the budget, in arrival
    order.  workers=0 does the host work inline."""
    pool = concurrent.futures.ThreadPoolExecutor(max_workers=workers) if workers else None

    def submit(fn, *args):
        if pool:
            return pool.submit(fn, *args)
        f = concurrent.futures.Future()
        f.set_result(fn(*args))
        return f

    def retire(s, r):
        d = digest(s)
        t = time.perf_counter() - t0 - r.arrival
        nbytes = s.device_bytes()
        s.close()
        return d, t, nbytes

    pending = collections.deque(sorted(reqs, key=lambda r: r.arrival))
    waiting = []
    active = {}           # Solver -> [request, steps left]
    staged = []           # (future Solver, request, bytes): being set up
    retiring = []         # (future (digest, latency, bytes), request)
    b = cl.Batch(budget)
    b.set_tile(tile)
    b.set_mode(mode)
    b.set_graph(graph)
    lat, res, conc = {}, {}, []
    held = 0              # bytes of contexts outside the batch (staged / retiring)
    t_launch = t_wait = 0.0
    t0 = time.perf_counter()
    while pending or waiting or active or staged or retiring:
        now = time.perf_counter() - t0
        while pending and pending[0].arrival <= now:
            waiting.append(pending.popleft())
        k = 0
        if active:
            k = min(chunk, min(v[1] for v in active.values()))
            a = time.perf_counter()
            b.launch(k)
            t_launch += time.perf_counter() - a
            conc.append(len(active))
        # admission while the batch runs
        waiting = admission_order(waiting, now, policy, age_s)
        capped = policy in ("sjf", "aging")
        inflight = sum(v[0].N ** 3 for v in active.values()) + sum(r.N ** 3 for _, r, _ in staged)
        while waiting:
            r = waiting[0]
            need = cl.footprint(pay[r.rid][0], precision)
            if b.used() + held + need > budget:
                if not (active or staged or retiring):
                    raise RuntimeError("request %d needs %d bytes, more than the budget" % (r.rid, need))
                break
            if capped and inflight and inflight + r.N ** 3 > cap_cells:
                break
            staged.append((submit(build, pay[r.rid], precision), r, need))
            held += need
            inflight += r.N ** 3
            waiting.pop(0)
        if k:
            a = time.perf_counter()
            b.wait()
            t_wait += time.perf_counter() - a
        # requests retired by the workers
        still = []
        for f, r in retiring:
            if f.done():
                res[r.rid], lat[r.rid], nbytes = f.result()
                held -= nbytes
            else:
                still.append((f, r))
        retiring = still
        if k:
            for s, v in list(active.items()):
                v[1] -= k
                if v[1] == 0:
                    b.remove(s)
                    held += s.device_bytes()
                    retiring.append((submit(retire, s, v[0]), v[0]))
                    del active[s]
        if not active:
            # nothing on the device: block on the host work instead of spinning
            if staged:
                concurrent.futures.wait([f for f, _, _ in staged], return_when=concurrent.futures.FIRST_COMPLETED)
            elif retiring:
                concurrent.futures.wait([f for f, _ in retiring], return_when=concurrent.futures.FIRST_COMPLETED)
            elif pending and not waiting:
                time.sleep(max(0.0, pending[0].arrival - (time.perf_counter() - t0)))
        still = []
        for f, r, need in staged:
            if f.done():
                s = f.result()
                held -= need
                if not b.add(s):
                    raise RuntimeError("staged request does not fit the budget")
                active[s] = [r, r.steps]
            else:
                still.append((f, r, need))
        staged = still
    wall = time.perf_counter() - t0
    b.close()
    if pool:
        pool.shutdown()
    return {"wall_s": wall, "makespan_s": wall - min(r.arrival for r in reqs), "latency_s": lat, "digests": res,
            "mean_concurrency": float(np.mean(conc)) if conc else 0.0, "max_concurrency": max(conc) if conc else 0,
            "host_launch_s": t_launch, "host_wait_s": t_wait}


def solo_rates(sizes=SIZES + (256,), steps=20, precision=cl.PREC_FP32BULK):
    """Single-context throughput per grid size (viscous Taylor-Green without
    the LES model, cell-stage updates per second over `steps` steps after two
    warm-up steps), and its fraction of the largest size's: how much of the
    device a solve of that size uses on its own."""
    out = {}
    for N in sizes:
        cfg, coor, Q = cases.tgv_case(N=N, Re=1000.0, Ma=0.3, les=0)
        s = cl.Solver(cfg, coor, precision=precision)
        s.set_state(Q)
        s.advance(2)
        t0 = time.perf_counter()
        s.advance(steps)
        dt = time.perf_counter() - t0
        s.close()
        out[N] = {"ms_per_step": 1e3 * dt / steps, "cell_stage_per_s": N ** 3 * 3 * steps / dt}
    top = out[max(sizes)]["cell_stage_per_s"]
    for N in sizes:
        out[N]["fraction_of_%d" % max(sizes)] = out[N]["cell_stage_per_s"] / top
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=256)
    ap.add_argument("--rate", type=float, default=64.0, help="Poisson arrival rate, requests per second")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--budget-gb", type=float, default=64.0)
    ap.add_argument("--chunk", type=int, default=10, help="steps between admission/retirement points")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--policy", choices=["aging", "sjf", "all"], default="aging")
    ap.add_argument("--age-s", type=float, default=2.0, help="waiting time after which a request goes first (aging)")
    ap.add_argument("--cap-mcells", type=float, default=6.0, help="nodes in flight (millions) under aging and sjf")
    ap.add_argument("--les-share", type=float, default=0.25, help="share of Taylor-Green requests with the LES model")
    ap.add_argument("--mode", choices=["launches", "persistent"], default="launches")
    ap.add_argument("--graph", type=int, choices=[0, 1], default=0,
                    help="replay the batched step as a hipGraph (1) or launch its kernels each step (0)")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    mode = cl.BATCH_PERSISTENT if args.mode == "persistent" else cl.BATCH_LAUNCHES
    reqs = make_requests(args.requests, args.rate, args.steps, args.seed, les_share=args.les_share)
    budget = int(args.budget_gb * (1 << 30))
    total = sum(work(r) for r in reqs)
    pay = payloads(reqs)
    rates = solo_rates()
    print(json.dumps({"solo_rates": rates}), flush=True)
    sample = [problem(Request(-1 - q, "tgv" if q % 2 else "vortex", N, 0.3, 1000.0, args.steps, 0.0))
              for q, N in enumerate(SIZES)]
    tile, ttimes = autotune(sample, budget)
    print(json.dumps({"autotune_ms_per_step": {TILES[t]: v for t, v in ttimes.items()}, "tile": TILES[tile]}),
          flush=True)
    seq = run_sequential(reqs, pay)
    print(json.dumps({"sequential_s": seq["wall_s"]}), flush=True)
    cap = int(args.cap_mcells * 1e6)
    bat = run_batched(reqs, pay, budget, tile, args.chunk, policy=args.policy, cap_cells=cap, age_s=args.age_s,
                      mode=mode, graph=bool(args.graph))
    lat = np.array([bat["latency_s"][r.rid] for r in reqs])
    slat = np.array([seq["latency_s"][r.rid] for r in reqs])
    same = sum(seq["digests"][r.rid] == bat["digests"][r.rid] for r in reqs)
    kinds = collections.Counter(r.kind for r in reqs)
    out = {
        "requests": len(reqs), "kinds": dict(kinds), "steps": args.steps, "rate_per_s": args.rate,
        "precision": "fp32bulk", "mode": args.mode, "graph": bool(args.graph),
        "tile": TILES[tile], "budget_gb": args.budget_gb, "cell_stage_updates": total,
        "sequential_wall_s": seq["wall_s"], "sequential_cell_stage_per_s": total / seq["wall_s"],
        "batched_makespan_s": bat["makespan_s"], "batched_cell_stage_per_s": total / bat["makespan_s"],
        "speedup": seq["wall_s"] / bat["makespan_s"],
        "policy": args.policy, "age_s": args.age_s if args.policy == "aging" else None,
        "cap_cells": cap if args.policy != "all" else None,
        "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
        "sequential_latency_p50_s": float(np.percentile(slat, 50)),
        "sequential_latency_p99_s": float(np.percentile(slat, 99)),
        "mean_concurrency": bat["mean_concurrency"], "max_concurrency": bat["max_concurrency"],
        "host_launch_s": bat["host_launch_s"], "host_wait_s": bat["host_wait_s"],
        "identical_results": int(same),
    }
    print(json.dumps(out), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write("# Serving: %d requests, %d steps, Poisson %.1f/s, FP32-bulk\n\n" % (len(reqs), args.steps, args.rate))
            f.write("Command: `python -m chimera_comp_les_amd.serving --requests %d --rate %g --steps %d "
                    "--budget-gb %g --chunk %d --policy %s --age-s %g --cap-mcells %g --les-share %g --mode %s --graph %d`\n\n"
                    % (len(reqs), args.rate, args.steps, args.budget_gb, args.chunk, args.policy, args.age_s,
                       args.cap_mcells, args.les_share, args.mode, args.graph))
            f.write("Single-context throughput per size (Taylor-Green, FP32-bulk, cell-stage updates/s):\n\n")
            f.write("| N | ms/step | cell-stage/s | fraction of 256^3 |\n|---|---|---|---|\n")
            for N, v in rates.items():
                f.write("| %d | %.3f | %.3e | %.2f |\n" % (N, v["ms_per_step"], v["cell_stage_per_s"],
                                                       v["fraction_of_256"]))
            f.write("\nAutotuner (ms per batched step of one request per size): %s -> %s\n\n"
                    % (json.dumps({TILES[t]: round(v, 4) for t, v in ttimes.items()}), TILES[tile]))
            f.write("```\n%s\n```\n" % json.dumps(out, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())
