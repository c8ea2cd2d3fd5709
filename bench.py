#!/usr/bin/env python3
"""Cell-stage updates/s of the Chimera_Comp_LES right-hand-side + SSP-RK path
on one MI355X.

Workload (default): Taylor-Green vortex, periodic [0, 2 pi]^3 at 256^3 unique
nodes, Re 1600, Ma 0.1, Pr 0.71, Sutherland viscosity, viscous fluxes plus the
dynamic Smagorinsky model, WENO3 reconstruction (hybrid central/Rusanov flux,
sigma = 1) and SSP-RK3, FP64, synthetic initial state, fixed dt from CFL 0.4.

One "step" is one physical time step: the LES model update and the three RK
stages (fused RHS+update kernel, boundary kernel).  value = K * stages *
256^3 / wall time of K steps, with the state resident in HBM throughout.

Also reported: the HBM roofline of the fused stage kernel (algorithmic bytes
per cell-stage / measured mean kernel duration from HIP events on the
library's stream) and of the whole step (algorithmic bytes of the stages and
the LES model / driver-timed step), the HBM traffic per stage launch from the
rocprofv3 counter summary when it was measured on the current kernel sources,
and the CPU restatement timed on this host (all cores at the same size, and
one thread on a bounded sample).

    python bench.py                                    # TGV-LES 256^3 FP64, WENO3/RK3
    python bench.py --case vortex                      # isentropic vortex 256^3, WENO3/Rusanov RK3
    python bench.py --scheme muscl                     # TGV-LES 256^3, MUSCL/RK2
    python bench.py --case vortex --scheme weno5roe    # isentropic vortex 256^3, WENO5 + Roe flux, RK3
    python bench.py --precision fp32bulk               # FP32 fluxes, FP64 metrics/RK/EOS
"""
import argparse
import glob
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per cell-stage of the fused stage kernel (viscous + LES):
# read Q 40, P 8, csi/eta/zet 72, aj 8, nvert 8, Qold 40, nu_t 8, sigma 24;
# write Q 40, P 8
B_STAGE_LES = 256
B_STAGE_INVISCID = 224
# LES model once per step: read Q 40, the stale state of the last CalcUCat
# (rho, rho u) 32, metrics 80, nvert 8; write Cs 8, nu_t 8
B_LES_STEP = 176
# sources whose change makes a stored traffic figure stale
TRAFFIC_SOURCES = ("chimera_comp_les_amd/csrc/cles_stage.hip", "chimera_comp_les_amd/csrc/cles_internal.hpp")


def source_sha():
    h = hashlib.sha256()
    for rel in TRAFFIC_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--scheme", choices=["weno", "muscl", "weno5roe"], default="weno",
                    help="weno: the reference's WENO3 (RK3); muscl: MUSCL-minmod (RK2); weno5roe: WENO5 with the "
                         "Roe flux (RK3, FP64), the schemes of the vortex configuration")
    ap.add_argument("--precision", choices=["fp64", "fp32bulk"], default="fp64")
    ap.add_argument("--case", choices=["tgv", "vortex", "sphere"], default="tgv")
    ap.add_argument("--cpu-N", type=int, default=None, help="CPU sample size (default: the GPU size)")
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="oracle steps of the all-core CPU sample (default: 30 stages, 8 steps for the sphere)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="rocprofv3 FETCH/WRITE summary used for roofline.traffic")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 512 MiB device-to-device copy three times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against a known byte count)")
    ap.add_argument("--store-traffic", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                    help="only turn four rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE of this bench; the same "
                         "of --copy-probe) into the --traffic entry of this workload, stamped with the "
                         "hash of the stage kernel sources")
    return ap.parse_args()


def copy_probe():
    import torch
    n = 64 << 20  # 64 Mi doubles = 512 MiB
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "launches": 3}), flush=True)
    return 0


def host_cpu():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count() or 1
    return model, os.cpu_count() or 1, usable


def oracle_rate(pyoracle, workload, args, N, threads, steps, mesh_case=False):
    """cell-stage updates/s of the oracle on `threads` threads: `steps` steps
    after one warm-up step (the first step of a fresh run has Cs = 0)"""
    from chimera_comp_les_amd import cases
    mesh = None
    if workload == "tgv":
        cfg, coor, Q = cases.tgv_case(N=N, fluxorder=2 if args.scheme == "muscl" else 3)
    elif workload == "sphere":
        cfg, coor, Q, V, T = cases.sphere_channel_case()
        mesh = (V, T)
    else:
        cfg, coor, Q = cases.vortex_case(N=N)
        cfg = dict(cfg, fluxorder=2 if args.scheme == "muscl" else 3)
    if args.scheme == "weno5roe":
        cfg = dict(cfg, weno5=1, roe=1)
    o = pyoracle.Oracle(cfg, coor, threads=threads)
    o.set_state(Q)
    del Q, coor
    if mesh is not None:
        o.ibm_search(*mesh)
    o.step(1)
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    o.close()
    return steps * cfg["fluxorder"] * cases.interior_cells(cfg) / dt, dt, cfg["fluxorder"]


def cpu_baseline(args, workload):
    """The oracle (plain C, -O3, OpenMP) on this host: all usable cores (at
    most 16, the host share of one GPU) at the GPU's size, and one thread on
    a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    model, nproc, usable = host_cpu()
    threads = max(1, min(16, usable))
    N = args.cpu_N or args.N
    stages = 2 if (workload != "sphere" and args.scheme == "muscl") else 3
    steps = args.cpu_steps or (8 if workload == "sphere" else max(1, 30 // stages))
    rate, dt, _ = oracle_rate(pyoracle, workload, args, N, threads, steps)
    size = "192x128x128" if workload == "sphere" else "%d^3" % N
    n1 = 128 if workload != "sphere" else None
    one = None
    if n1:
        r1, d1, _ = oracle_rate(pyoracle, workload, args, n1, 1, 1)
        one = {"value": r1, "cores": 1, "sample": "%d^3, 1 step after 1 warm-up step, %.2f s" % (n1, d1)}
    return {"value": rate, "unit": "cell-stage/s", "cores": threads, "kind": "port",
            "sample": "CPU restatement (oracle/, C, -O3, OpenMP %d threads), same workload at %s, "
                      "%d steps (%d stages) after 1 warm-up step, %.2f s" % (threads, size, steps, steps * stages, dt),
            "one_thread": one, "host": {"model": model, "nproc": nproc, "usable": usable}}


def store_traffic(args, key, kernel_prefix="stage_kernel"):
    """bytes per stage-kernel launch from the FETCH_SIZE / WRITE_SIZE runs,
    each scaled by what the counter reads for the 512 MiB copy"""
    import csv

    def per_dispatch(d, counter, match):
        vals = {}
        for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as f:
                for row in csv.DictReader(f):
                    if row.get("Counter_Name") != counter or match not in row.get("Kernel_Name", ""):
                        continue
                    did = row.get("Dispatch_Id")
                    vals[did] = vals.get(did, 0.0) + float(row["Counter_Value"])
        if not vals:
            raise SystemExit("no %s rows for %s under %s" % (counter, match, d))
        return sum(vals.values()) / len(vals), len(vals)

    fetch, nf = per_dispatch(args.store_traffic[0], "FETCH_SIZE", kernel_prefix)
    write, nw = per_dispatch(args.store_traffic[1], "WRITE_SIZE", kernel_prefix)
    cfetch, _ = per_dispatch(args.store_traffic[2], "FETCH_SIZE", "copyBuffer")
    cwrite, _ = per_dispatch(args.store_traffic[3], "WRITE_SIZE", "copyBuffer")
    copy_kib = (64 << 20) * 8 / 1024.0
    fs, ws = copy_kib / cfetch, copy_kib / cwrite
    entry = {"kernel": kernel_prefix, "dispatches": [nf, nw], "fetch_bytes": fetch * fs * 1024,
             "write_bytes": write * ws * 1024, "fetch_scale": fs, "write_scale": ws,
             "bytes_per_launch": (fetch * fs + write * ws) * 1024, "source_sha": source_sha()}
    tr = {}
    if os.path.exists(args.traffic):
        with open(args.traffic) as f:
            tr = json.load(f)
    tr[key] = entry
    with open(args.traffic, "w") as f:
        json.dump(tr, f, indent=1)
    print(json.dumps({key: entry}), flush=True)
    return 0


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2
    if args.copy_probe:
        return copy_probe()
    tkey = "%s_%s_%s_%s" % (args.case, args.N if args.case != "sphere" else "c4", args.scheme, args.precision)
    if args.store_traffic:
        return store_traffic(args, tkey)
    import chimera_comp_les_amd as cl
    from chimera_comp_les_amd import cases
    fluxorder = 2 if args.scheme == "muscl" else 3
    schemes = dict(weno5=1, roe=1) if args.scheme == "weno5roe" else {}
    if schemes and args.precision != "fp64":
        print("the WENO5 and Roe schemes run in the FP64 form only", file=sys.stderr)
        return 2
    if args.case == "tgv":
        cfg, coor, Q = cases.tgv_case(N=args.N, fluxorder=fluxorder)
        # with the LES form's sigma = 1 the hybrid flux is the central flux
        # alone (the reference multiplies the upwind part by 1 - sigma = 0)
        workload = ("taylor-green %d^3 Re1600 Ma0.1 Pr0.71 Sutherland dynamic-Smagorinsky central "
                    "(sigma=1) flux SSP-RK%d" % (args.N, fluxorder))
        b_stage = B_STAGE_LES
    elif args.case == "sphere":
        cfg, coor, Q, V, T = cases.sphere_channel_case()
        workload = ("immersed sphere D=1 in a channel, 128x128x192 nodes (x periodic, y walls tanh 2.2, z "
                    "inflow/outflow), icosphere 20480 triangles jittered 1e-3 D, Re_D 300 Ma 0.3 Sutherland "
                    "dynamic-Smagorinsky isothermal-2 walls SSP-RK3")
        b_stage = B_STAGE_LES
    else:
        cfg, coor, Q = cases.vortex_case(N=args.N)
        cfg = dict(cfg, fluxorder=fluxorder)
        flux = {"weno": "WENO3/Rusanov", "muscl": "MUSCL/Rusanov", "weno5roe": "WENO5 (Jiang-Shu)/Roe"}[args.scheme]
        workload = "isentropic-vortex %d^3 curvilinear inviscid %s SSP-RK%d" % (args.N, flux, fluxorder)
        b_stage = B_STAGE_INVISCID
    cfg = dict(cfg, **schemes)
    prec = cl.PREC_FP64 if args.precision == "fp64" else cl.PREC_FP32BULK
    s = cl.Solver(cfg, coor, precision=prec)
    s.set_state(Q)
    extra = {}
    if args.case == "sphere":
        t_s = time.perf_counter()
        n_ib = s.ibm_search(V, T)
        nv = s.nvert()
        extra = {"ib_nodes": n_ib, "solid_nodes": int((nv > 2.5).sum()),
                 "fluid_cells": int((nv[1:-1, 1:-1, 1:-1] < 0.5).sum()),
                 "ibm_search_s": time.perf_counter() - t_s}
    del Q, coor
    stages = cfg["fluxorder"]
    cells = cases.interior_cells(cfg)
    s.advance(args.warmup)
    t0 = time.perf_counter()
    s.advance(args.steps)  # returns after hipStreamSynchronize
    t1 = time.perf_counter()
    wall = t1 - t0
    value = args.steps * stages * cells / wall
    norms = s.norms()
    if not np.all(np.isfinite(norms)):
        print("non-finite state after the timed steps", file=sys.stderr)
        return 1
    ms_stage = s.time_stage_kernel(20)
    achieved = cells * b_stage / (ms_stage * 1e-3) / 1e9
    ms_les = s.time_les(3) if cfg.get("les") else 0.0
    # whole step: the stages and the LES model, over the driver-timed step
    step_bytes = stages * cells * b_stage + (cells * B_LES_STEP if cfg.get("les") else 0)
    step_achieved = step_bytes / (wall / args.steps) / 1e9
    # counter traffic, only when measured on the current stage kernel sources
    traffic, traffic_src = None, "not measured for this workload"
    if os.path.exists(args.traffic):
        try:
            with open(args.traffic) as f:
                tr = json.load(f)
            if tkey in tr:
                if tr[tkey].get("source_sha") == source_sha():
                    traffic = tr[tkey]["bytes_per_launch"]
                    traffic_src = "rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated, sources %s" % source_sha()
                else:
                    traffic_src = "stale (measured on other stage kernel sources)"
        except (OSError, ValueError, KeyError):
            traffic = None
    cpu = None if args.no_cpu else cpu_baseline(args, args.case)
    out = {
        "metric": "cell-stage updates/s",
        "value": value,
        "unit": "cell-stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32-bulk/f64",
        "data": "synthetic",
        "config": dict({"workload": workload, "nodes": [cfg["mx"], cfg["my"], cfg["mz"]], "cells": cells,
                        "stages_per_step": stages, "dt": cfg["dt"], "precision": args.precision}, **extra),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "stage_kernel", "kernel_ms": ms_stage, "bytes_per_cell": b_stage,
                     "step": {"bytes": step_bytes, "achieved": step_achieved,
                              "frac": step_achieved / HBM_PEAK_GBS,
                              "bytes_per_cell_stage": step_bytes / (stages * cells)}},
        "les_model_ms_per_step": ms_les,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
