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
library's stream) and the CPU restatement timed on this host.
"""
import argparse
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
# LES model once per step: read Q 40, P 8, metrics 80, nvert 8; write Cs 8, nu_t 8
B_LES_STEP = 152


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--scheme", choices=["weno", "muscl"], default="weno")
    ap.add_argument("--precision", choices=["fp64", "fp32bulk"], default="fp64")
    ap.add_argument("--case", choices=["tgv", "vortex", "sphere"], default="tgv")
    ap.add_argument("--cpu-N", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="oracle steps of the CPU sample (default: 30, 8 for the sphere case)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="rocprofv3 FETCH/WRITE summary used for roofline.traffic")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a 512 MiB device-to-device copy three times (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against a known byte count)")
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


def cpu_baseline(args, workload):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from chimera_comp_les_amd import cases
    threads = min(16, os.cpu_count() or 1)
    N = args.cpu_N
    mesh = None
    if workload == "tgv":
        cfg, coor, Q = cases.tgv_case(N=N, fluxorder=3 if args.scheme == "weno" else 2)
    elif workload == "sphere":
        cfg, coor, Q, V, T = cases.sphere_channel_case()
        mesh = (V, T)
        N = "192x128x128"
    else:
        cfg, coor, Q = cases.vortex_case(N=N)
    o = pyoracle.Oracle(cfg, coor, threads=threads)
    o.set_state(Q)
    if mesh is not None:
        o.ibm_search(*mesh)
    o.step(1)  # warm-up (first step of a fresh run has Cs = 0)
    if args.cpu_steps is None:
        args.cpu_steps = 8 if workload == "sphere" else 30
    t0 = time.perf_counter()
    o.step(args.cpu_steps)
    dt = time.perf_counter() - t0
    cells = cases.interior_cells(cfg)
    rate = args.cpu_steps * cfg["fluxorder"] * cells / dt
    o.close()
    return {"value": rate, "unit": "cell-stage/s", "cores": threads, "kind": "port",
            "sample": "CPU restatement (oracle/, C, -O3, OpenMP %d threads), same workload at %s, "
                      "%d steps after 1 warm-up step, %.2f s" % (threads, N if isinstance(N, str) else "%d^3" % N,
                                                               args.cpu_steps, dt)}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        return 2
    if args.copy_probe:
        return copy_probe()
    import chimera_comp_les_amd as cl
    from chimera_comp_les_amd import cases
    fluxorder = 3 if args.scheme == "weno" else 2
    if args.case == "tgv":
        cfg, coor, Q = cases.tgv_case(N=args.N, fluxorder=fluxorder)
        workload = ("taylor-green %d^3 Re1600 Ma0.1 Pr0.71 Sutherland dynamic-Smagorinsky "
                    "%s SSP-RK%d" % (args.N, "WENO3" if fluxorder == 3 else "MUSCL", fluxorder))
        b_stage = B_STAGE_LES
    elif args.case == "sphere":
        cfg, coor, Q, V, T = cases.sphere_channel_case()
        workload = ("immersed sphere D=1 in a channel, 128x128x192 nodes (x periodic, y walls tanh 2.2, z "
                    "inflow/outflow), icosphere 20480 triangles jittered 1e-3 D, Re_D 300 Ma 0.3 Sutherland "
                    "dynamic-Smagorinsky isothermal-2 walls SSP-RK3")
        b_stage = B_STAGE_LES
    else:
        cfg, coor, Q = cases.vortex_case(N=args.N)
        workload = "isentropic-vortex %d^3 curvilinear inviscid WENO3/Rusanov SSP-RK3" % args.N
        b_stage = B_STAGE_INVISCID
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
    traffic = None
    if os.path.exists(args.traffic):
        try:
            with open(args.traffic) as f:
                tr = json.load(f)
            key = "%s_%s_%s" % (args.case, args.N if args.case != "sphere" else "c4", args.precision)
            if key in tr:
                traffic = tr[key]["bytes_per_launch"]
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
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "stage_kernel", "kernel_ms": ms_stage, "bytes_per_cell": b_stage},
        "les_model_ms_per_step": ms_les,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
