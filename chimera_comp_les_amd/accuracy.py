"""Accuracy of the FP32-bulk form against the FP64 path on the Taylor-Green
vortex (Re 1600, Ma 0.1, Pr 0.71, Sutherland, dynamic Smagorinsky, WENO3 /
LES central flux, SSP-RK3).

Both forms start from the same state on one device and advance in lock step.
Every step records the kinetic energy K = sum(rho |u|^2 / 2 / aj) / V of each;
every `every` steps the dissipation rate eps = -dK/dt (central difference over
the neighbouring steps) and the relative L2 errors of rho, rho*u (vector) and
rho*E over the unique interior nodes are reported.

    python -m chimera_comp_les_amd.accuracy --N 384 --steps 1000 --every 50 --out report.md
"""
import argparse
import json
import sys
import time

import numpy as np

import chimera_comp_les_amd as cl
from chimera_comp_les_amd import cases


def rel_l2(a, b):
    a = a[1:-1, 1:-1, 1:-1]
    b = b[1:-1, 1:-1, 1:-1]
    out = {}
    for name, sl in (("rho", np.s_[..., 0]), ("rhoU", np.s_[..., 1:4]), ("rhoE", np.s_[..., 4])):
        out[name] = float(np.linalg.norm((a[sl] - b[sl]).ravel()) / np.linalg.norm(b[sl].ravel()))
    return out


def lockstep(N, steps, every, log=None):
    """Advance the FP64 and FP32-bulk forms of the N^3 Taylor-Green LES case
    in lock step; returns (rows, summary): per sample step the dissipation
    rates, their relative error and the field errors."""
    cfg, coor, Q = cases.tgv_case(N=N, fluxorder=3)
    s64 = cl.Solver(cfg, coor, precision=cl.PREC_FP64)
    s32 = cl.Solver(cfg, coor, precision=cl.PREC_FP32BULK)
    try:
        s64.set_state(Q)
        s32.set_state(Q)
        del Q, coor
        vol = (2 * np.pi) ** 3
        dt = cfg["dt"]
        K64, K32, t64, t32 = [s64.integrals()[5] / vol], [s32.integrals()[5] / vol], 0.0, 0.0
        fields = {}
        for n in range(1, steps + 2):
            a = time.perf_counter()
            s64.advance(1)
            b = time.perf_counter()
            s32.advance(1)
            c = time.perf_counter()
            t64 += b - a
            t32 += c - b
            K64.append(s64.integrals()[5] / vol)
            K32.append(s32.integrals()[5] / vol)
            if n % every == 0 and n <= steps:
                fields[n] = rel_l2(s32.state()[0], s64.state()[0])
                if log:
                    log(json.dumps({"step": n, **fields[n]}))
    finally:
        s64.close()
        s32.close()
    rows = []
    for n in range(every, steps + 1, every):
        e64 = -(K64[n + 1] - K64[n - 1]) / (2 * dt)
        e32 = -(K32[n + 1] - K32[n - 1]) / (2 * dt)
        rows.append({"step": n, "t": n * dt, "K64": K64[n], "eps64": e64, "eps32": e32,
                     "eps_rel_err": abs(e32 - e64) / abs(e64), "K_rel_err": abs(K32[n] - K64[n]) / K64[n],
                     **{"l2_" + k: v for k, v in fields[n].items()}})
    summary = {"N": N, "steps": steps, "dt": dt,
               "max_eps_rel_err": max(r["eps_rel_err"] for r in rows),
               "max_l2_rho": max(r["l2_rho"] for r in rows), "max_l2_rhoU": max(r["l2_rhoU"] for r in rows),
               "max_l2_rhoE": max(r["l2_rhoE"] for r in rows),
               "s_per_step_fp64": t64 / (steps + 1), "s_per_step_fp32bulk": t32 / (steps + 1)}
    return rows, summary


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=384)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    rows, summary = lockstep(args.N, args.steps, args.every, log=lambda m: print(m, flush=True))
    print(json.dumps(summary), flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write("# FP32-bulk vs FP64: Taylor-Green %d^3, %d steps\n\n" % (args.N, args.steps))
            f.write("Command: `python -m chimera_comp_les_amd.accuracy --N %d --steps %d --every %d`\n\n"
                    % (args.N, args.steps, args.every))
            f.write("| step | t | K (FP64) | eps FP64 | eps FP32-bulk | eps rel. err | K rel. err | "
                    "L2 rho | L2 rhoU | L2 rhoE |\n|---:|---:|---:|---:|---:|---:|---:|---:|---:|---:|\n")
            for r in rows:
                f.write("| %d | %.4f | %.8f | %.6e | %.6e | %.2e | %.2e | %.2e | %.2e | %.2e |\n"
                        % (r["step"], r["t"], r["K64"], r["eps64"], r["eps32"], r["eps_rel_err"], r["K_rel_err"],
                           r["l2_rho"], r["l2_rhoU"], r["l2_rhoE"]))
            f.write("\nSummary: `%s`\n" % json.dumps(summary))
    return 0


if __name__ == "__main__":
    sys.exit(main())
