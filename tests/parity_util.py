"""Shared helpers of the parity tests: small cases and tolerance checks.

Tolerances (FP64 path vs the CPU restatement, BASELINE bar 1e-12):
  * right-hand side: |gpu - cpu| <= 1e-12 * ||cpu_c||_inf per component c
    (the RHS is a difference of face fluxes, so it is normalised by the
    component's scale, not by its own value at cells where it cancels);
  * state after time steps: |gpu - cpu| <= 1e-12 * max(|cpu|, 1e-3 ||cpu_c||_inf)
    per cell and component.
"""
import numpy as np

from chimera_comp_les_amd import cases

RHS_TOL = 1e-12
STATE_TOL = 1e-12


def rhs_err(gpu, cpu):
    scale = np.maximum(np.abs(cpu).reshape(-1, 5).max(axis=0), 1e-300)
    return float((np.abs(gpu - cpu) / scale).max())


def state_err(gpu, cpu):
    flat = np.abs(cpu).reshape(-1, cpu.shape[-1]).max(axis=0)
    floor = np.maximum(1e-3 * flat, 1e-300)
    den = np.maximum(np.abs(cpu), floor)
    return float((np.abs(gpu - cpu) / den).max())


def small_vortex(N=12, curvilinear=True, shape=None):
    """Isentropic vortex on a periodic (curvilinear) grid; shape=(nx,ny,nz) unique nodes."""
    if shape is None:
        cfg, coor, Q = cases.vortex_case(N=N, curvilinear=curvilinear)
        return cfg, coor, Q
    nx, ny, nz = shape
    L = 1.0
    xs = [(np.arange(n + 2) - 1.0) * (L / n) for n in (nx, ny, nz)]
    Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    amp = 0.1
    hx, hy, hz = L / nx, L / ny, L / nz
    w = 2 * np.pi
    x = X + amp * hx * np.sin(w * Y) * np.sin(w * Z)
    y = Y + amp * hy * np.sin(w * X) * np.sin(w * Z)
    z = Z + amp * hz * np.sin(w * X) * np.sin(w * Y)
    coor = np.stack([x, y, z], axis=-1)
    Q = cases.isentropic_vortex(coor)
    cfg = dict(mx=nx + 2, my=ny + 2, mz=nz + 2, bctype=cases.PERIODIC, invicid=1, les=0, fluxorder=3,
               gamma=1.4, Ma=0.5, dt=cases.cfl_dt(Q, min(hx, hy, hz), 0.5))
    return cfg, coor, Q


def perturbed(Q, seed, amp=1e-3):
    """Seeded multiplicative perturbation that keeps the seams node-periodic."""
    rng = np.random.default_rng(seed)
    f = 1.0 + amp * rng.standard_normal(Q.shape[:3] + (1,))
    return cases.fill_seams(Q * f)
