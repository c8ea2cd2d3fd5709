"""Shared helpers of the parity tests: small cases and tolerance checks.

Tolerances (FP64 path vs the CPU restatement, a 1e-12 relative bar):
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


def channel_case(nx=8, ny=10, nz=12, bc=(7, 7, 1, 1, 45, 8), iso=0, les=0, invicid=0, Ma=0.3, seed=0,
                 height=1.0):
    """Stretched channel of the given height: periodic in i, walls in j,
    inflow/outflow in k (the only orientation the reference supports), uniform
    flow along k with a seeded perturbation."""
    rng = np.random.default_rng(seed)
    xi = (np.arange(nx + 2) - 1.0) / nx
    eta = np.linspace(0.0, 1.0, ny + 2)
    y = height * 0.5 * (1 + np.tanh(2.2 * (eta - 0.5)) / np.tanh(1.1))   # tanh clustering at the walls
    z = np.linspace(0.0, 2.0, nz + 2)
    Z, Y, X = np.meshgrid(z, y, xi, indexing="ij")
    coor = np.stack([X, Y, Z], axis=-1)
    gamma = 1.4
    rho = 1.0 + 1e-2 * rng.standard_normal(X.shape)
    w = 1.0 + 1e-2 * rng.standard_normal(X.shape)
    u = 1e-2 * rng.standard_normal(X.shape)
    v = 1e-2 * rng.standard_normal(X.shape)
    p = 1.0 / (gamma * Ma * Ma) * (1.0 + 1e-3 * rng.standard_normal(X.shape))
    Q = cases.conserved(rho, u, v, w, p, gamma)
    # node-periodic in i
    Q[:, :, 0] = Q[:, :, -2]
    Q[:, :, -1] = Q[:, :, 1]
    cfg = dict(mx=nx + 2, my=ny + 2, mz=nz + 2, bctype=tuple(bc), invicid=invicid, les=les, fluxorder=3,
               gamma=gamma, Ma=Ma, ren=300.0, Pr=0.72, sutherland=1, isothermal=iso, Twall=1.0,
               dt=2e-4)
    return cfg, coor, Q
