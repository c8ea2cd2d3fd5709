"""Grids, initial states and run options of the benchmark and test cases.

Node convention (periodic directions): a direction with N unique nodes has
m = N + 2 nodes; node 0 is the image of node m-2 and node m-1 the image of
node 1 (Chimera_Comp_LES rhs.c:1107-1227 copies them that way).  Coordinates
of node n are xi_n = (n - 1) * h, h = L / N, so the seam nodes carry
translated coordinates and the metric stencils at nodes 1 and m-2 are
regular.

All arrays are in the DMDA global layout of the reference: shape
(mz, my, mx[, ncomp]) with i fastest, conserved variables interleaved as
(rho, rhoU, rhoV, rhoW, rhoE).
"""
import numpy as np

GAMMA = 1.4


def _axis(N, L):
    h = L / N
    return (np.arange(N + 2, dtype=np.float64) - 1.0) * h, h


def fill_seams(a):
    """Make a node-periodic field: node 0 <- node m-2, node m-1 <- node 1, per axis."""
    a = np.array(a, copy=True)
    for ax in range(3):
        sl0 = [slice(None)] * a.ndim
        src = [slice(None)] * a.ndim
        sl0[ax], src[ax] = 0, a.shape[ax] - 2
        a[tuple(sl0)] = a[tuple(src)]
        sl0[ax], src[ax] = a.shape[ax] - 1, 1
        a[tuple(sl0)] = a[tuple(src)]
    return a


def cartesian_grid(N, L=1.0):
    x, h = _axis(N, L)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return np.stack([X, Y, Z], axis=-1), h


def wavy_grid(N, L=1.0, amp=0.1):
    """x_i = xi_i + amp*h*sin(2 pi xi_j / L) sin(2 pi xi_k / L): a periodic curvilinear grid."""
    xi, h = _axis(N, L)
    Z, Y, X = np.meshgrid(xi, xi, xi, indexing="ij")
    w = 2.0 * np.pi / L
    x = X + amp * h * np.sin(w * Y) * np.sin(w * Z)
    y = Y + amp * h * np.sin(w * X) * np.sin(w * Z)
    z = Z + amp * h * np.sin(w * X) * np.sin(w * Y)
    return np.stack([x, y, z], axis=-1), h


def conserved(rho, u, v, w, p, gamma=GAMMA):
    Q = np.empty(rho.shape + (5,))
    Q[..., 0] = rho
    Q[..., 1] = rho * u
    Q[..., 2] = rho * v
    Q[..., 3] = rho * w
    Q[..., 4] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return Q


def isentropic_vortex(coor, mach=0.5, beta=5.0, radius=0.1, center=(0.5, 0.5), gamma=GAMMA):
    """Shu's isentropic vortex (strength beta, rho_inf = p_inf = 1, T = p/rho),
    advected along x at Mach `mach`; lengths scaled by `radius`; invariant in z."""
    x, y = coor[..., 0], coor[..., 1]
    dx = (x - center[0]) / radius
    dy = (y - center[1]) / radius
    r2 = dx * dx + dy * dy
    u_inf = mach * np.sqrt(gamma)
    f = beta / (2.0 * np.pi) * np.exp(0.5 * (1.0 - r2))
    u = u_inf - f * dy
    v = f * dx
    w = np.zeros_like(u)
    T = 1.0 - (gamma - 1.0) * beta * beta / (8.0 * gamma * np.pi ** 2) * np.exp(1.0 - r2)
    rho = T ** (1.0 / (gamma - 1.0))
    p = rho * T
    return fill_seams(conserved(rho, u, v, w, p, gamma))


def taylor_green(coor, Ma=0.1, rho0=1.0, gamma=GAMMA, variant="velocity"):
    """Taylor-Green vortex on [0, 2 pi]^3.

    variant="velocity": u = sin x cos y cos z, v = -cos x sin y cos z, w = 0,
      p = p0 + rho0/16 (cos 2x + cos 2y)(cos 2z + 2), p0 = rho0/(gamma Ma^2),
      rho = rho0 p / p0 (uniform temperature).
    variant="reference": -init1 9 of the reference (rhs.c:278-291): the
      momentum (not velocity) takes the sine profile, p0 = 100, rho = gamma p Ma^2.
    """
    x, y, z = coor[..., 0], coor[..., 1], coor[..., 2]
    s = (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    if variant == "reference":
        p = 100.0 + s / 16.0
        rho = gamma * p * Ma * Ma
        Q = np.empty(x.shape + (5,))
        Q[..., 0] = rho
        Q[..., 1] = np.sin(x) * np.cos(y) * np.cos(z)
        Q[..., 2] = -np.cos(x) * np.sin(y) * np.cos(z)
        Q[..., 3] = 0.0
        Q[..., 4] = p / (gamma - 1.0) + 0.5 * (Q[..., 1] ** 2 + Q[..., 2] ** 2) / rho
        return fill_seams(Q)
    p0 = rho0 / (gamma * Ma * Ma)
    p = p0 + rho0 / 16.0 * s
    rho = rho0 * p / p0
    u = np.sin(x) * np.cos(y) * np.cos(z)
    v = -np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u)
    return fill_seams(conserved(rho, u, v, w, p, gamma))


def cfl_dt(Q, h, cfl, gamma=GAMMA):
    """Fixed time step from a CFL number on a (near-)uniform grid of spacing h:
    dt = cfl * h / max(|u| + |v| + |w| + 3 c)."""
    rho = Q[..., 0]
    u, v, w = Q[..., 1] / rho, Q[..., 2] / rho, Q[..., 3] / rho
    p = (gamma - 1.0) * (Q[..., 4] - 0.5 * rho * (u * u + v * v + w * w))
    c = np.sqrt(gamma * p / rho)
    return float(cfl * h / np.max(np.abs(u) + np.abs(v) + np.abs(w) + 3.0 * c))


PERIODIC = (7, 7, 7, 7, 7, 7)


def vortex_case(N=64, steps=100, cfl=0.5, mach=0.5, curvilinear=True):
    """Inviscid isentropic vortex on the periodic unit cube (WENO + Rusanov, SSP-RK3)."""
    coor, h = wavy_grid(N) if curvilinear else cartesian_grid(N)
    Q = isentropic_vortex(coor, mach=mach)
    cfg = dict(mx=N + 2, my=N + 2, mz=N + 2, bctype=PERIODIC, invicid=1, les=0, fluxorder=3,
               gamma=GAMMA, Ma=mach, dt=cfl_dt(Q, h, cfl), steps=steps, name="vortex%d" % N)
    return cfg, coor, Q


def tgv_case(N=256, steps=500, cfl=0.4, Re=1600.0, Ma=0.1, Pr=0.71, les=2, fluxorder=3,
             variant="velocity"):
    """Viscous Taylor-Green vortex, Sutherland viscosity, dynamic Smagorinsky."""
    coor, h = cartesian_grid(N, 2.0 * np.pi)
    Q = taylor_green(coor, Ma=Ma, variant=variant)
    cfg = dict(mx=N + 2, my=N + 2, mz=N + 2, bctype=PERIODIC, invicid=0, les=les,
               fluxorder=fluxorder, gamma=GAMMA, Ma=Ma, ren=Re, Pr=Pr, sutherland=1,
               Tref=273.15, dt=cfl_dt(Q, h, cfl), steps=steps, name="tgv%d" % N)
    return cfg, coor, Q


def interior_cells(cfg):
    """Unique interior nodes updated per stage (periodic seams not counted)."""
    return (cfg["mx"] - 2) * (cfg["my"] - 2) * (cfg["mz"] - 2)
