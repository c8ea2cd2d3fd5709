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


# ---------------------------------------------------------------- immersed body

def xorshift64star(seed, n):
    """n uniform numbers in [0, 1) from the xorshift64* generator
    (x ^= x >> 12; x ^= x << 25; x ^= x >> 27; out = x * 2685821657736338717)."""
    mask = (1 << 64) - 1
    x = seed & mask or 1
    out = np.empty(n)
    for q in range(n):
        x ^= x >> 12
        x ^= (x << 25) & mask
        x ^= x >> 27
        r = (x * 2685821657736338717) & mask
        out[q] = (r >> 11) * (1.0 / (1 << 53))
    return out


def icosphere(level, radius=0.5, center=(0.0, 0.0, 0.0), jitter=0.0, seed=20261203):
    """Closed triangulated sphere: the icosahedron with every triangle split in
    four `level` times (20 * 4**level triangles, 10 * 4**level + 2 vertices),
    vertices on the sphere, triangles ordered so that (p2-p1) x (p3-p1) points
    outward.  jitter > 0 moves every vertex coordinate by jitter * (u - 0.5),
    u from xorshift64*(seed).  Returns (verts (n_v, 3), tris (n_elmt, 3))."""
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, dtype=np.float64) / np.linalg.norm(v) for v in verts]
    tris = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
            (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
            (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache = {}

        def mid(a, b):
            key = (a, b) if a < b else (b, a)
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        new = []
        for a, b, c in tris:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            new += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        tris = new
    V = np.array(verts) * radius
    T = np.array(tris, dtype=np.int32)
    cen = V[T].mean(axis=1)
    nrm = np.cross(V[T[:, 1]] - V[T[:, 0]], V[T[:, 2]] - V[T[:, 0]])
    flip = (nrm * cen).sum(axis=1) < 0
    T[flip] = T[flip][:, [0, 2, 1]]
    if jitter > 0:
        V = V + jitter * (xorshift64star(seed, V.size).reshape(V.shape) - 0.5)
    return V + np.asarray(center, dtype=np.float64), T


def box_surface(lo, hi):
    """Closed axis-aligned box as 12 triangles with outward normals."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    V = np.array([(x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0),
                  (x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)], dtype=np.float64)
    T = np.array([(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4),
                  (2, 3, 7), (2, 7, 6), (1, 2, 6), (1, 6, 5), (3, 0, 4), (3, 4, 7)], dtype=np.int32)
    return V, T


def sphere_channel_case(nx=128, ny=128, nz=192, Lx=6.0, Ly=6.0, Lz=12.0, D=1.0, zc=3.0, level=5,
                        jitter=1e-3, seed=20261203, Re=300.0, Ma=0.3, les=2, isothermal=2, cfl=0.4,
                        bc=(7, 7, 1, 1, 45, 8)):
    """Sphere of diameter D in a channel: x periodic along i (nx unique
    nodes), walls at y = 0 and y = Ly along j (ny nodes, tanh clustering
    with stretching 2.2), inflow at z = 0 and outflow at z = Lz along k (nz
    nodes; the reference's inflow/outflow conditions act on k and rhoW).
    Uniform inflow rho = 1, w = 1, T = 1 as the initial state; the sphere is
    an icosphere of 20 * 4**level triangles jittered by jitter * D.
    Returns (cfg, coor, Q, verts, tris)."""
    x = (np.arange(nx + 2) - 1.0) * (Lx / nx)
    eta = np.linspace(0.0, 1.0, ny)
    y = Ly * 0.5 * (1 + np.tanh(2.2 * (eta - 0.5)) / np.tanh(1.1))
    z = np.linspace(0.0, Lz, nz)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    coor = np.stack([X, Y, Z], axis=-1)
    p0 = 1.0 / (GAMMA * Ma * Ma)
    one = np.ones(X.shape)
    Q = conserved(one, 0 * one, 0 * one, one, p0 * one, GAMMA)
    verts, tris = icosphere(level, 0.5 * D, (0.5 * Lx, 0.5 * Ly, zc), jitter * D, seed)
    h = min(Lx / nx, float(np.diff(y).min()), Lz / (nz - 1))
    cfg = dict(mx=nx + 2, my=ny, mz=nz, bctype=tuple(bc), invicid=0, les=les, fluxorder=3, gamma=GAMMA,
               Ma=Ma, ren=Re / D, Pr=0.72, sutherland=1, isothermal=isothermal, Twall=1.0,
               dt=cfl * h / (1.0 + 1.0 / Ma))
    return cfg, coor, Q, verts, tris
