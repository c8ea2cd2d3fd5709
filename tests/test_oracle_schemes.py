"""Pins of the oracle's build-only schemes (CPU only): WENO5 of Jiang & Shu
and the Roe flux, selectable beside the reference's WENO3 and Rusanov flux
(cles_params.weno5, cles_params.roe).  No reference function exists for
them, so each is pinned by what the mathematics fixes:

* WENO5: exact on linear data, the finite-volume face value 1/6 of q_j = j^2
  (the linear weights 0.1, 0.6, 0.3 on equal smoothness indicators), the
  smooth stencil at a jump, and fifth-order convergence of the spatial
  operator on a smooth one-dimensional wave given by cell averages;
* Roe: consistency F(Q, Q) = F(Q), an exactly resolved stationary contact,
  Roe's shock property (the flux of two states joined by one shock is the
  upwind side's physical flux, exactly), and convergence to the exact
  solution of Sod's shock tube.
"""
import ctypes as C
import math

import numpy as np
import pytest

import pyoracle
from chimera_comp_les_amd import cases

L = pyoracle.lib()
GAMMA = 1.4


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def weno5(q):
    a = np.ascontiguousarray(q, dtype=np.float64)
    return L.orc_weno5_ql(_dp(a))


def roe(sl, sr, h, gamma=GAMMA):
    """sl, sr: (rho, rhoU, rhoV, rhoW, rhoE, p)"""
    ql = np.ascontiguousarray(sl[:5], dtype=np.float64)
    qr = np.ascontiguousarray(sr[:5], dtype=np.float64)
    hh = np.ascontiguousarray(h, dtype=np.float64)
    out = np.zeros(5)
    L.orc_roe(_dp(ql), float(sl[5]), _dp(qr), float(sr[5]), _dp(hh), float(gamma), _dp(out))
    return out


def phys_flux(s, h):
    U = (h[0] * s[1] + h[1] * s[2] + h[2] * s[3]) / s[0]
    return np.array([s[0] * U, s[1] * U + s[5] * h[0], s[2] * U + s[5] * h[1], s[3] * U + s[5] * h[2],
                     (s[4] + s[5]) * U])


def state(rho, u, p, gamma=GAMMA):
    u = np.asarray(u, dtype=np.float64)
    return np.array([rho, rho * u[0], rho * u[1], rho * u[2], p / (gamma - 1) + 0.5 * rho * u @ u, p])


# ------------------------------------------------------------------ WENO5
def test_weno5_known_answers():
    for a, b in ((1.0, 0.25), (-3.0, 2.0), (100.0, -1e-3)):
        q = a + b * np.arange(-2, 3)
        assert weno5(q) == pytest.approx(a + 0.5 * b, rel=1e-14, abs=1e-14)
    # q_j = j^2 around j = 0: equal indicators, w = d, the finite-volume
    # face value of cell averages of x^2 - 1/12, i.e. 1/6 at x = 1/2
    assert weno5(np.arange(-2, 3) ** 2.0) == pytest.approx(1.0 / 6.0, rel=1e-13)
    # a jump between the last two points: the left stencil (q0..q2) is chosen
    assert weno5([1.0, 1.0, 1.0, 1.0, 50.0]) == pytest.approx(1.0, abs=1e-6)
    assert weno5([1.0, 1.0, 1.0, 50.0, 50.0]) == pytest.approx(1.0, abs=1e-6)


def _wave_rhs(N, weno5_on, roe_on=0):
    """spatial operator of a density wave rho = 1 + 0.2 sin(2 pi x), u = 1,
    p = 1 along x, given as cell averages; returns (x, rhs_rho, exact)"""
    Lx = 1.0
    h = Lx / N
    x = (np.arange(N + 2) - 1.0) * h
    ny = 5
    ys = (np.arange(ny + 2) - 1.0) * h
    Z, Y, X = np.meshgrid(ys, ys, x, indexing="ij")
    coor = np.stack([X, Y, Z], axis=-1)
    k = 2 * np.pi
    fac = math.sin(math.pi * h) / (math.pi * h)            # cell average of sin
    rho = 1.0 + 0.2 * np.sin(k * X) * fac
    Q = cases.conserved(rho, 1.0 + 0 * X, 0 * X, 0 * X, 1.0 + 0 * X)
    # rhoE as a cell average: p/(g-1) + rho u^2/2 is linear in rho here
    Q = cases.fill_seams(Q)
    cfg = dict(mx=N + 2, my=ny + 2, mz=ny + 2, bctype=cases.PERIODIC, invicid=1, les=0, fluxorder=3,
               gamma=GAMMA, weno5=weno5_on, roe=roe_on, dt=1e-3)
    o = pyoracle.Oracle(cfg, coor)
    o.set_state(Q)
    r = o.form_rhs(0)[3, 3, :, 0]
    o.close()
    # exact semi-discrete finite-volume operator: -(rho(x+h/2) - rho(x-h/2)) / h
    exact = -0.2 * (np.sin(k * (x + h / 2)) - np.sin(k * (x - h / 2))) / h
    return x, r, exact


@pytest.mark.parametrize("roe_on", [0, 1])
def test_weno5_fifth_order_spatial_operator(roe_on):
    """Error of the density residual against the exact finite-volume operator
    at nodes away from the seam faces (first order there, as in the
    reference) and from the wave's extrema (where Jiang-Shu weights lose
    order): halving h divides it by about 2^5."""
    errs = []
    for N in (32, 64):
        x, r, ex = _wave_rhs(N, 1, roe_on)
        sel = np.zeros(x.size, bool)
        sel[4:-4] = True
        sel &= np.abs(np.cos(2 * np.pi * x)) > 0.6
        errs.append(np.abs(r - ex)[sel].max())
    order = math.log2(errs[0] / errs[1])
    assert order > 4.5, (errs, order)
    # the reference's truncated WENO3 does not reach that
    x, r, ex = _wave_rhs(64, 0, roe_on)
    sel = np.zeros(x.size, bool)
    sel[4:-4] = True
    sel &= np.abs(np.cos(2 * np.pi * x)) > 0.6
    assert np.abs(r - ex)[sel].max() > 20 * errs[1]


# -------------------------------------------------------------------- Roe
def test_roe_consistency():
    rng = np.random.default_rng(3)
    for _ in range(20):
        s = state(0.5 + rng.random(), rng.standard_normal(3), 0.5 + rng.random())
        h = rng.standard_normal(3)
        assert np.allclose(roe(s, s, h), phys_flux(s, h), rtol=1e-13, atol=1e-13)


def test_roe_resolves_a_stationary_contact():
    """equal pressure, zero velocity, a density jump: no wave but the
    contact, whose speed is zero, so the flux is the exact (p h) -- the
    Rusanov flux would add SA/2 times the jump"""
    h = np.array([0.3, -0.2, 0.9])
    sl, sr = state(1.0, (0, 0, 0), 2.0), state(0.2, (0, 0, 0), 2.0)
    assert np.allclose(roe(sl, sr, h), [0, 2.0 * h[0], 2.0 * h[1], 2.0 * h[2], 0], atol=1e-14)
    out = np.zeros(5)
    L.orc_rusanov(_dp(np.ascontiguousarray(sl[:5])), sl[5], _dp(np.ascontiguousarray(sr[:5])), sr[5],
                  _dp(np.ascontiguousarray(h)), _dp(out))
    assert abs(out[0]) > 0.1


def test_roe_rankine_hugoniot():
    """Roe's shock property: for two states joined by one shock of speed S
    (F(R) - F(L) = S (Q_R - Q_L)) the Roe flux is F(L) when S > 0 and F(R)
    when S < 0, exactly.  States: a Mach 2 normal shock (upstream rho = 1,
    p = 1, downstream from the Rankine-Hugoniot relations) seen from frames
    in which it moves at S = +-0.8 (|S| above the entropy fix's 0.1 c, which
    acts on slow acoustic waves only), with any tangential velocity and any
    face orientation."""
    g = GAMMA
    M1 = 2.0
    u1 = M1 * math.sqrt(g)
    r2 = (g + 1) * M1 ** 2 / ((g - 1) * M1 ** 2 + 2)
    p2 = 1 + 2 * g / (g + 1) * (M1 ** 2 - 1)
    u2 = u1 / r2
    rng = np.random.default_rng(5)
    for S in (-0.8, 0.8, -0.8, 0.8):
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        t = np.cross(n, rng.standard_normal(3))
        t *= 0.3 / np.linalg.norm(t)
        # lab frame: the shock moves at S along n; the flow enters it from the left
        sl, sr = state(1.0, (u1 + S) * n + t, 1.0), state(r2, (u2 + S) * n + t, p2)
        h = 0.7 * n
        fl, fr = phys_flux(sl, h), phys_flux(sr, h)
        assert np.allclose(fr - fl, 0.7 * S * (sr[:5] - sl[:5]), rtol=1e-12, atol=1e-12)
        assert np.allclose(roe(sl, sr, h), fl if S > 0 else fr, rtol=1e-12, atol=1e-12)


def test_roe_moving_contact_and_shear():
    """a contact with a tangential slip moving at the normal velocity V
    (equal pressure and normal velocity, density and tangential velocity
    jump): a single linearly degenerate wave of speed V, so the Roe flux is
    again the upwind side's physical flux -- the entropy and shear terms of
    the dissipation carry it"""
    rng = np.random.default_rng(7)
    for V in (0.6, -0.6, 1.3, -1.3):
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        t1 = np.cross(n, rng.standard_normal(3))
        t2 = np.cross(n, rng.standard_normal(3))
        sl, sr = state(1.0, V * n + 0.4 * t1, 1.5), state(0.3, V * n - 0.7 * t2, 1.5)
        h = 1.3 * n
        fl, fr = phys_flux(sl, h), phys_flux(sr, h)
        assert np.allclose(roe(sl, sr, h), fl if V > 0 else fr, rtol=1e-12, atol=1e-12)


def sod_exact(z, t, z0=0.5, g=GAMMA):
    """exact solution of Sod's problem (rho, u, p) at positions z, time t:
    the star pressure from Newton's method on f_L(p) + f_R(p) + du = 0
    (Toro, Riemann Solvers, ch. 4), then the wave pattern sampled"""
    rL, uL, pL, rR, uR, pR = 1.0, 0.0, 1.0, 0.125, 0.0, 0.1
    cL, cR = math.sqrt(g * pL / rL), math.sqrt(g * pR / rR)

    def f(p, r, pk, ck):
        if p > pk:  # shock
            A, B = 2 / ((g + 1) * r), (g - 1) / (g + 1) * pk
            return (p - pk) * math.sqrt(A / (p + B)), math.sqrt(A / (B + p)) * (1 - (p - pk) / (2 * (B + p)))
        return (2 * ck / (g - 1) * ((p / pk) ** ((g - 1) / (2 * g)) - 1),
                1 / (r * ck) * (p / pk) ** (-(g + 1) / (2 * g)))
    p = 0.5 * (pL + pR)
    for _ in range(50):
        fl, dl = f(p, rL, pL, cL)
        fr, dr = f(p, rR, pR, cR)
        p -= (fl + fr + uR - uL) / (dl + dr)
    fl, _ = f(p, rL, pL, cL)
    fr, _ = f(p, rR, pR, cR)
    us = 0.5 * (uL + uR) + 0.5 * (fr - fl)
    rsL = rL * (p / pL) ** (1 / g)                                           # rarefaction
    rsR = rR * ((p / pR + (g - 1) / (g + 1)) / ((g - 1) / (g + 1) * p / pR + 1))  # shock
    S = uR + cR * math.sqrt((g + 1) / (2 * g) * p / pR + (g - 1) / (2 * g))
    cs = cL * (p / pL) ** ((g - 1) / (2 * g))
    out = np.empty((z.size, 3))
    for n, zz in enumerate(z):
        xi = (zz - z0) / t
        if xi < uL - cL:
            out[n] = rL, uL, pL
        elif xi < us - cs:
            c = 2 / (g + 1) * (cL + (g - 1) / 2 * (uL - xi))
            u = 2 / (g + 1) * (cL + (g - 1) / 2 * uL + xi)
            out[n] = rL * (c / cL) ** (2 / (g - 1)), u, pL * (c / cL) ** (2 * g / (g - 1))
        elif xi < us:
            out[n] = rsL, us, p
        elif xi < S:
            out[n] = rsR, us, p
        else:
            out[n] = rR, uR, pR
    return out


def _sod(N, t_end=0.2):
    """Sod's tube along z (k-min bctype 44 and k-max 45: extrapolation), 5
    periodic nodes in x and y; WENO5 + Roe, SSP-RK3"""
    h = 1.0 / (N - 1)
    z = np.arange(N) * h
    ny = 5
    ys = (np.arange(ny + 2) - 1.0) * h
    Z, Y, X = np.meshgrid(z, ys, ys, indexing="ij")
    coor = np.stack([X, Y, Z], axis=-1)
    left = Z < 0.5
    rho = np.where(left, 1.0, 0.125)
    p = np.where(left, 1.0, 0.1)
    Q = cases.conserved(rho, 0 * Z, 0 * Z, 0 * Z, p)
    steps = int(round(t_end / (0.2 * h)))
    dt = t_end / steps
    cfg = dict(mx=ny + 2, my=ny + 2, mz=N, bctype=(7, 7, 7, 7, 44, 45), invicid=1, les=0, fluxorder=3,
               gamma=GAMMA, weno5=1, roe=1, dt=dt)
    o = pyoracle.Oracle(cfg, coor)
    o.set_state(Q)
    o.step(steps)
    Qo, Po = o.state()
    o.close()
    return z, Qo[:, 3, 3, 0], Po[:, 3, 3]


def test_sod_converges_to_the_exact_solution():
    errs = []
    for N in (101, 201):
        z, rho, p = _sod(N)
        ex = sod_exact(z, 0.2)
        errs.append((np.abs(rho - ex[:, 0]).mean(), np.abs(p - ex[:, 2]).mean()))
    (er1, ep1), (er2, ep2) = errs
    assert er2 < 0.8 * er1 and ep2 < 0.8 * ep1, errs
    assert er2 < 0.01 and ep2 < 0.01, errs
