"""GPU parity of the build-only schemes (options weno5 and roe, FP64 form):
WENO5 of Jiang & Shu where the reference uses WENO3, the Roe flux in place
of the Rusanov flux, alone and together, in formConvection and in the
upwind part of the hybrid LES flux; against the oracle's restatement
(pinned in tests/test_oracle_schemes.py) at the FP64 bar."""
import numpy as np
import pytest

import chimera_comp_les_amd as cl
import pyoracle
from chimera_comp_les_amd import cases
from parity_util import (RHS_TOL, STATE_TOL, bar, channel_case, field_err, oracle_envelope, perturbed, rhs_err,
                         small_vortex, state_err)

pytestmark = pytest.mark.gpu

SCHEMES = [(1, 0), (0, 1), (1, 1)]


def _pair(cfg, coor, Q):
    o = pyoracle.Oracle(cfg, coor)
    o.set_state(Q)
    g = cl.Solver(cfg, coor)
    g.set_state(Q)
    return o, g


@pytest.mark.parametrize("weno5,roe", SCHEMES)
def test_schemes_vortex(weno5, roe):
    cfg, coor, Q = small_vortex(shape=(16, 14, 12))
    cfg.update(weno5=weno5, roe=roe)
    Q = perturbed(Q, 3)
    o, g = _pair(cfg, coor, Q)
    e = rhs_err(g.form_rhs(0), o.form_rhs(0))
    assert e <= RHS_TOL, e
    o.step(4)
    g.advance(4)
    Qc, Pc = o.state()
    Qg, Pg = g.state()
    g.close()
    assert state_err(Qg, Qc) <= STATE_TOL
    assert state_err(Pg[..., None], Pc[..., None]) <= STATE_TOL


@pytest.mark.parametrize("weno5,roe", SCHEMES)
def test_schemes_channel(weno5, roe):
    cfg, coor, Q = channel_case(nx=10, ny=12, nz=14, bc=(7, 7, 1, 1, 45, 8), iso=2)
    cfg.update(weno5=weno5, roe=roe)
    o, g = _pair(cfg, coor, Q)
    e = rhs_err(g.form_rhs(0), o.form_rhs(0))
    assert e <= RHS_TOL, e
    o.step(3)
    g.advance(3)
    Qc, Pc = o.state()
    Qg, Pg = g.state()
    g.close()
    assert state_err(Qg, Qc) <= STATE_TOL


def test_schemes_in_the_hybrid_les_flux():
    cfg, coor, Q = cases.tgv_case(N=14)
    cfg.update(s_x=0.5, s_y=0.3, s_z=0.7, weno5=1, roe=1)
    Q = perturbed(Q, 4, amp=1e-2)
    o, g = _pair(cfg, coor, Q)
    e = rhs_err(g.form_rhs(0), o.form_rhs(0))
    assert e <= RHS_TOL, e
    o.step(3)
    g.advance(3)
    Qc, _ = o.state()
    Qg, _ = g.state()
    g.close()
    assert state_err(Qg, Qc) <= STATE_TOL


def test_sod_tube():
    """Sod's shock tube along z with WENO5 + Roe: the device's steps equal
    the oracle's (whose convergence to the exact solution is pinned), each
    conserved field within 1e-12 of its maximum or the oracle's round-off
    envelope (the nonlinear weights at the discontinuities amplify round-off)."""
    import test_oracle_schemes as S
    N = 101
    h = 1.0 / (N - 1)
    z = np.arange(N) * h
    ys = (np.arange(7) - 1.0) * h
    Z, Y, X = np.meshgrid(z, ys, ys, indexing="ij")
    coor = np.stack([X, Y, Z], axis=-1)
    left = Z < 0.5
    Q = cases.conserved(np.where(left, 1.0, 0.125), 0 * Z, 0 * Z, 0 * Z, np.where(left, 1.0, 0.1))
    cfg = dict(mx=7, my=7, mz=N, bctype=(7, 7, 7, 7, 44, 45), invicid=1, les=0, fluxorder=3, gamma=1.4,
               weno5=1, roe=1, dt=0.2 * h)
    def run(Qi):
        o = pyoracle.Oracle(cfg, coor)
        o.set_state(Qi)
        o.step(50)
        out = o.state()[0]
        o.close()
        return [out[..., c] for c in range(5)]
    base, env = oracle_envelope(run, Q)
    g = cl.Solver(cfg, coor)
    g.set_state(Q)
    g.advance(50)
    Qg, _ = g.state()
    g.close()
    for c in (0, 3, 4):
        assert field_err(Qg[..., c], base[c]) <= bar(env[c]), (c, field_err(Qg[..., c], base[c]), env[c])
    ex = S.sod_exact(z, 50 * 0.2 * h)
    assert np.abs(Qg[:, 3, 3, 0] - ex[:, 0]).mean() < 0.02
