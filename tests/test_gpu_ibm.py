"""GPU parity of the immersed body: classification, interpolation stencils and
time steps with the interpolation, HIP path through the C ABI against the CPU
restatement.

Classification (Nvert) and the stencils -- every integer (node, triangle,
ICP mode, donor indices) and every floating-point field (nearest point,
distances, weights) -- must be bit-identical; the state after time steps
within the FP64 bar of parity_util.  The surface forces of a given state are
bitwise (zero steps); after steps the inputs differ at round-off, and each
output vector (F, Cp, A, A_tot, M, Mdpdn, Ia) is held to 1e-12 of its
magnitude or the oracle's round-off envelope (the moment about the centre of
a nearly symmetric body cancels).
"""
import numpy as np
import pytest

import chimera_comp_les_amd as cl
import pyoracle
from chimera_comp_les_amd import cases
from parity_util import ENVELOPE_FACTOR, STATE_TOL, ULP_NOISE, state_err

pytestmark = pytest.mark.gpu

INT_FIELDS = ("ni", "nj", "nk", "cell", "imode", "i1", "j1", "k1", "i2", "j2", "k2", "i3", "j3", "k3")
FP_FIELDS = ("d_s", "d_i", "cs1", "cs2", "cs3", "cr1", "cr2", "cr3")


FORCE_GROUPS = ((0, 3), (3, 6), (6, 9), (9, 10), (10, 13), (13, 16), (16, 19))
CENTER = (1.2, 1.2, 1.4)


def _group_err(a, b):
    return np.array([np.linalg.norm(a[i:j] - b[i:j]) / max(np.linalg.norm(b[i:j]), 1e-300) for i, j in FORCE_GROUPS])


def oracle_forces(cfg, coor, Q, V, T, steps_list):
    """oracle forces after each number of steps in steps_list, and their
    round-off envelope per output vector (one-ulp perturbations of Q)"""
    def run(Qi):
        o = pyoracle.Oracle(cfg, coor)
        o.set_state(Qi)
        o.ibm_search(V, T)
        out, done = [], 0
        for n in steps_list:
            o.step(n - done)
            done = n
            out.append(o.ibm_forces(CENTER))
        o.close()
        return out
    base = run(Q)
    env = [np.zeros(len(FORCE_GROUPS)) for _ in steps_list]
    for seed in (11, 12):
        rng = np.random.default_rng(seed)
        r = run(Q * (1.0 + ULP_NOISE * rng.standard_normal(Q.shape)))
        env = [np.maximum(e, _group_err(x, y)) for e, x, y in zip(env, r, base)]
    return base, env


def assert_forces(fg, fo, env):
    tol = np.maximum(1e-12, ENVELOPE_FACTOR * env)
    err = _group_err(fg, fo)
    assert np.all(err <= tol), np.c_[err, tol]


def sphere(iso=2, les=0, level=2, jitter=1e-3, twoD=0):
    cfg, coor, Q, V, T = cases.sphere_channel_case(nx=16, ny=16, nz=24, Lx=2.4, Ly=2.4, Lz=3.2, zc=1.4,
                                                   level=level, jitter=jitter, isothermal=iso, les=les)
    if twoD:
        cfg = dict(cfg, twoD=twoD)
    return cfg, coor, Q, V, T


def compare_search(cfg, coor, Q, V, T):
    o = pyoracle.Oracle(cfg, coor)
    o.set_state(Q)
    g = cl.Solver(cfg, coor)
    g.set_state(Q)
    n_o = o.ibm_search(V, T)
    n_g = g.ibm_search(V, T)
    assert n_g == n_o
    assert np.array_equal(g.nvert(), o.nvert())
    Lo, Lg = o.ibm_list(), g.ibm_list()
    for f in INT_FIELDS:
        assert np.array_equal(Lo[f], Lg[f]), f
    for f in FP_FIELDS:
        assert np.array_equal(Lg[f], Lo[f]), f
    assert np.array_equal(Lg["pmin"], Lo["pmin"])
    return o, g


@pytest.mark.parametrize("iso", [2, 1])
def test_search_sphere(iso):
    cfg, coor, Q, V, T = sphere(iso=iso)
    o, g = compare_search(cfg, coor, Q, V, T)
    assert len(g.ibm_list()) > 0
    g.close()


def test_search_box_with_ray_retries():
    # flat faces split along a diagonal: rays through the shared edge are
    # singular and are cast again with the next srand seed
    N = 12
    cfg = dict(mx=N + 2, my=N + 2, mz=N + 2, bctype=cases.PERIODIC, invicid=0, les=0, fluxorder=3,
               gamma=1.4, Ma=0.3, ren=100.0, isothermal=2, Twall=1.0, dt=1e-3)
    coor, _ = cases.cartesian_grid(N, 1.0)
    V, T = cases.box_surface((-0.5, -0.3, -0.4), (1.5, 1.6, 0.3 + 0.37 / N))
    Q = cases.conserved(np.ones(coor.shape[:3]), 0 * coor[..., 0], 0 * coor[..., 0], 0 * coor[..., 0],
                        np.full(coor.shape[:3], 1 / (1.4 * 0.09)), 1.4)
    o, g = compare_search(cfg, coor, Q, V, T)
    g.close()


@pytest.mark.parametrize("iso,les", [(2, 0), (2, 2), (1, 0)])
def test_steps_sphere(iso, les):
    cfg, coor, Q, V, T = sphere(iso=iso, les=les)
    o, g = compare_search(cfg, coor, Q, V, T)
    o.step(3)
    g.advance(3)
    Qc, Pc = o.state()
    Qg, Pg = g.state()
    g.close()
    assert np.all(np.isfinite(Qc))
    assert state_err(Qg, Qc) <= STATE_TOL
    assert state_err(Pg[..., None], Pc[..., None]) <= STATE_TOL


@pytest.mark.parametrize("iso,les,steps", [(2, 0, 0), (2, 0, 3), (2, 2, 2), (1, 0, 2)])
def test_surface_forces_sphere(iso, les, steps):
    """Calc_forces_SI2 on the device: the per-node gather runs in parallel and
    the sums follow the list order in one thread, so the nineteen outputs are
    bitwise those of the oracle after the same steps (zero steps: no CalcUCat
    yet, pressure only)."""
    cfg, coor, Q, V, T = sphere(iso=iso, les=les)
    o, g = compare_search(cfg, coor, Q, V, T)
    o.close()
    fos, env = oracle_forces(cfg, coor, Q, V, T, (steps, steps + 1))
    if steps:
        g.advance(steps)
    fg = g.ibm_forces(CENTER)
    fo = fos[0]
    assert np.all(np.isfinite(fo)) and (steps == 0 or np.any(fo[0:3] != 0))
    if steps == 0:  # the same state: the gather and the ordered fold are bitwise
        assert np.array_equal(fg, fo), np.c_[fg, fo, fg - fo]
    else:
        assert_forces(fg, fo, env[0])
    # once per step, as Struc_Solver calls it: the work buffers are reused
    g.advance(1)
    fg = g.ibm_forces(CENTER)
    g.close()
    assert_forces(fg, fos[1], env[1])


@pytest.mark.parametrize("twoD", [1, 3])
def test_steps_and_forces_sphere_two_d(twoD):
    """-twoD with the immersed body: the interpolation normal and the surface
    velocity drop one component (ibm_iso.c:2662-2664, 2744-2746); with 1 also
    the flow-side rules (no xi fluxes, u = 0). Steps and the nineteen surface
    force outputs against the oracle."""
    cfg, coor, Q, V, T = sphere(iso=2, les=2, twoD=twoD)
    o, g = compare_search(cfg, coor, Q, V, T)
    o.step(2)
    g.advance(2)
    Qc, Pc = o.state()
    Qg, Pg = g.state()
    fos, env = oracle_forces(cfg, coor, Q, V, T, (2,))
    fg = g.ibm_forces(CENTER)
    g.close()
    assert state_err(Qg, Qc) <= STATE_TOL
    assert state_err(Pg[..., None], Pc[..., None]) <= STATE_TOL
    assert_forces(fg, fos[0], env[0])
