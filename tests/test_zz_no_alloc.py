"""The allocation-free surface-force entry point (cles_ibm_forces_noalloc)
against cles_ibm_forces, through the C ABI: bitwise equal outputs on two
calls in a row, the second on a body with fewer interpolation nodes (the
static scratch keeps the first call's records), and the node limit."""
import numpy as np
import pytest

import chimera_comp_les_amd as cl
from chimera_comp_les_amd import cases

pytestmark = pytest.mark.gpu


def _body(nx, ny, nz, level, D=1.0, steps=2):
    cfg, coor, Q, V, T = cases.sphere_channel_case(nx=nx, ny=ny, nz=nz, Lx=2.4, Ly=2.4, Lz=3.2, zc=1.4,
                                                   D=D, level=level, jitter=1e-3, isothermal=2, les=2)
    g = cl.Solver(cfg, coor)
    g.set_state(Q)
    n = g.ibm_search(V, T)
    g.advance(steps)   # a flow with a CalcUCat velocity and a pressure field
    return g, n


def test_noalloc_forces_equal_the_allocating_entry_point():
    center = (1.2, 1.2, 1.4)
    runs = []
    for shape in ((16, 16, 24, 2), (12, 12, 16, 1)):
        g, n = _body(*shape)
        a = g.ibm_forces_noalloc(center)
        b = g.ibm_forces(center)
        c = g.ibm_forces_noalloc(center)
        assert np.array_equal(a, b) and np.array_equal(c, b)
        assert np.any(b != 0.0)
        runs.append(n)
        g.close()
    assert 0 < runs[1] < runs[0] <= cl.FORCE_NOALLOC_MAX_NODES


def test_noalloc_forces_reject_oversized_lists():
    """a body whose list exceeds the static scratch is refused with
    CLES_EINVAL before any device work; the allocating form still runs"""
    g, n = _body(128, 128, 160, 3, D=2.0, steps=0)
    assert n > cl.FORCE_NOALLOC_MAX_NODES
    with pytest.raises(RuntimeError, match="static scratch"):
        g.ibm_forces_noalloc((1.2, 1.2, 1.4))
    assert np.all(np.isfinite(g.ibm_forces((1.2, 1.2, 1.4))))
    g.close()
