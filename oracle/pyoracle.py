"""ctypes binding of the CPU restatement (oracle/liboracle.so).

Test infrastructure only: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the reference the results are
compared with, never as the thing measured or shipped.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class V3(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class OrcParams(C.Structure):
    _fields_ = [
        ("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int),
        ("bctype", C.c_int * 6),
        ("ren", C.c_double), ("Ma", C.c_double), ("Pr", C.c_double),
        ("gamma", C.c_double), ("dt", C.c_double), ("st", C.c_double),
        ("Tref", C.c_double), ("max_cs", C.c_double), ("Twall", C.c_double),
        ("s_x", C.c_double), ("s_y", C.c_double), ("s_z", C.c_double),
        ("s_x_min", C.c_double), ("s_y_min", C.c_double), ("s_z_min", C.c_double),
        ("sigma_thrsh", C.c_double), ("ratio", C.c_double),
        ("les", C.c_int), ("invicid", C.c_int), ("fluxorder", C.c_int),
        ("eno_order", C.c_int), ("central", C.c_int), ("sutherland", C.c_int),
        ("isothermal", C.c_int), ("calc_sigma", C.c_int), ("abc", C.c_int),
        ("twoD", C.c_int), ("testfilter_ik", C.c_int),
        ("tistart", C.c_int), ("rstart", C.c_int),
        ("i_homo_filter", C.c_int), ("j_homo_filter", C.c_int), ("k_homo_filter", C.c_int),
        ("weno5", C.c_int), ("roe", C.c_int),
    ]


IBM_INFO_DTYPE = np.dtype(
    [(n, np.int32) for n in ("ni", "nj", "nk", "cell", "imode", "i1", "j1", "k1", "i2", "j2", "k2",
                             "i3", "j3", "k3")]
    + [("pmin", np.float64, (3,))]
    + [(n, np.float64) for n in ("d_s", "d_i", "cs1", "cs2", "cs3", "cr1", "cr2", "cr3")],
    align=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcParams), dp]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_threads.argtypes = [C.c_int]
        for name in ("orc_set_nvert", "orc_set_state"):
            getattr(L, name).argtypes = [C.c_void_p, dp]
        L.orc_get_state.argtypes = [C.c_void_p, dp, dp]
        L.orc_get_metrics.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.orc_get_les.argtypes = [C.c_void_p, dp, dp, dp]
        L.orc_set_time.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_time.argtypes = [C.c_void_p]
        L.orc_get_time.restype = C.c_int
        L.orc_eq_of_state.argtypes = [C.c_void_p]
        L.orc_form_rhs.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_form_bcs.argtypes = [C.c_void_p, C.c_double]
        L.orc_bcs_tail.argtypes = [C.c_void_p]
        L.orc_runge_kutta.argtypes = [C.c_void_p]
        L.orc_les_update.argtypes = [C.c_void_p]
        L.orc_flow_step.argtypes = [C.c_void_p]
        L.orc_do_averaging.argtypes = [C.c_void_p]
        L.orc_ibm_forces.argtypes = [C.c_void_p, dp, dp]
        L.orc_get_averages.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.orc_norms.argtypes = [C.c_void_p, dp]
        L.orc_les_gradients.argtypes = [C.c_void_p, dp, dp]
        L.orc_les_gradients_literal.argtypes = [C.c_void_p, dp, dp]
        L.orc_les_debug.argtypes = [C.c_void_p, dp, dp]
        for name in ("orc_weno_ql", "orc_weno_qr", "orc_muscl_ql", "orc_muscl_qr"):
            f = getattr(L, name)
            f.restype = C.c_double
            f.argtypes = [C.c_double] * 3
        for name in ("orc_eno_ql", "orc_eno_qr"):
            f = getattr(L, name)
            f.restype = C.c_double
            f.argtypes = [C.c_double] * 3 + [C.c_int]
        L.orc_weno5_ql.restype = C.c_double
        L.orc_weno5_ql.argtypes = [dp]
        L.orc_roe.argtypes = [dp, C.c_double, dp, C.c_double, dp, C.c_double, dp]
        L.orc_sigma_cent.restype = C.c_double
        L.orc_sigma_cent.argtypes = [C.c_double] * 5
        L.orc_minmod.restype = C.c_double
        L.orc_minmod.argtypes = [C.c_double, C.c_double]
        L.orc_fp_2.restype = C.c_double
        L.orc_fp_2.argtypes = [dp, dp]
        L.orc_fp4.restype = C.c_double
        L.orc_fp4.argtypes = [dp]
        L.orc_rusanov.argtypes = [dp, C.c_double, dp, C.c_double, dp, dp]
        ip = C.POINTER(C.c_int)
        L.orc_get_nvert.argtypes = [C.c_void_p, dp]
        L.orc_ibm_search.restype = C.c_int
        L.orc_ibm_search.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, C.c_int, ip, ip, ip, dp]
        L.orc_ibm_count.restype = C.c_int
        L.orc_ibm_count.argtypes = [C.c_void_p]
        L.orc_ibm_get_list.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_ibm_set_velocity.argtypes = [C.c_void_p, dp, dp]
        L.orc_ibm_interpolation.argtypes = [C.c_void_p, C.c_int]
        L.orc_intsect_triangle.restype = C.c_int
        L.orc_intsect_triangle.argtypes = [dp] * 5 + [dp] * 3
        L.orc_point_in_triangle.restype = C.c_int
        L.orc_point_in_triangle.argtypes = [V3] * 4 + [C.c_double] * 3
        L.orc_pow15.restype = C.c_double
        L.orc_pow15.argtypes = [C.c_double]
        L.orc_glibc_first_rand.restype = C.c_double
        L.orc_glibc_first_rand.argtypes = [C.c_int]
        L.orc_jacobi3.argtypes = [dp, dp]
        L.orc_q_criteria.argtypes = [C.c_int] * 3 + [dp] * 8
        _LIB = L
    return _LIB


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(cfg):
    """cfg: dict with the option names of OrcParams (missing -> reference defaults)."""
    p = OrcParams()
    defaults = dict(ren=3000.0, Ma=2.81, Pr=0.72, gamma=1.4, dt=0.01, st=1.0, Tref=273.15,
                    max_cs=0.05, Twall=1.0, s_x=1.0, s_y=1.0, s_z=1.0, s_x_min=0.0,
                    s_y_min=0.0, s_z_min=0.0, sigma_thrsh=1e-5, ratio=0.0, les=0, invicid=0,
                    fluxorder=3, eno_order=2, central=2, sutherland=0, isothermal=0,
                    calc_sigma=0, abc=0, twoD=0, testfilter_ik=0, tistart=0, rstart=0,
                    i_homo_filter=0, j_homo_filter=0, k_homo_filter=0, weno5=0, roe=0)
    d = dict(defaults)
    d.update({k: v for k, v in cfg.items() if k in defaults})
    p.mx, p.my, p.mz = cfg["mx"], cfg["my"], cfg["mz"]
    for n, b in enumerate(cfg["bctype"]):
        p.bctype[n] = int(b)
    for k, v in d.items():
        setattr(p, k, v)
    return p


class Oracle:
    """Owns one orc_ctx.  Arrays use the global DMDA layout (k, j, i, comp)."""

    def __init__(self, cfg, coor, nvert=None, threads=None):
        self.L = lib()
        if threads:
            self.L.orc_set_threads(int(threads))
        self.cfg = cfg
        self.shape = (cfg["mz"], cfg["my"], cfg["mx"])
        self.p = make_params(cfg)
        coor = np.ascontiguousarray(coor, dtype=np.float64)
        assert coor.shape == self.shape + (3,)
        self.h = self.L.orc_create(C.byref(self.p), _ptr(coor))
        if nvert is not None:
            nv = np.ascontiguousarray(nvert, dtype=np.float64)
            assert nv.shape == self.shape
            self.L.orc_set_nvert(self.h, _ptr(nv))

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        assert Q.shape == self.shape + (5,)
        self.L.orc_set_state(self.h, _ptr(Q))

    def state(self):
        Q = np.empty(self.shape + (5,))
        P = np.empty(self.shape)
        self.L.orc_get_state(self.h, _ptr(Q), _ptr(P))
        return Q, P

    def metrics(self):
        csi = np.empty(self.shape + (3,))
        eta = np.empty(self.shape + (3,))
        zet = np.empty(self.shape + (3,))
        aj = np.empty(self.shape)
        self.L.orc_get_metrics(self.h, _ptr(csi), _ptr(eta), _ptr(zet), _ptr(aj))
        return csi, eta, zet, aj

    def les_fields(self):
        cs = np.empty(self.shape)
        nut = np.empty(self.shape)
        sig = np.empty(self.shape + (3,))
        self.L.orc_get_les(self.h, _ptr(cs), _ptr(nut), _ptr(sig))
        return cs, nut, sig

    def form_rhs(self, istage=0):
        r = np.empty(self.shape + (5,))
        self.L.orc_form_rhs(self.h, int(istage), _ptr(r))
        return r

    def eq_of_state(self):
        self.L.orc_eq_of_state(self.h)

    def runge_kutta(self):
        self.L.orc_runge_kutta(self.h)

    def form_bcs(self, alfa):
        """FormBCS(user, alfa) on the current Q (= lQ = Qold after set_state),
        then the G2L and EqOfState that RungeKutta runs after it."""
        self.L.orc_form_bcs(self.h, float(alfa))
        self.L.orc_bcs_tail(self.h)

    def les_update(self):
        self.L.orc_les_update(self.h)

    def step(self, n=1, averaging=False):
        """n time steps (Flow_Solver); with averaging, Do_averaging after each."""
        for _ in range(n):
            self.L.orc_flow_step(self.h)
            if averaging:
                self.L.orc_do_averaging(self.h)

    def ibm_forces(self, center=(0.0, 0.0, 0.0)):
        """Calc_forces_SI2: F(3), Cp(3), A(3), A_tot, M(3), Mdpdn(3), Ia(3)."""
        c = np.ascontiguousarray(center, dtype=np.float64)
        out = np.empty(19)
        self.L.orc_ibm_forces(self.h, _ptr(c), _ptr(out))
        return out

    def averages(self):
        """Do_averaging sums: Q (5), P, u^2 (3), (uv, vw, wu) per node."""
        q = np.empty(self.shape + (5,))
        p = np.empty(self.shape)
        u2 = np.empty(self.shape + (3,))
        uc = np.empty(self.shape + (3,))
        self.L.orc_get_averages(self.h, _ptr(q), _ptr(p), _ptr(u2), _ptr(uc))
        return q, p, u2, uc

    @property
    def ti(self):
        return self.L.orc_get_time(self.h)

    @ti.setter
    def ti(self, v):
        self.L.orc_set_time(self.h, int(v))

    def norms(self):
        out = np.empty(2)
        self.L.orc_norms(self.h, _ptr(out))
        return out

    def les_update_debug(self):
        """les_update() and return (LM, MM) of the dynamic model."""
        lm = np.zeros(self.shape)
        mm = np.zeros(self.shape)
        self.L.orc_les_debug(self.h, _ptr(lm), _ptr(mm))
        self.L.orc_les_update(self.h)
        self.L.orc_les_debug(self.h, None, None)
        return lm, mm

    def les_gradients(self, literal=False):
        A = np.empty(self.shape + (9,))
        S = np.empty(self.shape)
        f = self.L.orc_les_gradients_literal if literal else self.L.orc_les_gradients
        f(self.h, _ptr(A), _ptr(S))
        return A, S

    def nvert(self):
        nv = np.empty(self.shape)
        self.L.orc_get_nvert(self.h, _ptr(nv))
        return nv

    def ibm_search(self, verts, tris, nvert_o=None):
        """Classify the grid against a closed triangulated surface and build the
        interpolation list.  verts (n_v, 3), tris (n_elmt, 3) 0-based."""
        v = np.ascontiguousarray(np.asarray(verts, dtype=np.float64).T)
        t = np.ascontiguousarray(np.asarray(tris, dtype=np.int32).T)
        self._mesh = (v, t)
        no = None if nvert_o is None else _ptr(np.ascontiguousarray(nvert_o, dtype=np.float64))
        n = self.L.orc_ibm_search(self.h, v.shape[1], _ptr(v[0]), _ptr(v[1]), _ptr(v[2]), t.shape[1],
                                  _iptr(t[0]), _iptr(t[1]), _iptr(t[2]), no)
        if n < 0:
            raise RuntimeError("orc_ibm_search failed (%d)" % n)
        return n

    def ibm_list(self):
        n = self.L.orc_ibm_count(self.h)
        out = np.zeros(n, dtype=IBM_INFO_DTYPE)
        if n:
            self.L.orc_ibm_get_list(self.h, out.ctypes.data_as(C.c_void_p))
        return out

    def ibm_set_velocity(self, u, uold):
        u = np.ascontiguousarray(u, dtype=np.float64)
        uold = np.ascontiguousarray(uold, dtype=np.float64)
        self.L.orc_ibm_set_velocity(self.h, _ptr(u), _ptr(uold))

    def ibm_interpolation(self, add_dundt=1):
        self.L.orc_ibm_interpolation(self.h, int(add_dundt))


def intsect_triangle(orig, d, v0, v1, v2):
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (orig, d, v0, v1, v2)]
    t, u, v = (np.zeros(1) for _ in range(3))
    r = lib().orc_intsect_triangle(*[_ptr(x) for x in a], _ptr(t), _ptr(u), _ptr(v))
    return r, float(t[0]), float(u[0]), float(v[0])


def point_in_triangle(p, p1, p2, p3, n):
    return lib().orc_point_in_triangle(V3(*p), V3(*p1), V3(*p2), V3(*p3), *[float(x) for x in n])


def glibc_first_rand(seed):
    return lib().orc_glibc_first_rand(int(seed))


def pow15(x):
    return lib().orc_pow15(float(x))


def jacobi3(a):
    """jacobi + eigsrt (data.c:5057-5203) of a 3x3 matrix: the eigenvalue
    estimates in descending order."""
    a = np.ascontiguousarray(a, dtype=np.float64).reshape(3, 3)
    d = np.empty(3)
    lib().orc_jacobi3(_ptr(a), _ptr(d))
    return d


def q_criteria(Q, P, csi, eta, zet, aj, nvert):
    """QCriteria (data.c:4815-5053) on global-layout arrays: returns (Q
    criterion with the pressure kept on boundary nodes, lambda2)."""
    mz, my, mx = Q.shape[:3]
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (Q, csi, eta, zet, aj, nvert)]
    qc = np.array(P, dtype=np.float64, order="C")
    l2 = np.empty((mz, my, mx))
    lib().orc_q_criteria(mx, my, mz, *[_ptr(x) for x in arrs], _ptr(qc), _ptr(l2))
    return qc, l2
