"""MI355X (gfx950) right-hand-side and time-advance path of Chimera_Comp_LES.

The product is the C ABI in include/cles.h, implemented by libcles.so (HIP
kernels for gfx950).  This module is a thin ctypes binding of that ABI that
mirrors the reference's call chain for tests and measurement:

    Solver(cfg, coor)            MG_Initial + FormMetrics      (init.c, metrics.c:3)
    .set_state(Q)                tail of SetInitialCondition   (rhs.c:424-433)
    .form_rhs(istage)            formRHS                       (rhs.c:4703)
    .runge_kutta()               RungeKutta                    (solvers.c:223)
    .les_update()                Compute_Smagorinsky_Constant_1 + Compute_eddy_viscosity_LES
                                                               (les.c:135, les.c:1477)
    .flow_step() / .advance(n)   Flow_Solver + Qold copy       (solvers.c:13, main.c:900-974)

There is no CPU fallback: constructing a Solver without libcles.so or without
a HIP device raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcles.so")

LAYOUT_GLOBAL = 0
LAYOUT_LOCAL = 1
PREC_FP64 = 0
PREC_FP32BULK = 1
BATCH_PERSISTENT = 0  # one persistent kernel per group runs every stage of a batched step
BATCH_LAUNCHES = 1    # one batched launch per group and stage (default)
# bodies with more interpolation nodes are refused by cles_ibm_forces_noalloc
FORCE_NOALLOC_MAX_NODES = 16384
# return codes of include/cles.h
OK, EINVAL, ENOMEM, EDEVICE, ESTATE = 0, -1, -2, -3, -4

ERRORS = {0: "ok", -1: "invalid argument", -2: "out of device memory", -3: "HIP runtime error",
          -4: "call out of sequence"}


class CParams(C.Structure):
    """Mirror of struct cles_params (include/cles.h)."""
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
        ("tistart", C.c_int), ("rstart", C.c_int), ("precision", C.c_int),
        ("averaging", C.c_int),
        ("i_homo_filter", C.c_int), ("j_homo_filter", C.c_int), ("k_homo_filter", C.c_int),
        ("weno5", C.c_int), ("roe", C.c_int),
    ]


EXPORTS = [
    "cles_default_params", "cles_create", "cles_destroy", "cles_set_nvert", "cles_set_state",
    "cles_get_state", "cles_get_metrics", "cles_get_les", "cles_eq_of_state", "cles_form_rhs",
    "cles_runge_kutta", "cles_les_update", "cles_flow_step", "cles_advance", "cles_norms",
    "cles_get_time", "cles_set_time", "cles_time_stage_kernel", "cles_time_les",
    "cles_device_bytes", "cles_last_error", "cles_get_nvert", "cles_ibm_search",
    "cles_ibm_set_velocity", "cles_ibm_get_list", "cles_integrals", "cles_footprint",
    "cles_batch_create", "cles_batch_destroy", "cles_batch_add", "cles_batch_remove", "cles_batch_size",
    "cles_batch_used", "cles_batch_set_tile", "cles_batch_advance", "cles_batch_time_step",
    "cles_batch_last_error", "cles_set_tile", "cles_batch_launch", "cles_batch_wait",
    "cles_set_state_qp", "cles_write_fields", "cles_read_fields", "cles_petsc_vec_write", "cles_petsc_vec_read",
    "cles_get_averages", "cles_ibm_forces", "cles_write_vtk", "cles_ibm_forces_noalloc",
    "cles_batch_set_mode", "cles_batch_set_spin_limit", "cles_batch_set_graph", "cles_write_vtk_opts",
    "cles_q_criteria", "cles_write_q_vtk",
]

class CVtkOpts(C.Structure):
    """cles_vtk_opts of include/cles.h: VTKOut's -skip, -theta, -Pres, -airfoil."""
    _fields_ = [("skip", C.c_int), ("theta", C.c_double), ("pres", C.c_int), ("airfoil", C.c_int)]


IBM_INFO_FIELDS = ("ni", "nj", "nk", "cell", "imode", "i1", "j1", "k1", "i2", "j2", "k2", "i3", "j3", "k3")


class CIbmInfo(C.Structure):
    """cles_ibm_info of include/cles.h."""
    _fields_ = [(n, C.c_int) for n in IBM_INFO_FIELDS] + [("pmin", C.c_double * 3)] + [
        (n, C.c_double) for n in ("d_s", "d_i", "cs1", "cs2", "cs3", "cr1", "cr2", "cr3")]


IBM_INFO_DTYPE = np.dtype(
    [(n, np.int32) for n in IBM_INFO_FIELDS] + [("pmin", np.float64, (3,))]
    + [(n, np.float64) for n in ("d_s", "d_i", "cs1", "cs2", "cs3", "cr1", "cr2", "cr3")], align=True)

_LIB = None


def load_library(path=LIB_PATH):
    """Load libcles.so and declare its signatures (no device is touched)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise RuntimeError("libcles.so is not built: run `python -m chimera_comp_les_amd.build`")
    L = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    L.cles_default_params.argtypes = [C.POINTER(CParams), C.c_int, C.c_int, C.c_int]
    L.cles_default_params.restype = None
    L.cles_create.argtypes = [C.POINTER(CParams), dp, C.POINTER(C.c_int)]
    L.cles_create.restype = vp
    L.cles_destroy.argtypes = [vp]
    L.cles_destroy.restype = None
    L.cles_set_nvert.argtypes = [vp, dp]
    L.cles_set_state.argtypes = [vp, dp, C.c_int]
    L.cles_get_state.argtypes = [vp, dp, dp, C.c_int]
    L.cles_get_metrics.argtypes = [vp, dp, dp, dp, dp]
    L.cles_get_les.argtypes = [vp, dp, dp, dp]
    L.cles_eq_of_state.argtypes = [vp]
    L.cles_form_rhs.argtypes = [vp, C.c_int, dp]
    L.cles_runge_kutta.argtypes = [vp]
    L.cles_les_update.argtypes = [vp]
    L.cles_flow_step.argtypes = [vp]
    L.cles_advance.argtypes = [vp, C.c_int]
    L.cles_norms.argtypes = [vp, dp]
    L.cles_get_time.argtypes = [vp]
    L.cles_set_time.argtypes = [vp, C.c_int]
    L.cles_time_stage_kernel.argtypes = [vp, C.c_int, dp]
    L.cles_time_les.argtypes = [vp, C.c_int, dp]
    L.cles_device_bytes.argtypes = [vp]
    L.cles_device_bytes.restype = C.c_size_t
    L.cles_last_error.argtypes = [vp]
    L.cles_last_error.restype = C.c_char_p
    ip = C.POINTER(C.c_int)
    L.cles_get_nvert.argtypes = [vp, dp]
    L.cles_integrals.argtypes = [vp, dp]
    L.cles_footprint.argtypes = [C.POINTER(CParams)]
    L.cles_footprint.restype = C.c_size_t
    L.cles_batch_create.argtypes = [C.c_size_t, C.POINTER(C.c_int)]
    L.cles_batch_create.restype = vp
    L.cles_batch_destroy.argtypes = [vp]
    L.cles_batch_destroy.restype = None
    L.cles_batch_add.argtypes = [vp, vp]
    L.cles_batch_remove.argtypes = [vp, vp]
    L.cles_batch_size.argtypes = [vp]
    L.cles_batch_used.argtypes = [vp]
    L.cles_batch_used.restype = C.c_size_t
    L.cles_batch_set_tile.argtypes = [vp, C.c_int]
    L.cles_batch_set_mode.argtypes = [vp, C.c_int]
    L.cles_batch_set_spin_limit.argtypes = [vp, C.c_uint]
    L.cles_batch_set_graph.argtypes = [vp, C.c_int]
    L.cles_batch_advance.argtypes = [vp, C.c_int]
    L.cles_batch_launch.argtypes = [vp, C.c_int]
    L.cles_batch_wait.argtypes = [vp]
    L.cles_batch_time_step.argtypes = [vp, C.c_int, dp]
    L.cles_batch_last_error.argtypes = [vp]
    L.cles_batch_last_error.restype = C.c_char_p
    L.cles_set_tile.argtypes = [vp, C.c_int]
    L.cles_ibm_search.argtypes = [vp, C.c_int, dp, dp, dp, C.c_int, ip, ip, ip, dp, ip]
    L.cles_ibm_set_velocity.argtypes = [vp, dp, dp]
    L.cles_ibm_get_list.argtypes = [vp, C.POINTER(CIbmInfo), C.c_int]
    L.cles_set_state_qp.argtypes = [vp, dp, dp, C.c_int]
    L.cles_get_averages.argtypes = [vp, dp, dp, dp, dp]
    L.cles_ibm_forces.argtypes = [vp, dp, dp]
    L.cles_ibm_forces_noalloc.argtypes = [vp, dp, dp]
    L.cles_write_fields.argtypes = [vp, C.c_char_p, C.c_int]
    L.cles_write_vtk.argtypes = [vp, C.c_char_p, C.c_int]
    L.cles_write_vtk_opts.argtypes = [vp, C.c_char_p, C.c_int, C.POINTER(CVtkOpts), dp, C.c_char_p, C.c_size_t,
                                      C.POINTER(C.c_size_t)]
    L.cles_read_fields.argtypes = [vp, C.c_char_p, C.c_int, C.c_int]
    L.cles_q_criteria.argtypes = [vp, dp, dp]
    L.cles_write_q_vtk.argtypes = [vp, C.c_char_p, C.c_int, C.c_int]
    L.cles_petsc_vec_write.argtypes = [C.c_char_p, dp, C.c_longlong, C.c_int]
    L.cles_petsc_vec_read.argtypes = [C.c_char_p, dp, C.c_longlong]
    _LIB = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(cfg):
    L = load_library()
    p = CParams()
    L.cles_default_params(C.byref(p), int(cfg["mx"]), int(cfg["my"]), int(cfg["mz"]))
    for n, b in enumerate(cfg.get("bctype", (7,) * 6)):
        p.bctype[n] = int(b)
    for name, ctype in CParams._fields_:
        if name in ("mx", "my", "mz", "bctype") or name not in cfg:
            continue
        setattr(p, name, float(cfg[name]) if ctype is C.c_double else int(cfg[name]))
    return p


class CLESError(RuntimeError):
    pass


class Solver:
    """One grid block resident on the current HIP device."""

    def __init__(self, cfg, coor, nvert=None, precision=None):
        self.L = load_library()
        cfg = dict(cfg)
        if precision is not None:
            cfg["precision"] = precision
        self.cfg = cfg
        self.shape = (cfg["mz"], cfg["my"], cfg["mx"])
        self.p = make_params(cfg)
        coor = np.ascontiguousarray(coor, dtype=np.float64)
        if coor.shape != self.shape + (3,):
            raise ValueError("coordinates must have shape %s" % (self.shape + (3,),))
        err = C.c_int(0)
        self.h = self.L.cles_create(C.byref(self.p), _ptr(coor), C.byref(err))
        if not self.h:
            raise CLESError("cles_create failed (%s): %s" % (ERRORS.get(err.value, err.value),
                                                             self.L.cles_last_error(None).decode()))
        if nvert is not None:
            nv = np.ascontiguousarray(nvert, dtype=np.float64)
            self._check(self.L.cles_set_nvert(self.h, _ptr(nv)))

    def _check(self, rc):
        if rc != 0:
            raise CLESError("%s: %s" % (ERRORS.get(rc, rc), self.L.cles_last_error(self.h).decode()))
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.cles_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # state
    def set_state(self, Q, layout=LAYOUT_GLOBAL):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        self._check(self.L.cles_set_state(self.h, _ptr(Q), layout))

    def set_state_qp(self, Q, P, layout=LAYOUT_GLOBAL):
        """Q and P as given (no EqOfState); Qold = Q."""
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        P = np.ascontiguousarray(P, dtype=np.float64)
        self._check(self.L.cles_set_state_qp(self.h, _ptr(Q), _ptr(P), layout))

    def write_fields(self, directory=None, block=0):
        """Q_P_Binary_Output: the reference's binary field files of the current step."""
        d = None if directory is None else os.fsencode(directory)
        self._check(self.L.cles_write_fields(self.h, d, int(block)))


    def write_vtk(self, directory=None, block=0):
        """VTKOut (data.c:385-770): Result%2.2d_%6.6d.vts of the current state
        in directory (default: the working directory); returns its path."""
        d = None if directory is None else os.fsencode(directory)
        self._check(self.L.cles_write_vtk(self.h, d, int(block)))
        name = "Result%2.2d_%6.6d.vts" % (int(block), self.ti)
        return name if directory is None else os.path.join(os.fspath(directory), name)

    def vtk_out(self, directory=None, block=0, skip=1, theta=0.0, pres=0, airfoil=0):
        """VTKOut with its options (data.c:385-855).  Without pres: writes the
        file of write_vtk with every skip-th x node (block 2 rotated by theta)
        and returns (path, tke), tke the sum the reference prints as "TKE".
        With pres: writes nothing and returns (None, text), text the surface
        pressure lines the reference prints (block 2 along x node 20, airfoil
        along x node 8)."""
        d = None if directory is None else os.fsencode(directory)
        o = CVtkOpts(int(skip), float(theta), int(pres), int(airfoil))
        tke = C.c_double(0.0)
        used = C.c_size_t(0)
        cap = 1 << 16
        while True:
            buf = C.create_string_buffer(cap)
            self._check(self.L.cles_write_vtk_opts(self.h, d, int(block), C.byref(o), C.byref(tke), buf, cap,
                                                   C.byref(used)))
            if used.value < cap:
                break
            cap = used.value + 1
        if pres:
            return None, buf.value.decode()
        name = "Result%2.2d_%6.6d.vts" % (int(block), self.ti)
        return (name if directory is None else os.path.join(os.fspath(directory), name)), tke.value

    def q_criteria(self):
        """QCriteria (data.c:4815-5053): (Q criterion, lambda2) of the current
        state, the boundary nodes holding the pressure and 0."""
        q = np.empty(self.shape)
        l2 = np.empty(self.shape)
        self._check(self.L.cles_q_criteria(self.h, _ptr(q), _ptr(l2)))
        return q, l2

    def write_q_vtk(self, directory=None, block=0, vc=0):
        """VTKOutQ (data.c:1562-1711): QCriteria%2.2d_%5.5d.vtk; returns its path."""
        d = None if directory is None else os.fsencode(directory)
        self._check(self.L.cles_write_q_vtk(self.h, d, int(block), int(vc)))
        name = "QCriteria%2.2d_%5.5d.vtk" % (int(block), self.ti)
        return name if directory is None else os.path.join(os.fspath(directory), name)

    def read_fields(self, ti, directory=None, block=0):
        """Q_P_Binary_Input + RestartApplication from the files of step ti."""
        d = None if directory is None else os.fsencode(directory)
        self._check(self.L.cles_read_fields(self.h, d, int(ti), int(block)))

    def state(self, layout=LAYOUT_GLOBAL):
        if layout == LAYOUT_GLOBAL:
            shp = self.shape
        else:
            g = [2 if self.cfg["bctype"][2 * d] == 7 else 0 for d in range(3)]
            shp = (self.shape[0] + 2 * g[2], self.shape[1] + 2 * g[1], self.shape[2] + 2 * g[0])
        Q = np.empty(shp + (5,))
        P = np.empty(shp)
        self._check(self.L.cles_get_state(self.h, _ptr(Q), _ptr(P), layout))
        return Q, P

    def metrics(self):
        csi = np.empty(self.shape + (3,))
        eta = np.empty(self.shape + (3,))
        zet = np.empty(self.shape + (3,))
        aj = np.empty(self.shape)
        self._check(self.L.cles_get_metrics(self.h, _ptr(csi), _ptr(eta), _ptr(zet), _ptr(aj)))
        return csi, eta, zet, aj

    def les_fields(self):
        cs = np.empty(self.shape)
        nut = np.empty(self.shape)
        sig = np.empty(self.shape + (3,))
        self._check(self.L.cles_get_les(self.h, _ptr(cs), _ptr(nut), _ptr(sig)))
        return cs, nut, sig

    # reference call chain
    def eq_of_state(self):
        self._check(self.L.cles_eq_of_state(self.h))

    def form_rhs(self, istage=0):
        r = np.empty(self.shape + (5,))
        self._check(self.L.cles_form_rhs(self.h, int(istage), _ptr(r)))
        return r

    def runge_kutta(self):
        self._check(self.L.cles_runge_kutta(self.h))

    def les_update(self):
        self._check(self.L.cles_les_update(self.h))

    def flow_step(self):
        self._check(self.L.cles_flow_step(self.h))

    def advance(self, n):
        self._check(self.L.cles_advance(self.h, int(n)))

    def norms(self):
        out = np.empty(2)
        self._check(self.L.cles_norms(self.h, _ptr(out)))
        return out

    @property
    def ti(self):
        return self.L.cles_get_time(self.h)

    @ti.setter
    def ti(self, v):
        self._check(self.L.cles_set_time(self.h, int(v)))

    # measurement
    def time_stage_kernel(self, nrep=10):
        ms = C.c_double(0)
        self._check(self.L.cles_time_stage_kernel(self.h, int(nrep), C.byref(ms)))
        return ms.value

    def time_les(self, nrep=3):
        ms = C.c_double(0)
        self._check(self.L.cles_time_les(self.h, int(nrep), C.byref(ms)))
        return ms.value

    def device_bytes(self):
        return int(self.L.cles_device_bytes(self.h))

    def integrals(self):
        """{mass, momentum x/y/z, energy, kinetic energy} summed as Q/aj over
        the unique interior nodes."""
        out = np.empty(6)
        self._check(self.L.cles_integrals(self.h, _ptr(out)))
        return out

    # ------------------------------------------------------------ immersed body
    def nvert(self):
        out = np.empty(self.shape)
        self._check(self.L.cles_get_nvert(self.h, _ptr(out)))
        return out

    def ibm_search(self, verts, tris, nvert_o=None):
        """Classify the grid against a closed triangulated surface (verts
        (n_v, 3), tris (n_elmt, 3) 0-based) and build the interpolation
        stencils; returns the number of interpolation nodes."""
        v = np.ascontiguousarray(np.asarray(verts, dtype=np.float64).T)
        t = np.ascontiguousarray(np.asarray(tris, dtype=np.int32).T)
        n = C.c_int(0)
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        no = None if nvert_o is None else np.ascontiguousarray(nvert_o, dtype=np.float64)
        self._check(self.L.cles_ibm_search(self.h, v.shape[1], _ptr(v[0]), _ptr(v[1]), _ptr(v[2]), t.shape[1],
                                           ip(t[0]), ip(t[1]), ip(t[2]), _ptr(no), C.byref(n)))
        self._n_ib = n.value
        return n.value

    def ibm_set_velocity(self, u, uold):
        u = np.ascontiguousarray(u, dtype=np.float64)
        uold = np.ascontiguousarray(uold, dtype=np.float64)
        self._check(self.L.cles_ibm_set_velocity(self.h, _ptr(u), _ptr(uold)))

    def ibm_list(self):
        n = getattr(self, "_n_ib", 0)
        out = np.zeros(n, dtype=IBM_INFO_DTYPE)
        arr = (CIbmInfo * max(n, 1))()
        self._check(self.L.cles_ibm_get_list(self.h, arr, n))
        if n:
            out[:] = np.frombuffer(arr, dtype=IBM_INFO_DTYPE, count=n)
        return out

    def ibm_forces(self, center=(0.0, 0.0, 0.0)):
        """Calc_forces_SI2: F(3), Cp(3), A(3), A_tot, M(3), Mdpdn(3), Ia(3)."""
        c = np.ascontiguousarray(center, dtype=np.float64)
        out = np.empty(19)
        self._check(self.L.cles_ibm_forces(self.h, _ptr(c), _ptr(out)))
        return out

    def ibm_forces_noalloc(self, center=(0.0, 0.0, 0.0)):
        """ibm_forces without device allocation (static per-GPU scratch)."""
        c = np.ascontiguousarray(center, dtype=np.float64)
        out = np.empty(19)
        self._check(self.L.cles_ibm_forces_noalloc(self.h, _ptr(c), _ptr(out)))
        return out

    def averages(self):
        """Do_averaging sums: Q (5), P, u^2 (3), (uv, vw, wu) per node."""
        q = np.empty(self.shape + (5,))
        p = np.empty(self.shape)
        u2 = np.empty(self.shape + (3,))
        uc = np.empty(self.shape + (3,))
        self._check(self.L.cles_get_averages(self.h, _ptr(q), _ptr(p), _ptr(u2), _ptr(uc)))
        return q, p, u2, uc

    def set_tile(self, tile):
        """Stage-kernel tile shape: 0 32x4x2, 1 16x4x4, 2 16x8x2, 3 64x2x2, 4 6x6x6, 5 8x5x5
        (256 threads), 6 4x4x3 (64 threads), 7 8x4x3 (128 threads)."""
        self._check(self.L.cles_set_tile(self.h, int(tile)))


def petsc_vec_write(path, x, block_size=1):
    """Write x (flattened, float64) as one PETSc binary Vec file."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    L = load_library()
    if L.cles_petsc_vec_write(os.fsencode(path), _ptr(x), x.size, int(block_size)) != 0:
        raise CLESError(L.cles_last_error(None).decode())


def petsc_vec_read(path, n):
    """Read a PETSc binary Vec file of exactly n values."""
    x = np.empty(int(n))
    L = load_library()
    if L.cles_petsc_vec_read(os.fsencode(path), _ptr(x), x.size) != 0:
        raise CLESError(L.cles_last_error(None).decode())
    return x


def footprint(cfg, precision=None):
    """Device bytes a Solver for cfg allocates (the batch admission size)."""
    p = make_params(cfg)
    if precision is not None:
        p.precision = precision
    return load_library().cles_footprint(C.byref(p))


class Batch:
    """Independent periodic solves advanced together by the batched kernels.
    Members are Solver objects with their state uploaded."""

    def __init__(self, hbm_budget):
        self.L = load_library()
        err = C.c_int(0)
        self.h = self.L.cles_batch_create(C.c_size_t(int(hbm_budget)), C.byref(err))
        if not self.h:
            raise CLESError("cles_batch_create failed (%d): %s" % (err.value, self.L.cles_last_error(None).decode()))
        self.members = []

    def _check(self, rc):
        if rc != 0:
            raise CLESError("%s (%d)" % (self.L.cles_batch_last_error(self.h).decode(), rc))

    def add(self, solver):
        """Admit a member; returns False when it does not fit the HBM budget."""
        rc = self.L.cles_batch_add(self.h, solver.h)
        if rc == ENOMEM:
            return False
        self._check(rc)
        self.members.append(solver)
        return True

    def remove(self, solver):
        self._check(self.L.cles_batch_remove(self.h, solver.h))
        self.members.remove(solver)

    def used(self):
        return self.L.cles_batch_used(self.h)

    def size(self):
        return self.L.cles_batch_size(self.h)

    def set_tile(self, tile):
        self._check(self.L.cles_batch_set_tile(self.h, int(tile)))

    def set_mode(self, mode):
        """BATCH_LAUNCHES (default) or BATCH_PERSISTENT"""
        self._check(self.L.cles_batch_set_mode(self.h, int(mode)))

    def set_graph(self, on):
        """replay a captured step (True) or launch each step's kernels (False)"""
        self._check(self.L.cles_batch_set_graph(self.h, int(bool(on))))

    def set_spin_limit(self, polls):
        """polls of a stage dependency wait of the persistent kernel before it gives up"""
        self._check(self.L.cles_batch_set_spin_limit(self.h, int(polls)))

    def advance(self, nsteps):
        self._check(self.L.cles_batch_advance(self.h, int(nsteps)))

    def launch(self, nsteps):
        """queue nsteps batched steps and return (see wait)"""
        self._check(self.L.cles_batch_launch(self.h, int(nsteps)))

    def wait(self):
        self._check(self.L.cles_batch_wait(self.h))

    def time_step(self, nrep=5):
        ms = C.c_double(0)
        self._check(self.L.cles_batch_time_step(self.h, int(nrep), C.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            self.L.cles_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
