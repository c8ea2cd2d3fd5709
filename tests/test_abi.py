"""The C ABI library loads and exports every entry point include/cles.h declares
(no device calls: this runs on the CPU)."""
import ctypes
import os
import re

import chimera_comp_les_amd as cl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "cles.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cles_[a-z_]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert declared_functions() == sorted(cl.EXPORTS)


def test_library_exports_every_symbol():
    L = cl.load_library()
    for name in declared_functions():
        assert hasattr(L, name), name
        assert ctypes.cast(getattr(L, name), ctypes.c_void_p).value


def test_params_struct_layout():
    # field order/size of cles_params as the header declares it
    src = open(os.path.join(ROOT, "include", "cles.h")).read()
    body = src[src.index("typedef struct cles_params {"):src.index("} cles_params;")]
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    names = []
    for decl in body.split(";")[:-1]:
        decl = decl.split("{")[-1].strip()
        typ, rest = decl.split(None, 1)
        for n in rest.split(","):
            names.append(n.strip().split("[")[0])
    assert names == [f[0] for f in cl.CParams._fields_]


def test_default_params_match_reference_defaults():
    p = cl.make_params(dict(mx=8, my=9, mz=10))
    assert (p.mx, p.my, p.mz) == (8, 9, 10)
    assert p.ren == 3000.0 and p.Ma == 2.81 and p.Pr == 0.72 and p.gamma == 1.4   # init.c:85-96
    assert p.dt == 0.01 and p.fluxorder == 3 and p.eno_order == 2 and p.central == 2
    assert p.max_cs == 0.05 and p.Tref == 273.15 and p.sigma_thrsh == 1e-5


def _create_err(**kw):
    L = cl.load_library()
    base = dict(mx=8, my=9, mz=10, bctype=(7, 7, 7, 7, 7, 7))
    base.update(kw)
    p = cl.make_params(base)
    err = ctypes.c_int(0)
    coor = (ctypes.c_double * 3)()
    ctx = L.cles_create(ctypes.byref(p), coor, ctypes.byref(err))
    assert not ctx
    return err.value, L.cles_last_error(None).decode()


def test_create_rejects_unsupported_options_before_device_use():
    # parameter validation runs on the host before any allocation
    e, msg = _create_err(bctype=(7, 7, 1, 1, 49, 8))   # 49 is a k-max type only (rhs.c:973)
    assert e == -1 and "bctype[4]=49" in msg
    e, msg = _create_err(bctype=(7, 3, 7, 7, 7, 7))
    assert e == -1 and "periodic" in msg
    e, msg = _create_err(fluxorder=4)
    assert e == -1 and "fluxorder" in msg
    e, msg = _create_err(les=2, invicid=1)
    assert e == -1 and "LES" in msg
    e, msg = _create_err(twoD=4)
    assert e == -1 and "twoD" in msg
    e, msg = _create_err(weno5=1, precision=1)
    assert e == -1 and "FP64" in msg


def test_create_rejects_grids_beyond_32_bit_plane_offsets():
    # a plane of more than 2^29 nodes would not fit the stage kernel's 32-bit byte offsets
    e, msg = _create_err(mx=1024, my=1024, mz=1024)
    assert e == -1 and "2^29" in msg
