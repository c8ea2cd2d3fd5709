"""The boundary as a C compiler sees it (CPU only).

* tests/c_abi/caller.c, a C99 program, is compiled with gcc against
  include/cles.h, linked to libcles.so and run: the reference defaults, the
  option validation of cles_create (before any device use) and the PETSc
  binary Vec writer/reader; its printed struct sizes and offsets must equal
  the ctypes mirror's.
* The reference-side binding of INTEGRATION.md (its C blocks) is compiled
  against tests/c_abi/petsc_stub.h, a minimal stand-in for the PETSc and
  reference types it touches, to prove that it builds.
"""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

import chimera_comp_les_amd as cl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.join(ROOT, "tests", "c_abi")
GCC = shutil.which("gcc")

pytestmark = pytest.mark.skipif(GCC is None, reason="gcc not found")


def test_c_caller_builds_and_runs(tmp_path):
    lib = os.path.join(ROOT, "chimera_comp_les_amd", "libcles.so")
    if not os.path.exists(lib):
        pytest.skip("libcles.so is not built")
    exe = str(tmp_path / "caller")
    cmd = [GCC, "-std=c99", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
           os.path.join(HERE, "caller.c"), "-o", exe, lib, "-Wl,-rpath," + os.path.dirname(lib)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe, str(tmp_path / "vec.dat")], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("OK 0"), out.stdout
    m = re.search(r"sizeof cles_params (\d+)", out.stdout)
    assert int(m.group(1)) == ctypes.sizeof(cl.CParams)
    m = re.search(r"offsetof ren (\d+) les (\d+) precision (\d+) averaging (\d+) k_homo_filter (\d+)", out.stdout)
    for name, off in zip(("ren", "les", "precision", "averaging", "k_homo_filter"), m.groups()):
        assert getattr(cl.CParams, name).offset == int(off), name
    m = re.search(r"sizeof cles_ibm_info (\d+) offsetof pmin (\d+) cr3 (\d+)", out.stdout)
    assert int(m.group(1)) == ctypes.sizeof(cl.CIbmInfo)
    assert cl.CIbmInfo.pmin.offset == int(m.group(2)) and cl.CIbmInfo.cr3.offset == int(m.group(3))
    m = re.search(r"sizeof cles_vtk_opts (\d+) offsetof theta (\d+) pres (\d+) airfoil (\d+)", out.stdout)
    assert int(m.group(1)) == ctypes.sizeof(cl.CVtkOpts)
    for name, off in zip(("theta", "pres", "airfoil"), m.groups()[1:]):
        assert getattr(cl.CVtkOpts, name).offset == int(off), name
    # the file the C program wrote is the PETSc binary format the Python side reads
    x = cl.petsc_vec_read(str(tmp_path / "vec.dat"), 37)
    assert all(x[n] == (n - 18) * 0.375 + 1e-300 * n for n in range(37))


def test_reference_binding_compiles(tmp_path):
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```c\n(.*?)```", text, re.S)
    assert len(blocks) >= 3
    src = tmp_path / "binding.c"
    body = "\n".join(b for b in blocks if "static PetscErrorCode" in b)
    src.write_text('#include "petsc_stub.h"\n' + body)
    r = subprocess.run([GCC, "-std=c99", "-Wall", "-Werror", "-Wno-unused-function", "-c",
                        "-I" + os.path.join(ROOT, "include"), "-I" + HERE, str(src), "-o",
                        str(tmp_path / "binding.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr + "\n" + body
