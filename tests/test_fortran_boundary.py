"""The Fortran side of the boundary: ow3d_iface.f90 (ISO_C_BINDING interface to include/ow3d.h)
compiles on its own, and ow3d_dropin.f90 (the reference-named replacements BuildLinearSystem,
MultiGridSolveLinearSystem, iterative_solution, VerticalFreeSurfaceVelocity) compiles against the
reference's own modules (DataTypes, GlobalVariables, MGLevels) and links with libow3d.so."""
import os
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FC = "/opt/rocm/lib/llvm/bin/flang"
FDIR = os.path.join(ROOT, "oceanwave3d-fortran90_amd", "fortran")
REFOBJ = os.path.join(ROOT, "oracle", "_ref", "obj")

pytestmark = pytest.mark.skipif(not os.path.exists(FC), reason="flang not installed")


def test_interface_module_compiles():
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([FC, "-c", os.path.join(FDIR, "ow3d_iface.f90"), "-module-dir", d, "-o", os.path.join(d, "i.o")],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


@pytest.mark.skipif(not os.path.isdir(REFOBJ), reason="reference modules (oracle/_ref) not built")
def test_dropin_wrappers_compile_against_reference_modules_and_link():
    with tempfile.TemporaryDirectory() as d:
        for src, obj in (("ow3d_iface.f90", "i.o"), ("ow3d_dropin.f90", "d.o")):
            r = subprocess.run([FC, "-c", os.path.join(FDIR, src), "-module-dir", d, "-I", d, "-I", REFOBJ,
                                "-o", os.path.join(d, obj)], capture_output=True, text=True)
            assert r.returncode == 0, r.stderr
        # the wrappers' external references are the C ABI (resolved by libow3d.so) and the reference's
        # module data (resolved by the reference objects it is linked into)
        nm = subprocess.run(["nm", "-u", os.path.join(d, "d.o")], capture_output=True, text=True).stdout
        undefined = [l.split()[-1] for l in nm.splitlines() if l.strip()]
        ow3d_syms = [u for u in undefined if u.startswith("ow3d_")]
        assert set(ow3d_syms) >= {"ow3d_create", "ow3d_laplace_apply", "ow3d_precond_apply", "ow3d_solve",
                                  "ow3d_vertical_velocity"}
        lib = os.path.join(ROOT, "oceanwave3d-fortran90_amd", "ow3d", "libow3d.so")
        exported = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True).stdout
        for s in ow3d_syms:
            assert f" {s}" in exported, s
        defined = subprocess.run(["nm", "--defined-only", os.path.join(d, "d.o")], capture_output=True, text=True).stdout
        for name in ("buildlinearsystem_", "multigridsolvelinearsystem_", "iterative_solution_", "verticalfreesurfacevelocity_"):
            assert name in defined, name
