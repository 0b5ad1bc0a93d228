"""Builds the Fortran test program tests/fortran/boundary_check.f90 against the interface module
(oceanwave3d-fortran90_amd/fortran/ow3d_iface.f90) and libow3d.so, into tests/fortran/_build/."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
FC = "/opt/rocm/lib/llvm/bin/flang"
FDIR = os.path.join(ROOT, "oceanwave3d-fortran90_amd", "fortran")
LIBDIR = os.path.join(ROOT, "oceanwave3d-fortran90_amd", "ow3d")
PROG = os.path.join(HERE, "_build", "boundary_check")


def build_program():
    out = os.path.dirname(PROG)
    os.makedirs(out, exist_ok=True)
    subprocess.run([FC, "-c", os.path.join(FDIR, "ow3d_iface.f90"), "-module-dir", out, "-o", os.path.join(out, "iface.o")],
                   check=True)
    subprocess.run([FC, "-I", out, os.path.join(HERE, "boundary_check.f90"), os.path.join(out, "iface.o"),
                    os.path.join(LIBDIR, "libow3d.so"), f"-Wl,-rpath,{LIBDIR}", "-o", PROG], check=True)
    return PROG


if __name__ == "__main__":
    print(build_program())
