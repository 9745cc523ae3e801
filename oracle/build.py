"""ORACLE build recipe (test infrastructure).

liboracle.so: the plain-C restatement in this folder (gcc, OpenMP).
_ref/libref_sym.so: the reference's own symmetry/Wigner/sort sources compiled
unchanged from /root/reference (only when that tree is present).  These files
need nothing beyond libc/libm and -DMAX_SYMMETRIES=5 (CMakeLists.txt:18,76);
the rest of the reference needs CBLAS/LAPACKE/HDF5 headers the image lacks and
is not built.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REFLIB = os.path.join(HERE, "_ref", "libref_sym.so")
REF_SOURCES = ["symmetries.c", "symmetry_z2.c", "symmetry_u1.c", "symmetry_su2.c",
               "symmetry_pg.c", "Wigner.c", "macros.c", "sort.c"]
CFLAGS = ["-O3", "-fPIC", "-shared", "-std=gnu11", "-fopenmp", "-march=x86-64-v3"]


def build_oracle():
    srcs = sorted(os.path.join(HERE, f) for f in os.listdir(HERE) if f.endswith(".c"))
    cmd = ["gcc"] + CFLAGS + srcs + ["-o", LIB + ".tmp", "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_ref(refroot="/root/reference"):
    src = os.path.join(refroot, "src")
    if not os.path.isdir(src):
        return None
    os.makedirs(os.path.dirname(REFLIB), exist_ok=True)
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=gnu11", "-DMAX_SYMMETRIES=5",
           "-I", os.path.join(refroot, "include")] + \
          [os.path.join(src, f) for f in REF_SOURCES] + ["-o", REFLIB + ".tmp", "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(REFLIB + ".tmp", REFLIB)
    return REFLIB


if __name__ == "__main__":
    print(build_oracle())
    print(build_ref(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
