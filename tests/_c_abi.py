"""Build recipe of tests/c_abi/solve_fixture.c against the in-tree libraries."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "qiotoolkit_amd")


def build(out_dir):
    exe = os.path.join(out_dir, "solve_fixture")
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c_abi", "solve_fixture.c"), "-L", PKG, "-lqiotk", "-lqmc",
           "-Wl,-rpath," + PKG, "-o", exe]
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, timeout=120)
    return exe
