"""Builds each MEX gateway of tom_amd/mex against the test runtime
(tests/mex_stub/mexrt.c) and libtomhip.so, one shared library per gateway in
tests/mex_stub/_build/, so the tests can call mexFunction end to end.

    python tests/mex_stub/build_gateways.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")


def build() -> list:
    os.makedirs(OUT, exist_ok=True)
    mex = os.path.join(ROOT, "tom_amd", "mex")
    libdir = os.path.join(ROOT, "tom_amd")
    built = []
    for f in sorted(os.listdir(mex)):
        if not f.endswith(".c"):
            continue
        name = os.path.splitext(f)[0]
        so = os.path.join(OUT, f"libgw_{name}.so")
        cmd = ["gcc", "-std=c99", "-O1", "-Wall", "-Werror", "-shared", "-fPIC", "-I", HERE, "-I",
               os.path.join(ROOT, "include"), os.path.join(mex, f), os.path.join(HERE, "mexrt.c"), "-L", libdir,
               "-ltomhip", "-Wl,-rpath,$ORIGIN/../../../tom_amd", "-lm", "-o", so + ".tmp"]
        subprocess.check_call(cmd)
        os.replace(so + ".tmp", so)
        built.append(so)
    return built


if __name__ == "__main__":
    for p in build():
        print(p)
    sys.exit(0)
