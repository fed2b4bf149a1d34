"""Build the CPU oracle (test infrastructure only).

    python oracle/build.py     # -> oracle/liboracle.so

g++ -O2 with -ffp-contract=off and no fast-math: one IEEE binary32 operation
per source operation, as the restatement requires.  OpenMP parallelises over
triangles; results do not depend on the thread count.
"""
import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def build() -> Path:
    out = HERE / "liboracle.so"
    tmp = HERE / "liboracle.so.tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
           "-fPIC", "-shared", "-Wall", "-Wextra", str(HERE / "crx_oracle.cpp"), "-o", str(tmp)]
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("oracle build failed")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print("built", build())
