"""Build the gfx950 rasterizer library in-tree.

    python curast-ooc_amd/build.py            # libcrx.so + libcrx_gen.so

Outputs go to curast-ooc_amd/lib/.  Kernels are compiled ahead of time for
gfx950 with IEEE float semantics: no contraction into FMAs, correctly rounded
division and square root, denormals kept, packed-FP32 instructions disabled.
"""
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
INCLUDE = PKG.parent / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CRX_OFFLOAD_ARCH", "gfx950")

FLOAT_FLAGS = [
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
]
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
] + FLOAT_FLAGS


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(
        line for line in proc.stdout.splitlines()
        if "packed-fp32-ops" not in line and line.strip() != "1 warning generated."
    )
    if proc.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("command failed: " + " ".join(cmd))
    if out.strip():
        sys.stderr.write(out + "\n")


def _atomic_link(cmd, target: Path):
    tmp = target.with_suffix(target.suffix + ".tmp")
    _run(cmd + ["-o", str(tmp)])
    os.replace(tmp, target)


def build(verbose: bool = False) -> None:
    LIB.mkdir(exist_ok=True)
    obj_dir = LIB / "obj"
    obj_dir.mkdir(exist_ok=True)
    objs = []
    for src in ("crx_kernels.hip", "crx_api.cpp", "crx_args.cpp", "crx_server.cpp", "crx_frame.cpp", "crx_cull.cpp"):
        obj = obj_dir / (Path(src).stem + ".o")
        cmd = [HIPCC, "-x", "hip"] + HIP_FLAGS + ["-pthread", "-c", str(CSRC / src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(str(obj))
    _atomic_link([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread"] + objs, LIB / "libcrx.so")
    gen_cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-Wall",
               f"-I{INCLUDE}", str(CSRC / "crx_gen.cpp")]
    _atomic_link(gen_cmd, LIB / "libcrx_gen.so")
    shutil.rmtree(obj_dir, ignore_errors=True)


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", LIB / "libcrx.so", LIB / "libcrx_gen.so")
