"""Build the engine in-tree: libpaising.so (kernels + host engine + C ABI) and the drop-in
executables PAisingSSC / PAisingMSC, all for gfx950 with hipcc."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libpaising.so")
BIN_DIR = os.path.join(HERE, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PA_OFFLOAD_ARCH", "gfx950")
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose=True):
    os.makedirs(BIN_DIR, exist_ok=True)
    objs = []
    for src in ("pa_kernels.hip", "pa_engine.cpp"):
        obj = os.path.join(CSRC, src.rsplit(".", 1)[0] + ".o")
        _run([HIPCC, *COMMON, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, *COMMON, "-shared", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    cli = os.path.join(CSRC, "pa_cli.cpp")
    for name, coding in (("PAisingSSC", 0), ("PAisingMSC", 1)):
        _run([HIPCC, *COMMON, "-x", "hip", f"-DPA_CLI_CODING={coding}", cli, "-o",
              os.path.join(BIN_DIR, name), f"-L{HERE}", "-lpaising", f"-Wl,-rpath,{HERE}"])
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
