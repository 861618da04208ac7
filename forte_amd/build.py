"""Build the gfx950 FCI library and its Python module in-tree.

    python forte_amd/build.py        (run by path: importing the package needs the extension)

1. every csrc/*.hip is compiled by hipcc for gfx950 and linked into
   forte_amd/_lib/libforte_fci_gfx950.so (the C ABI of include/forte_fci.h);
2. csrc/forte_amd_py.cpp is compiled by g++ against pybind11 into forte_amd/_forte_amd*.so,
   linked to the library with an $ORIGIN rpath.
"""
import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
OBJDIR = os.path.join(HERE, "_build")
LIB = os.path.join(LIBDIR, "libforte_fci_gfx950.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # stream 0 is each host thread's own stream: solvers driven from different threads overlap
    "-fgpu-default-stream=per-thread",
    "-Wno-unused-result",
]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l)
    if p.returncode != 0:
        raise RuntimeError(f"command failed ({p.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_library(jobs=8):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    objs = [os.path.join(OBJDIR, f[:-4] + ".o") for f in srcs]
    cmds = [[HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o] for s, o in zip(srcs, objs)]
    with ThreadPoolExecutor(max_workers=min(jobs, len(cmds))) as ex:
        list(ex.map(_run, cmds))
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", LIB + ".tmp", *objs])
    os.replace(LIB + ".tmp", LIB)
    return LIB


def build_module():
    import pybind11

    ext = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(HERE, "_forte_amd" + ext)
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC",
           "-I" + sysconfig.get_paths()["include"], "-I" + pybind11.get_include(),
           os.path.join(CSRC, "forte_amd_py.cpp"), "-o", out + ".tmp",
           "-L" + LIBDIR, "-lforte_fci_gfx950", "-Wl,-rpath,$ORIGIN/_lib"]
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build():
    lib = build_library()
    mod = build_module()
    return lib, mod


if __name__ == "__main__":
    for p in build():
        print(p)
    sys.exit(0)
