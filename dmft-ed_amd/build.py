#!/usr/bin/env python3
"""Build the gfx950 library libdmed.so (HIP + C++) and the Fortran layer.

    python dmft-ed_amd/build.py            # library + Fortran module/driver
    python dmft-ed_amd/build.py --lib      # library only

Outputs stay in-tree (dmft-ed_amd/libdmed.so, dmft-ed_amd/fortran/build/) so
they travel with the source tree to the GPU host; no build system is needed
beyond hipcc and amdflang.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdmed.so")
FDIR = os.path.join(HERE, "fortran")
FBUILD = os.path.join(FDIR, "build")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
FLANG = shutil.which("amdflang") or os.path.join(ROCM, "llvm", "bin", "amdflang")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-variable", "-Wno-unused-value",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-value", f"-I{ROCM}/include",
             "-D__HIP_PLATFORM_AMD__"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_lib(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))
    cmds = []
    objs = []
    for f in srcs:
        src = os.path.join(CSRC, f)
        obj = os.path.join(OBJ, f + ".o")
        objs.append(obj)
        if f.endswith(".hip"):
            cmds.append([HIPCC] + HIP_FLAGS + ["-x", "hip", "-c", src, "-o", obj])
        else:
            cmds.append([HIPCC] + CXX_FLAGS + ["-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    """DMFT_ED module (ISO_C_BINDING over libdmed.so) and the Bethe drivers
    (single site and inequivalent sites)."""
    os.makedirs(FBUILD, exist_ok=True)
    srcs = ["dmed_cabi.f90", "dmft_ed.f90", "ed_hxv_gpu.f90"]
    objs = []
    for f in srcs:
        src = os.path.join(FDIR, f)
        if not os.path.exists(src):
            return None
        obj = os.path.join(FBUILD, f.replace(".f90", ".o"))
        _run([FLANG, "-O2", "-fPIC", "-J", FBUILD, "-c", src, "-o", obj])
        objs.append(obj)
    exe = None
    for name in ("ed_hm_bethe", "ed_hm_bethe_ineq"):
        drv = os.path.join(FDIR, name + ".f90")
        if os.path.exists(drv):
            out = os.path.join(FBUILD, name)
            _run([FLANG, "-O2", "-J", FBUILD, drv] + objs +
                 ["-o", out, f"-L{HERE}", "-ldmed", f"-Wl,-rpath,{HERE}",
                  "-Wl,-rpath,$ORIGIN/../.."])
            exe = exe or out
    return exe


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    print(build_lib(a.j))
    if not a.lib:
        print(build_fortran())


if __name__ == "__main__":
    sys.exit(main())
