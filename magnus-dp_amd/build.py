"""Build the MI355X (gfx950) library and its Python extension in-tree.

Outputs (git-ignored, carried with the tree):
  magnus-dp_amd/magnus/libmagnus_hip.so   C ABI of include/magnus_capi.h
  magnus-dp_amd/magnus/_core*.so          pybind11 module over the C ABI

Usage: python magnus-dp_amd/build.py [-j N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import sysconfig
from pathlib import Path

PKG_ROOT = Path(__file__).resolve().parent
REPO = PKG_ROOT.parent
CSRC = PKG_ROOT / "csrc"
OUT = PKG_ROOT / "magnus"
BUILD = PKG_ROOT / "_build"
INCLUDE = REPO / "include"

ARCH = os.environ.get("MAGNUS_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# packed-FP32 instructions stay disabled in all device code
COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-parameter", "-Wno-unused-result",
    f"-I{INCLUDE}", f"-I{CSRC}",
]

DEVICE_SOURCES = [
    "kernels_lm.hip", "kernels_wb.hip", "kernels_sample.hip", "kernels_expm.hip", "kernels_uneven.hip",
    "kernels_wbd.hip", "kernels_ctrl.hip", "kernels_mlp.hip", "kernels_wbd8.hip", "noalloc.hip", "kernels_walk.hip", "kernels_walkseg.hip",
]
HOST_SOURCES = ["capi.cpp", "uneven.cpp", "mlp.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def _compile(src: Path, obj: Path, hip: bool):
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    if hip:
        cmd[1:1] = ["-x", "hip"]
    _run(cmd)
    return obj


def build(jobs: int = 8, verbose: bool = False) -> None:
    BUILD.mkdir(exist_ok=True)
    sys.path.insert(0, str(CSRC))
    import gen_gl_table  # noqa: E402

    gen_gl_table.main()
    objs = []
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = [ex.submit(_compile, CSRC / s, BUILD / (s + ".o"), True) for s in DEVICE_SOURCES]
        futs += [ex.submit(_compile, CSRC / s, BUILD / (s + ".o"), False) for s in HOST_SOURCES]
        for f in futs:
            objs.append(f.result())
    lib = OUT / "libmagnus_hip.so"
    tmp = lib.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, lib)

    import pybind11

    ext = OUT / ("_core" + sysconfig.get_config_var("EXT_SUFFIX"))
    tmp = ext.with_suffix(".tmp")
    _run([
        HIPCC, "-O2", "-fPIC", "-shared", "-std=c++17", "-fvisibility=hidden",
        f"-I{INCLUDE}", f"-I{pybind11.get_include()}", f"-I{sysconfig.get_paths()['include']}",
        str(CSRC / "pybind_core.cpp"), f"-L{OUT}", "-lmagnus_hip", "-Wl,-rpath,$ORIGIN",
        "-o", str(tmp),
    ])
    os.replace(tmp, ext)
    if verbose:
        print(f"built {lib}\nbuilt {ext}")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    args = ap.parse_args()
    build(args.j, verbose=True)
