"""Build the gfx950 engine in-tree: fermiphasespace_amd/libfps_ftwa.so and the fps_mi355x CLI.

    python -m fermiphasespace_amd.build        (or fermiphasespace_amd.build.build())

Device code is compiled for gfx950 only.  Every HIP translation unit gets
`-Xclang -target-feature -Xclang -packed-fp32-ops` (no packed FP32 instructions).
The two kernel precisions compile as separate translation units, in parallel.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OUT = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfps_ftwa.so")
CLI = os.path.join(PKG, "fps_mi355x")
MAKE_CLI = os.path.join(PKG, "fps_make_system")  # the system generator (makeSitesSystem)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass warns that it does not know the packed-fp32 feature: not an error
    noise = "packed-fp32-ops"
    msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if noise not in l).strip()
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{msg}")
    if msg and os.environ.get("FPS_BUILD_VERBOSE"):
        print(msg, file=sys.stderr)


def build(verbose: bool = False) -> str:
    """Compile every translation unit and link the shared library; returns its path."""
    os.makedirs(OUT, exist_ok=True)
    hip_units = [
        ("ftwa_kernels.hip", "ftwa_kernels_f32.o", ["-DFTWA_REAL=1"]),
        ("ftwa_kernels.hip", "ftwa_kernels_f64.o", ["-DFTWA_REAL=2"]),
        ("ftwa_stencil.hip", "ftwa_stencil_f32.o", ["-DFTWA_REAL=1"]),
        ("ftwa_stencil.hip", "ftwa_stencil_f64.o", ["-DFTWA_REAL=2"]),
        ("ftwa_herm.hip", "ftwa_herm.o", []),
        ("ftwa_herm_seq.hip", "ftwa_herm_seq.o", []),
        ("ftwa_herm_wide.hip", "ftwa_herm_wide.o", []),
        ("ftwa_gpsr.hip", "ftwa_gpsr.o", []),
        ("ftwa_peaks.hip", "ftwa_peaks.o", []),
        ("ftwa_adjoint.hip", "ftwa_adjoint.o", []),
        ("ftwa_dense.hip", "ftwa_dense.o", []),
        ("ftwa_api.cpp", "ftwa_api.o", ["-x", "hip"]),
    ]
    cxx_units = [("json.cpp", "json.o"), ("system.cpp", "system.o"), ("eigen.cpp", "eigen.o"),
                 ("solver.cpp", "solver.o"), ("lattice_gen.cpp", "lattice_gen.o")]
    jobs = []
    for src, obj, extra in hip_units:
        jobs.append([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", os.path.join(OUT, obj)])
    cxx = shutil.which("g++") or "g++"
    for src, obj in cxx_units:
        jobs.append([cxx, *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", os.path.join(OUT, obj)])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for f in [ex.submit(_run, j) for j in jobs]:
            f.result()
    objs = [os.path.join(OUT, o) for _, o, _ in hip_units] + [os.path.join(OUT, o) for _, o in cxx_units]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp, *objs, "-L/opt/rocm/lib", "-lrocblas",
          "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    _run([cxx, *CXX_FLAGS, os.path.join(CSRC, "fps_main.cpp"), "-o", CLI + ".tmp", "-L" + PKG, "-lfps_ftwa",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(CLI + ".tmp", CLI)
    _run([cxx, *CXX_FLAGS, os.path.join(CSRC, "make_system_main.cpp"), "-o", MAKE_CLI + ".tmp", "-L" + PKG,
          "-lfps_ftwa", "-Wl,-rpath,$ORIGIN"])
    os.replace(MAKE_CLI + ".tmp", MAKE_CLI)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
