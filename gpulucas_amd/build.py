"""Build libgpulucas.so (HIP, gfx950) in-tree with hipcc.

    python -m gpulucas_amd.build          # library
    python -m gpulucas_amd.build --oracle # also the CPU oracle (make -C oracle)
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libgpulucas.so")
# the same library with the kernels' phase stamps compiled in (profiling only)
PHASES_LIB = os.path.join(PKG, "libgpulucas_phases.so")
CLI = os.path.join(PKG, "gpulucas")

SOURCES = ["lucas_kernels.hip", "lucas_host.cpp", "lucas_engine.cpp", "lucas_capi.cpp", "lucas_noalloc.hip",
           "lucas_line.hip"]
# kernel files built a second time, with phase stamps, for the profiling library
PHASED = ["lucas_kernels.hip", "lucas_line.hip"]
HEADERS = ["lucas_desc.h", "lucas_host.h", "lucas_engine.h", "lucas_device.h", "lucas_line.h"]

ARCH = os.environ.get("GPULUCAS_ARCH", "gfx950")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc"))

# Packed FP32 instructions are disabled for all device code on these hosts.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + INCLUDE,
    "-I" + CSRC,
    "-Wall",
]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)


def build_library(verbose: bool = False, jobs: int = 4) -> str:
    """Compile every source (in parallel) and link libgpulucas.so, and
    libgpulucas_phases.so from the same host objects and the kernels built
    with LUCAS_PHASE_STAMPS."""
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, procs = [], []
    # the kernel files once more with the phase stamps compiled in
    phases_objs = []
    for src in PHASED:
        obj = os.path.join(objdir, src.rsplit(".", 1)[0] + "_phases.o")
        phases_objs.append(obj)
        cmd = [HIPCC] + HIP_FLAGS + ["-DLUCAS_PHASE_STAMPS", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    for src in SOURCES:
        obj = os.path.join(objdir, src.rsplit(".", 1)[0] + ".o")
        objs.append(obj)
        cmd = [HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
        if src.endswith(".cpp"):
            # host-only translation units: no device pass
            cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-Wall", "-I" + INCLUDE, "-I" + CSRC,
                   "-I" + os.path.join(ROCM, "include"), "-D__HIP_PLATFORM_AMD__", "-x", "c++",
                   "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for cmd, p in procs:
        out, _ = p.communicate()
        out = "\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l)
        if p.returncode != 0:
            failed.append(f"{' '.join(cmd)}\n{out}")
        elif verbose and out.strip():
            print(out)
    if failed:
        raise RuntimeError("compile failed:\n" + "\n\n".join(failed))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"], verbose)
    os.replace(tmp, LIB)
    tmp = PHASES_LIB + ".tmp"
    plain = [o for o, src in zip(objs, SOURCES) if src not in PHASED]
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + phases_objs + plain + ["-lpthread"],
         verbose)
    os.replace(tmp, PHASES_LIB)
    build_cli(verbose)
    return LIB


def build_cli(verbose: bool = False) -> str:
    """The command-line driver `gpulucas_amd/gpulucas`, linked to the library."""
    tmp = CLI + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + INCLUDE, os.path.join(CSRC, "lucas_cli.cpp"),
          "-o", tmp, "-L" + PKG, "-lgpulucas", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, CLI)
    return CLI


def build_oracle(verbose: bool = False) -> None:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], verbose)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    build_library(a.verbose)
    if a.oracle:
        build_oracle(a.verbose)
    print(LIB)


if __name__ == "__main__":
    sys.exit(main())
