"""Build the in-tree native libraries.

* taylor_amd/libmega_hip.so  -- the product: HIP kernels for gfx950 + C-ABI
  (include/mega.h), compiled with hipcc.
* taylor_amd/host/mega_drive -- the Fortran host (bind(C) interfaces and the
  time loop), compiled with flang against libmega_hip.so.
* oracle/liboracle.so       -- the CPU checker (test infrastructure).
* oracle/_ref/              -- the reference program compiled from its own
  sources, only where those sources exist (this build container).

Packed FP32 instructions are disabled for all device code
(-target-feature -packed-fp32-ops)."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmega_hip.so")
REF_SRC = "/root/reference"

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def hipcc() -> str:
    for cand in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_hip(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, f) for f in sorted(os.listdir(CSRC)) if f.endswith(".hip")]
    tmp = LIB + ".tmp"
    cmd = [hipcc(), *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC, *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_host(verbose: bool = False) -> str | None:
    """The Fortran host (taylor_amd/host/mega_drive): the bind(C) interface
    module, the reader of the reference's input files, the writer of its
    VTK frames and the time-loop driver, linked against libmega_hip.so."""
    fc = shutil.which("flang") or ("/opt/rocm/lib/llvm/bin/flang" if os.path.exists("/opt/rocm/lib/llvm/bin/flang") else None)
    if not fc:
        return None
    import tempfile
    host = os.path.join(HERE, "host")
    out = os.path.join(host, "mega_drive")
    with tempfile.TemporaryDirectory() as td:
        cmds = [[fc, "-O2", "-c", "-module-dir", td, os.path.join(host, "mega_iface.f90"), "-o",
                 os.path.join(td, "mega_iface.o")],
                [fc, "-O2", "-c", "-module-dir", td, "-I", td, os.path.join(host, "mega_deck.f90"), "-o",
                 os.path.join(td, "mega_deck.o")],
                [fc, "-O2", "-c", "-module-dir", td, "-I", td, os.path.join(host, "mega_vtk.f90"), "-o",
                 os.path.join(td, "mega_vtk.o")],
                [fc, "-O2", "-module-dir", td, "-I", td, os.path.join(host, "mega_drive.f90"),
                 os.path.join(td, "mega_iface.o"), os.path.join(td, "mega_deck.o"), os.path.join(td, "mega_vtk.o"),
                 "-L", HERE, "-lmega_hip", "-Wl,-rpath,$ORIGIN/..", "-o", out + ".tmp"]]
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


def build_oracle(verbose: bool = False) -> None:
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)
    if os.path.isdir(REF_SRC) and shutil.which("flang"):
        subprocess.run(["make", "-s", "-j8", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF={REF_SRC}"], check=True)


def build_all(verbose: bool = False) -> None:
    build_hip(verbose)
    build_host(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
