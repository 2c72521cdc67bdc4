"""Build recipes: libsdca.so (HIP, gfx950), the sdca CLI, and the CPU oracle.

Everything is compiled in-tree so the built files travel with the repository:
  superdca_amd/libsdca.so        hipcc --offload-arch=gfx950 (kernels + C ABI + host pipeline)
  superdca_amd/sdca              C++ CLI linked against libsdca.so (rpath $ORIGIN)
  oracle/liboracle_plmdca.so     gcc, FP64 CPU restatement (test infrastructure only)
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "superdca_amd")
CSRC = os.path.join(PKG, "csrc")
ORACLE = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 instructions are kept off in all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

LIB_SOURCES = ["sdca_kernels.hip", "sdca_gram.hip", "sdca_subspace.hip", "sdca_pairs.hip", "sdca_mlp.hip", "sdca_noalloc.hip", "sdca_api.cpp", "sdca_warm.cpp",
               "sdca_host.cpp", "sdca_store.cpp"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in proc.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        sys.stderr.write(out + "\n")


def build_lib() -> str:
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    cmds = []
    for src in LIB_SOURCES:
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-I", os.path.join(ROOT, "include"), "-c",
                     os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        list(ex.map(_run, cmds))
    out = os.path.join(PKG, "libsdca.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_cli() -> str:
    out = os.path.join(PKG, "sdca")
    tmp = out + ".tmp"
    _run([HIPCC, "-O3", "-std=c++17", "-I", os.path.join(ROOT, "include"),
          os.path.join(CSRC, "sdca_cli.cpp"), "-L", PKG, "-lsdca",
          "-Wl,-rpath,$ORIGIN", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle() -> str:
    out = os.path.join(ORACLE, "liboracle_plmdca.so")
    tmp = out + ".tmp"
    # -ffp-contract=off: no fused multiply-adds, like the reference's SSE/AVX arithmetic
    _run(["gcc", "-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fPIC", "-shared", "-pthread",
          os.path.join(ORACLE, "plmdca_oracle.c"), "-lm", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    with ThreadPoolExecutor(max_workers=2) as ex:
        f_lib = ex.submit(build_lib)
        f_orc = ex.submit(build_oracle)
        f_lib.result()
        f_orc.result()
    build_cli()


if __name__ == "__main__":
    build_all()
    print("built:", os.path.join(PKG, "libsdca.so"), os.path.join(PKG, "sdca"),
          os.path.join(ORACLE, "liboracle_plmdca.so"))
