"""In-tree build of the native libraries and executables.

    python -m nomad_amd.build            # everything
    python -m nomad_amd.build --only hip

Products (all git-ignored, shipped with the tree):
    nomad_amd/lib/libnomad_hip.so    HIP kernels + C ABI (include/nomad_hip.h)
    nomad_amd/lib/libnomad_tools.so  .dat I/O, generators (include/nomad_tools.h)
    nomad_amd/bin/nomad_float        drop-in host driver, fp32 storage
    nomad_amd/bin/nomad_double       drop-in host driver, fp64 storage
    nomad_amd/bin/nomad_convert, nomad_synthgen_conf, nomad_synthgen, nomad_gen
                                     the reference's data tools (Scripts/convert.py,
                                     synthgen-conf, synthgen, gen.sh)
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
LIBDIR = PKG / "lib"
BINDIR = PKG / "bin"
OBJDIR = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("NOMAD_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: every fused multiply-add in the kernels is written out
# explicitly, so the arithmetic is the sequence the CPU oracle restates.
# -packed-fp32-ops off: no v_pk_*_f32 instructions are generated.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument", f"-I{INCLUDE}", f"-I{CSRC}",
]
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{CSRC}"]

HIP_SOURCES = ["sgd_kernels.hip", "async_kernel.hip", "eval_kernels.hip", "ccd.hip", "objective.hip", "capi.hip",
               "autotune.hip"]
TOOLS_SOURCES = ["datfile.cpp", "synth.cpp", "convert.cpp", "checkpoint.cpp", "tools_capi.cpp"]
# command-line front ends of libnomad_tools.so: one program, one name per reference tool
TOOL_NAMES = ["nomad_convert", "nomad_synthgen_conf", "nomad_synthgen", "nomad_gen"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build_tools(jobs: int) -> Path:
    OBJDIR.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in TOOLS_SOURCES:
        obj = OBJDIR / (Path(src).stem + ".o")
        objs.append(obj)
        cmds.append([CXX, *CXX_FLAGS, "-pthread", "-c", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_run, cmds))
    out = LIBDIR / "libnomad_tools.so"
    _run([CXX, "-shared", "-pthread", *map(str, objs), "-o", str(out)])
    return out


def build_hip(jobs: int) -> Path:
    OBJDIR.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = OBJDIR / (Path(src).stem + ".hip.o")
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(_run, cmds))
    out = LIBDIR / "libnomad_hip.so"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(out)])
    return out


def build_bin() -> list[Path]:
    BINDIR.mkdir(exist_ok=True)
    outs = []
    main = CSRC / "nomad_main.cpp"
    if not main.exists():
        return outs
    for name, storage in (("nomad_float", "NOMAD_STORE_FP32"), ("nomad_double", "NOMAD_STORE_FP64")):
        out = BINDIR / name
        _run([CXX, *CXX_FLAGS, f"-DNOMAD_DEFAULT_STORAGE={storage}", "-pthread", str(main),
              "-L", str(LIBDIR), "-lnomad_hip", "-lnomad_tools",
              f"-Wl,-rpath,$ORIGIN/../lib", "-o", str(out)])
        outs.append(out)
    first = BINDIR / TOOL_NAMES[0]
    _run([CXX, *CXX_FLAGS, str(CSRC / "tools_main.cpp"), "-L", str(LIBDIR), "-lnomad_tools",
          f"-Wl,-rpath,$ORIGIN/../lib", "-o", str(first)])
    outs.append(first)
    for name in TOOL_NAMES[1:]:
        out = BINDIR / name
        out.write_bytes(first.read_bytes())
        out.chmod(0o755)
        outs.append(out)
    return outs


def build_all(only: str | None = None, jobs: int = 8) -> None:
    if only in (None, "tools"):
        build_tools(jobs)
    if only in (None, "hip"):
        build_hip(jobs)
    if only in (None, "bin"):
        build_bin()


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", choices=["tools", "hip", "bin"])
    ap.add_argument("-j", "--jobs", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    build_all(a.only, a.jobs)


if __name__ == "__main__":
    main()
