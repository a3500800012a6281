"""Build recipe for the native library ``gradsolve_amd/_lib/libgradsolve_amd.so`` (gfx950).

``hipcc`` compiles every ``csrc/*.hip`` translation unit for gfx950 only and links them into
one shared library in-tree, so the built file travels with the source tree. Packed-FP32
instruction selection is disabled on every HIP compile; nothing here is compiled for any
other GPU architecture.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libgradsolve_amd.so")
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found: the gfx950 library cannot be built")
    return path


# Translation units compiled more than once, with extra flags per variant: the kernel file
# is built with the default fused multiply-add contraction (namespace gsx) and again with
# contraction off and the library's pow and IEEE division (namespace gsx_strict).
STRICT_FLAGS = ["-DGSX_NS=gsx_strict", "-DGSX_STRICT_ARITH", "-ffp-contract=off"]
VARIANTS = {
    "gsx_tsit5.hip": [("", []), ("strict", STRICT_FLAGS)],
    "gsx_rb23.hip": [("", []), ("strict", STRICT_FLAGS)],
    "gsx_rodas5p.hip": [("", []), ("strict", STRICT_FLAGS)],
}


def sources() -> list[tuple[str, str, list[str]]]:
    """(source, object-name suffix, extra flags) for every object of the library."""
    out = []
    for src in sorted(glob.glob(os.path.join(CSRC, "*.hip"))):
        for suffix, extra in VARIANTS.get(os.path.basename(src), [("", [])]):
            out.append((src, suffix, extra))
    return out


# Headers embedded into the library for fields compiled at run time (gsx_runtime.hip hands
# them to hiprtc by name, so the library needs no header files where it runs).
RT_HEADERS = ["gsx_platform.h", "gsx_args.h", "gsx_core.h", "gsx_f32.h", "gsx_rb23.h", "gsx_wide.h",
              "gsx_rb23_f32.h",
              "gsx_tsit5_kernels.h", "gsx_rb23_kernels.h", "gsx_rodas5p_tab.h",
              "gsx_rodas5p.h", "gsx_rodas5p_kernels.h"]
RT_DELIM = "GSXRTHDR"


def write_rt_headers(gendir: str) -> str:
    """Generate gsx_rt_headers.inc: the RT_HEADERS texts as raw string literals."""
    os.makedirs(gendir, exist_ok=True)
    parts = ["// generated by gradsolve_amd/build.py from csrc/ -- do not edit", ""]
    names, srcs = [], []
    for h in RT_HEADERS:
        with open(os.path.join(CSRC, h)) as fh:
            text = fh.read()
        if ")" + RT_DELIM + '"' in text:
            raise RuntimeError(f"{h} contains the raw-string delimiter {RT_DELIM}")
        names.append(f'"{h}"')
        srcs.append(f'R"{RT_DELIM}({text}){RT_DELIM}"')
    parts.append(f"static const int kRtHeaderCount = {len(RT_HEADERS)};")
    parts.append("static const char* const kRtHeaderNames[] = {" + ", ".join(names) + "};")
    parts.append("static const char* const kRtHeaderSources[] = {")
    parts.extend(f"    {x}," for x in srcs)
    parts.append("};")
    out = os.path.join(gendir, "gsx_rt_headers.inc")
    text = "\n".join(parts) + "\n"
    old = None
    if os.path.exists(out):
        with open(out) as fh:
            old = fh.read()
    if old != text:
        with open(out, "w") as fh:
            fh.write(text)
    return gendir


def build(verbose: bool = False, jobs: int = 8) -> str:
    """Compile each translation unit to an object (in parallel), then link the .so."""
    os.makedirs(LIBDIR, exist_ok=True)
    objdir = os.path.join(LIBDIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    gendir = write_rt_headers(os.path.join(objdir, "gen"))
    procs = []
    objs = []
    for src, suffix, extra in sources():
        name = os.path.basename(src) + (f".{suffix}" if suffix else "") + ".o"
        obj = os.path.join(objdir, name)
        objs.append(obj)
        cmd = [hipcc(), *HIPCC_FLAGS, *extra, f"-I{gendir}", "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True)))
        if len(procs) >= jobs:
            _drain(procs)
    _drain(procs)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           "-lhiprtc"]
    if verbose:
        print(" ".join(cmd), flush=True)
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if out.returncode != 0:
        raise RuntimeError(f"link failed:\n{out.stdout}")
    os.replace(tmp, LIB)
    return LIB


def _drain(procs):
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif out and ("warning" in out or "error" in out):
            sys.stderr.write(out)
    procs.clear()
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError(f"hipcc failed:\n{msg}")


if __name__ == "__main__":
    print(build(verbose=True))
