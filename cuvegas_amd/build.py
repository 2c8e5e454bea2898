"""Build libcuvegas_amd.so in-tree with hipcc for gfx950.

Steps:
  1. embed vg_fill.hip, vg_philox.h and vg_math.h as C string literals (csrc/vg_embed.inc),
     so the library carries the sampling-kernel source it hands to hipRTC;
  2. compile the ahead-of-time step kernels (vg_steps.hip), the backward replay
     entry point that allocates nothing (vg_noalloc.hip), the probes of the
     device's instruction costs (vg_rates.hip) and the host driver (vg_host.cpp);
  3. link the shared library against libhiprtc;
  4. build the standalone program cuvegas_main (vg_main.cpp, with the integrand
     library vg_functions.h embedded) against the shared library.

Usage: python cuvegas_amd/build.py   (run as a file: importing the package needs the library)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcuvegas_amd.so")
MAIN = os.path.join(HERE, "cuvegas_main")
BUILD = os.path.join(HERE, "_build")
ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# Packed FP32 instructions are disabled for every device compile.
DEVICE_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _embed(names_and_files, out_path):
    parts = ["// Generated by cuvegas_amd/build.py -- do not edit.\n"]
    for name, path in names_and_files:
        with open(path, "r", encoding="utf-8") as f:
            text = f.read()
        delim = "VGSRC"
        assert f"){delim}\"" not in text
        # split into chunks to stay clear of string-literal length limits
        chunks = [text[i:i + 12000] for i in range(0, len(text), 12000)] or [""]
        body = "\n".join(f'R"{delim}({c}){delim}"' for c in chunks)
        parts.append(f"static const char {name}[] =\n{body};\n")
    new = "".join(parts)
    old = None
    if os.path.exists(out_path):
        with open(out_path, "r", encoding="utf-8") as f:
            old = f.read()
    if old != new:
        with open(out_path, "w", encoding="utf-8") as f:
            f.write(new)


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build(verbose: bool = True) -> str:
    os.makedirs(BUILD, exist_ok=True)
    _embed([("kVgFillSrc", os.path.join(CSRC, "vg_fill.hip")),
            ("kVgPhiloxSrc", os.path.join(CSRC, "vg_philox.h")),
            ("kVgMathSrc", os.path.join(CSRC, "vg_math.h"))],
           os.path.join(CSRC, "vg_embed.inc"))
    steps_o = os.path.join(BUILD, "vg_steps.o")
    host_o = os.path.join(BUILD, "vg_host.o")
    noalloc_o = os.path.join(BUILD, "vg_noalloc.o")
    rates_o = os.path.join(BUILD, "vg_rates.o")
    _run([HIPCC, *DEVICE_FLAGS, "-c", os.path.join(CSRC, "vg_steps.hip"), "-o", steps_o])
    _run([HIPCC, *DEVICE_FLAGS, "-c", os.path.join(CSRC, "vg_noalloc.hip"), "-o", noalloc_o])
    _run([HIPCC, *DEVICE_FLAGS, "-c", os.path.join(CSRC, "vg_rates.hip"), "-o", rates_o])
    _run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall",
          "-Wno-unused-result", "-c", "-x", "c++", os.path.join(CSRC, "vg_host.cpp"),
          "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-o", host_o])
    tmp = OUT + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", steps_o, noalloc_o, rates_o, host_o,
          "-o", tmp, "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, OUT)
    # 4. the standalone program (the reference's src/main.cu caller) over the C ABI
    _embed([("kVgFunctionsSrc", os.path.join(CSRC, "vg_functions.h"))], os.path.join(CSRC, "vg_main_embed.inc"))
    tmp = MAIN + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", os.path.join(CSRC, "vg_main.cpp"), "-o", tmp,
          f"-L{HERE}", "-lcuvegas_amd", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, MAIN)
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
