"""Build recipes for the native libraries (run on the build host, not at import).

libkfusion.so  -- HIP kernels + C ABI (include/kfusion.h), gfx950 only.
libkfsynth.so  -- host C++ synthetic depth sequences (OpenMP).
kf_gfusion_example -- C++ drop-in example driving the gFusion class.
"""
import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-fPIC",
    "-std=c++17",
    # FP32 results follow the reference expression order bit for bit
    "-ffp-contract=off",
    # packed FP32 instructions are not used on this hardware
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_kfusion():
    srcs = [os.path.join(CSRC, f) for f in ("kf_kernels.hip", "kf_capi.hip", "kf_sdf.hip", "kf_mcubes.hip", "kf_io.cpp")]
    srcs = [s for s in srcs if os.path.exists(s)]
    out = os.path.join(PKG, "libkfusion.so")
    tmp = out + ".tmp"
    _run([HIPCC] + HIP_FLAGS + ["-shared", "-Wl,--no-undefined"] + srcs + ["-o", tmp])
    os.replace(tmp, out)
    return out


def build_synth():
    out = os.path.join(PKG, "libkfsynth.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-fPIC", "-shared", "-fopenmp", "-std=c++17",
          os.path.join(CSRC, "kf_synth.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_example():
    src = os.path.join(CSRC, "kf_gfusion_example.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(PKG, "kf_gfusion_example")
    _run([HIPCC, "--offload-arch=gfx950", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
          "-I", CSRC, src, "-o", out, "-L", PKG, "-lkfusion", "-lkfsynth",
          "-Wl,-rpath,$ORIGIN"])
    return out


def build_all():
    build_kfusion()
    build_synth()
    build_example()
