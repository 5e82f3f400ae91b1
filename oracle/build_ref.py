"""Builds oracle/_ref/libkfref.so: the reference's compute shaders
(shaders/depthToVert.cs, vertToNorm.cs, track.cs, reduce.cs, integrate.cs,
raycast.cs, trackSDF.cs, reduceSDF.cs and mip.cs of the reference tree)
bound into ref_llvmpipe.c, which runs them on Mesa's llvmpipe (headless,
host cores).

Test infrastructure only.  The shader files are read where they lie in the
reference tree at build time; the generated C file lives in a temporary
directory and only the shared library is written, to oracle/_ref/ (kept out
of the repository's history).  Without the reference tree, the Mesa headers
or the swrast driver nothing is built and the tests that use the library
skip.

    python oracle/build_ref.py [reference_root]
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_ref")
OUT = os.path.join(OUT_DIR, "libkfref.so")
SHADERS = ("depthToVert", "vertToNorm", "track", "reduce", "integrate", "raycast", "trackSDF", "reduceSDF",
           "mip")
DRI_HEADER = "/usr/include/GL/internal/dri_interface.h"
SWRAST = "/usr/lib/x86_64-linux-gnu/dri/swrast_dri.so"


def c_string(text):
    """`text` as a C string literal, one source line per literal line."""
    out = []
    for line in text.splitlines(True):
        esc = line.replace("\\", "\\\\").replace('"', '\\"').replace("\r", "").replace("\n", "\\n")
        esc = esc.replace("\t", "\\t")
        out.append('    "%s"' % esc)
    return "\n".join(out) if out else '""'


def build(ref_root="/root/reference", quiet=False):
    """Returns the library path, or None (with the reason printed) when the
    reference tree, the Mesa loader header or the swrast driver is absent."""
    srcdir = os.path.join(ref_root, "shaders")
    missing = [p for p in [srcdir, DRI_HEADER, SWRAST] if not os.path.exists(p)]
    if missing:
        if not quiet:
            print("build_ref: not built, missing %s" % ", ".join(missing))
        return None
    os.makedirs(OUT_DIR, exist_ok=True)
    with tempfile.TemporaryDirectory() as tmp:
        gen = os.path.join(tmp, "kfr_shaders.c")
        with open(gen, "w") as f:
            f.write("/* generated: the reference's compute shader sources */\n")
            for name in SHADERS:
                with open(os.path.join(srcdir, name + ".cs"), encoding="utf-8", errors="replace") as s:
                    text = s.read()
                f.write("const char *kfr_src_%s =\n%s;\n" % (name, c_string(text)))
        tmp_out = os.path.join(tmp, "libkfref.so")
        cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fvisibility=hidden", "-std=gnu11", "-Wall",
               os.path.join(HERE, "ref_llvmpipe.c"), gen, "-ldl", "-o", tmp_out]
        subprocess.run(cmd, check=True)
        os.replace(tmp_out, OUT)
    if not quiet:
        print("build_ref: wrote", OUT)
    return OUT


if __name__ == "__main__":
    sys.exit(0 if build(*sys.argv[1:2]) else 1)
