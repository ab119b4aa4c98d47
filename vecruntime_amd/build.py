"""Builds libvrgpu.so (gfx950) from vecruntime_amd/csrc with hipcc, in-tree.

    python -m vecruntime_amd.build            # incremental
    python -m vecruntime_amd.build --force    # rebuild every object
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvrgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VR_OFFLOAD_ARCH", "gfx950")

# Packed fp32 instructions are kept out of the generated code (-packed-fp32-ops).
FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]

SOURCES = ["vr_runtime.hip", "vr_strings.hip", "vr_agg.hip", "vr_join.hip", "vr_serve.hip", "vr_jit.hip", "vr_filter.hip",
           "vr_partition.hip"]

# Device headers compiled into the library as text for the run-time specialization (vr_jit.hip):
# (name the kernel source includes them by, path)
JIT_HEADERS = [("../../include/vr.h", os.path.join(os.path.dirname(HERE), "include", "vr.h"))] + [
    (f, os.path.join(CSRC, f)) for f in ("vr_types.h", "vr_device.h", "vr_agg_kernels.h", "vr_part_types.h", "vr_agg_part.h")]
JIT_INC = os.path.join(OBJ, "vr_jit_headers.inc")


def _write_jit_headers():
    """vr_jit_headers.inc: the headers as raw string literals (rewritten only when they change)."""
    parts = ["// generated by vecruntime_amd/build.py from the device headers; do not edit\n"]
    names, srcs = [], []
    for name, path in JIT_HEADERS:
        text = open(path).read()
        assert ')VRJIT"' not in text
        names.append('"%s"' % name)
        srcs.append('R"VRJIT(%s)VRJIT"' % text)
    parts.append("static const char* kJitHeaderNames[] = {%s};\n" % ", ".join(names))
    parts.append("static const char* kJitHeaderSrcs[] = {\n%s};\n" % ",\n".join(srcs))
    out = "".join(parts)
    if not os.path.exists(JIT_INC) or open(JIT_INC).read() != out:
        with open(JIT_INC, "w") as f:
            f.write(out)


def _headers():
    hs = [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    inc = os.path.join(os.path.dirname(HERE), "include")
    hs += [os.path.join(inc, f) for f in os.listdir(inc) if f.endswith(".h")]
    return hs


def _stale(src, obj):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(p) > t for p in [src] + _headers())


def _compile(src, force):
    obj = os.path.join(OBJ, os.path.splitext(os.path.basename(src))[0] + ".o")
    if force or _stale(src, obj):
        cmd = [HIPCC] + FLAGS + ["-I", OBJ, "-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("hipcc failed on %s:\n%s" % (src, r.stderr[-6000:]))
    return obj


def build(force=False, jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    _write_jit_headers()
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    with ThreadPoolExecutor(max_workers=min(jobs, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    if force or not os.path.exists(LIB) or any(os.path.getmtime(o) > os.path.getmtime(LIB) for o in objs):
        tmp = LIB + ".tmp"
        cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-lhiprtc"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("link failed:\n" + r.stderr[-6000:])
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
