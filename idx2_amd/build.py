"""Build libidx2g.so (HIP, gfx950) in-tree: idx2_amd/_build/libidx2g.so."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libidx2g.so")
SRC = [os.path.join(HERE, "csrc", f) for f in ("idx2g.hip", "layout.cpp", "common.cpp")]
DEPS = SRC + [os.path.join(HERE, "csrc", f) for f in (
    "device.hpp", "kernels.hpp", "brick32.hpp", "code32.hpp", "dec32.hpp", "brick2d.hpp", "code2d.hpp",
    "dec2d.hpp", "host.hpp", "layout.hpp", "encode_kernels.hip",
    "decode_kernels.hip", "encode.inc", "files.inc", "decode.inc", "api.inc", "server.inc", "tune.inc",
    "noalloc.inc", "pool.inc")] + [os.path.join(ROOT, "include", "idx2g.h")]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # FP64 lifting must follow the reference's operation order: no contraction
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I" + os.path.join(ROOT, "include"),
]


# Profiling variant with the encode kernel's phase clocks (-DIDX2G_PHASE_TIMERS,
# idx2g_phase_times); never loaded by the product.
PHASE_DIR = os.path.join(HERE, "_build_phase")
PHASE_LIB = os.path.join(PHASE_DIR, "libidx2g.so")


def build(force=True, verbose=False, phase_timers=False):
    out_dir, lib = (PHASE_DIR, PHASE_LIB) if phase_timers else (OUT_DIR, LIB)
    flags = FLAGS + (["-DIDX2G_PHASE_TIMERS"] if phase_timers else [])
    os.makedirs(out_dir, exist_ok=True)
    if not force and os.path.exists(lib):
        t = os.path.getmtime(lib)
        if all(os.path.getmtime(d) <= t for d in DEPS):
            return lib
    objs = []
    for src in SRC:
        obj = os.path.join(out_dir, os.path.basename(src) + ".o")
        cmd = [HIPCC, *flags, "-c", "-x", "hip" if src.endswith(".hip") else "c++", src, "-o", obj]
        if src.endswith(".cpp"):
            cmd = [HIPCC, *[f for f in flags if f != "-shared"], "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        subprocess.run(cmd, check=True)
        objs.append(obj)
    tmp = lib + ".tmp"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp, "-ldl"], check=True)
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    print(build(verbose=True))
