"""Build recipe for libfrhip.so (the HIP/gfx950 FR engine behind include/fr.h).

Compiles in-tree with hipcc; the resulting shared library sits next to this
file so that it travels with the repository.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfrhip.so")

# translation units, compiled in parallel (the stage kernels per precision,
# the adjoint and the batch kernels each in their own)
SOURCES = ["fr_stage64_3d.hip", "fr_stage32_3d.hip", "fr_stage_dbg.hip", "fr_adjoint_v3.hip", "fr_adjoint_v2.hip", "fr_stage64.hip",
           "fr_stage32.hip", "fr_adjoint.hip",
           "fr_batch.hip", "fr_adjoint_v3ad.hip", "fr_api.hip", "fr_host.cpp", "fr_output.cpp", "fr_tri_ops.cpp",
           "fr_tri.hip", "fr_error.hip"]
HEADERS = ["fr_ctx.hpp", "fr_host.hpp", "fr_kernels.hpp", "fr_layout.hpp", "fr_physics.hpp", "fr_adjoint.hpp",
           "fr_adjoint_vis.hpp", "fr_dual.hpp", "fr_batch.hpp", "fr_stage_impl.hpp", "fr_tri.hpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=None, lib=LIB, objdir=None, defines=()):
    """Compile libfrhip.so for gfx950: every translation unit to an object in
    parallel, then one link.  Always rebuilds.  lib / objdir / defines (-D
    switches, e.g. the stage kernel's FR_LDS_PAD) build a variant elsewhere."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = objdir or os.path.join(HERE, "_obj")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(ROOT, "include")]

    def compile_one(src):
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc()] + HIPCC_FLAGS + ["-D" + d for d in defines] + inc + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        return src, obj, r

    jobs = jobs or min(len(SOURCES), max(1, (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        results = list(pool.map(compile_one, SOURCES))
    failed = [(src, r) for src, _, r in results if r.returncode != 0]
    for src, r in failed:
        sys.stderr.write(f"--- {src}\n" + r.stdout + r.stderr)
    if failed:
        raise RuntimeError("hipcc failed building libfrhip.so")
    tmp = lib + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC"] + [obj for _, obj, _ in results] + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking libfrhip.so failed")
    os.replace(tmp, lib)
    return lib


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="build libfrhip.so (or a variant of it elsewhere)")
    ap.add_argument("--lib", default=LIB, help="output path of the shared library")
    ap.add_argument("-D", dest="defines", action="append", default=[],
                    help="a build switch NAME=VALUE, e.g. FR_MFMA=0 for the VALU line contractions")
    args = ap.parse_args()
    objdir = None if args.lib == LIB else args.lib + ".obj"
    print(build(verbose=True, lib=args.lib, objdir=objdir, defines=args.defines))
