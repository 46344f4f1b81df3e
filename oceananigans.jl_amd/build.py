"""Build the HIP library in-tree: oceananigans.jl_amd/liboceananigans_mi355x.so (gfx950)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "liboceananigans_mi355x.so")
SOURCES = ["om_api.cpp", "om_kernels.hip", "om_tend_tiled.hip", "om_fft.hip", "om_cg.hip", "om_closure_tiled.hip",
           "om_diagnostics.hip", "om_learned.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions are disabled on this path
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


# (source, extra flags, object suffix): each .hip source is compiled once per precision.  The FP64
# object is built without FMA contraction (fused multiply-adds only where the source writes fma, at
# the reference's @muladd sites), so that its rounding follows the reference's.  The FP32 object
# contracts a·b + c within one expression only: whether a product and a sum fuse is then decided
# by the source, not by how a kernel's code is laid out, so one kernel body compiled into the
# single-model launch and into the ensemble's member-table launch rounds identically.
UNITS = [("om_api.cpp", [], "")]
for _s in SOURCES[1:]:
    UNITS.append((_s, ["-DOM_ONLY_FLOAT", "-ffp-contract=on"], "_f32"))
    UNITS.append((_s, ["-DOM_ONLY_DOUBLE", "-ffp-contract=off"], "_f64"))


def build(verbose=False):
    from concurrent.futures import ThreadPoolExecutor
    jobs = []
    for s, extra, suffix in UNITS:
        src = os.path.join(CSRC, s)
        obj = os.path.join(CSRC, os.path.splitext(s)[0] + suffix + ".o")
        jobs.append(([HIPCC, *FLAGS, *extra, "-x", "hip", "-c", src, "-o", obj], obj))

    def run(job):
        cmd, obj = job
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        return obj

    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(run, jobs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    for o in objs:
        os.remove(o)
    return OUT


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(OUT)
