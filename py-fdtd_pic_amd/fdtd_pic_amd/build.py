"""Build libfdtdpic.so in-tree for gfx950 (hipcc, no CMake).

    python -m fdtd_pic_amd.build            # from py-fdtd_pic_amd/

Every HIP translation unit under csrc/ is compiled (in parallel) with
`--offload-arch=gfx950 -O3` and `-target-feature -packed-fp32-ops` (no packed
FP32 VALU ops are generated), all but the FP32 field units with
`-ffp-contract=off` (only the fused products written out in the source), and
linked into fdtd_pic_amd/libfdtdpic.so.
"""
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
CSRC = os.path.join(PKG_ROOT, "csrc")
OUT = os.path.join(HERE, "libfdtdpic.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]
# -ffp-contract=off: the compiler fuses no product on its own; the FP64 field
# update's fused products are written out (fmr), so an FP64 cell's result does
# not depend on which tile-class loop or (K, S) tiling computed it.  The FP32
# field units keep contraction (fewer register moves in their step loop).
NO_CONTRACT = ["-ffp-contract=off"]


def flags_for(src):
    """Compile flags of one translation unit."""
    if os.path.basename(src).startswith("stencil_f32_"):
        return FLAGS
    return FLAGS + NO_CONTRACT


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r.returncode


def build(verbose=False, jobs=None):
    """Compile every translation unit in parallel (the field-step instances
    are split by precision and slab kind), then link the shared library."""
    import tempfile
    from concurrent.futures import ThreadPoolExecutor
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, os.cpu_count() or 1))
    with tempfile.TemporaryDirectory(prefix="fdtdpic_build_") as tmpd:
        objs = [os.path.join(tmpd, os.path.basename(src)[:-4] + ".o") for src in srcs]
        cmds = [[HIPCC] + flags_for(src) + ["-c", src, "-o", obj] for src, obj in zip(srcs, objs)]
        if verbose:
            for c in cmds:
                print(" ".join(c), flush=True)
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            codes = list(ex.map(_run, cmds))
        if any(codes):
            raise RuntimeError("hipcc failed")
        tmp = OUT + ".tmp"
        link = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", tmp]
        if verbose:
            print(" ".join(link), flush=True)
        if _run(link):
            raise RuntimeError("hipcc link failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
