"""Build libbsglmm.so (HIP, gfx950) in-tree with hipcc.

Run as ``python -m bsglmm_amd.build`` or through ``__graft_entry__.build()``.
The shared library lands next to this file so it travels with the tree.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libbsglmm.so")
SOURCES = ["bsg_layout.cpp", "bsg_ctx.cpp", "bsg_io.cpp", "bsg_cli.cpp", "bsg_kernels.hip", "bsg_probe.hip"]
EXE = os.path.join(HERE, "BinCar")

# Packed FP32 instructions are disabled for all device code: some of their
# forms have given wrong results on these machines.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in ordinary VGPRs: no AGPR copies around each MFMA
    "-mllvm", "-amdgpu-mfma-vgpr-form=1",
]


def hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False, jobs=None):
    """Compile every source to an object, then link libbsglmm.so."""
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [cc] + HIP_FLAGS + ["-I", os.path.join(ROOT, "include"), "-c",
                                  os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 feature; that warning is harmless
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and l.strip()]
        if p.returncode != 0:
            failed.append(src)
            print("\n".join(lines), file=sys.stderr)
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError("hipcc failed on: " + ", ".join(failed))
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    # the BinCar executable: a main() over the library, found next to it at run time
    tmp = EXE + ".tmp"
    cmd = ["g++", "-O2", "-o", tmp, os.path.join(CSRC, "bincar_exe.cpp"), "-L", HERE, "-lbsglmm",
           "-Wl,-rpath,$ORIGIN"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, EXE)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
