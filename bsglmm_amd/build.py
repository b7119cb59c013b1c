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


# Host-side AddressSanitizer + UndefinedBehaviorSanitizer (the device code is
# built as usual). The library is linked without the sanitizer runtime, which
# the process that loads it preloads, e.g.
#   BSGLMM_LIB=<out> LD_PRELOAD=<clang>/lib/linux/libclang_rt.asan-x86_64.so \
#   ASAN_OPTIONS=detect_leaks=0 python -m pytest tests/test_io.py tests/test_layout.py
HOST_SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                 "-Xarch_host", "-fno-omit-frame-pointer", "-g"]


def build(verbose=False, jobs=None, host_sanitize_out=None):
    """Compile every source to an object, then link libbsglmm.so (or, with
    host_sanitize_out, a host-sanitized library at that path, and no BinCar)."""
    cc = hipcc()
    objdir = os.path.join(HERE, "_build_hostsan" if host_sanitize_out else "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        flags = HIP_FLAGS + (HOST_SANITIZE if host_sanitize_out else [])
        cmd = [cc] + flags + ["-I", os.path.join(ROOT, "include"), "-c",
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
    if host_sanitize_out:
        subprocess.run([cc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", host_sanitize_out] + objs + ["-lz"],
                       check=True)
        return host_sanitize_out
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
    out = None
    if "--host-sanitize" in sys.argv:
        out = os.path.abspath(sys.argv[sys.argv.index("--host-sanitize") + 1])
    print(build(verbose="-v" in sys.argv, host_sanitize_out=out))
