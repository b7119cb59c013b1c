"""The C ABI library loads, exports every entry point include/*.h
declares, and its host-only parts work without a GPU."""
import ctypes as C
import json
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import bsglmm_amd
from bsglmm_amd import BsgError, Layout, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")


def header_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(bsg_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_function():
    lib = C.CDLL(bsglmm_amd.LIB_PATH)
    names = header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bsglmm_amd.lib()  # every binding resolves too


def test_kiss_and_devseed_match_reference():
    with open(os.path.join(GOLD, "ref_kiss.json")) as fh:
        f = json.load(fh)
    seed = list(f["seed"])
    for h in f["draws_hex"][:200]:
        assert bsglmm_amd.kiss(seed) == float.fromhex(h)
    seed = list(f["seed"])
    assert bsglmm_amd.devseed(seed) == f["devseed"]
    assert seed == f["state_after_devseed"]


def test_opts_defaults_follow_the_gpu_path():
    o = bsglmm_amd.default_opts()
    assert (o.nchains, o.xtz_mode, o.df_offset) == (1, 0, 19)  # df = V+19, covarGPU.cu:1104
    assert abs(o.trunc_eps - 1e-6) < 1e-12 and abs(o.pivot_min - 1e-6) < 1e-12


def test_create_rejects_bad_arguments():
    d = synth.generate("c1", nsubs=8, ncovar=2)
    nx, ny, nz = d["grid"]
    L = Layout(d["mask"], nx, ny, nz)
    with pytest.raises(BsgError):
        bsglmm_amd.Sampler(L, np.ones((9, 8), np.float32), np.zeros((8, L.totvox), np.uint8))  # P > 8
    with pytest.raises(BsgError):
        bsglmm_amd.Sampler(L, d["covar"], d["data"], xtz_mode=7)
    with pytest.raises(BsgError):
        bsglmm_amd.Sampler(L, d["covar"], d["data"], nchains=0)
    # the BF16 MFMA X'Z inside the draws (xtz_mode 3) keeps no Z, draws by inversion, takes no WM image
    for kw in (dict(z_store=0), dict(z_store=1), dict(z_store=2, latent_method=1)):
        with pytest.raises(BsgError, match="xtz_mode 3"):
            bsglmm_amd.Sampler(L, d["covar"], d["data"], xtz_mode=3, **kw)
    with pytest.raises(BsgError, match="xtz_mode 3"):
        bsglmm_amd.Sampler(L, d["covar"], d["data"], wm=np.ones(L.totvox, np.float32), xtz_mode=3, z_store=2)
    with pytest.raises(BsgError, match="z_store"):
        bsglmm_amd.Sampler(L, d["covar"], d["data"], xtz_mode=1, z_store=2)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
