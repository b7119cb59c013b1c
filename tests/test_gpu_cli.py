"""The BinCar program end to end on a small synthetic study, plus the two
per-save summaries it uses (plane means, per-subject log-likelihood).

The CLI is checked against the library driven directly with the same seed:
the sweep is deterministic given the Philox key (devseed of seed.dat,
main.cu:413), so the posterior maps it writes must equal the accumulators of
a Sampler run through the same iterations (mcmc.cpp:311, 584: sweeps 0..MaxIter,
saves after the sweep of every iteration above BurnIn that SAVE_ITER divides).
The empirical maps are recomputed from the data; the log-likelihood is
checked in FP64 with the same A&S 26.2.17 CDF, itself pinned against
scipy's normal CDF to its published 7.5e-8 bound.
"""
import os
import re
import subprocess

import numpy as np
import pytest
import scipy.special as sp

import bsglmm_amd as B
from bsglmm_amd import Layout, Sampler, synth
from nifti_fixture import make_nifti

pytestmark = pytest.mark.gpu
BINCAR = os.path.join(os.path.dirname(B.LIB_PATH), "BinCar")
SEED3 = [5118772087888418775, 289433066200759127, 1964621368]


def study(tmp_path, nsubs=40):
    """C1 mask and data; design with two groups (NTypes 2) and age."""
    d = synth.generate("c1", nsubs=nsubs, ncovar=3)
    nx, ny, nz = d["grid"]
    mask = d["mask"].reshape(nz + 2, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1].astype(np.float64)
    img = tmp_path / "images"
    img.mkdir()
    make_nifti(str(img / "mask.nii.gz"), mask, 64)
    L = Layout(d["mask"], nx, ny, nz)
    e = L.export()
    inner = e["host_idx"].reshape(nz + 2, ny + 2, nx + 2)[1:-1, 1:-1, 1:-1]
    inside = inner >= 0
    rng = np.random.default_rng(4)
    wm8 = rng.integers(0, 256, size=mask.shape).astype(np.uint8)
    make_nifti(str(img / "wm.nii"), wm8.astype(np.float64), 2)
    wm = np.zeros(L.totvox, np.float32)
    wm[inner[inside]] = wm8[inside].astype(np.float32) / 255.0
    grp = (np.arange(nsubs) >= nsubs // 2).astype(np.float32)
    covar = np.vstack([1 - grp, grp, d["covar"][1]]).astype(np.float32)
    rows = ["ID grpA grpB age Image"]
    for i in range(nsubs):
        vol = np.zeros(mask.shape, np.float64)
        vol[inside] = d["data"][i][inner[inside]]
        make_nifti(str(img / f"s{i}.nii.gz"), vol, 16)  # FLOAT32 lesion map
        rows.append(f"sub{i} {covar[0, i]:g} {covar[1, i]:g} {float(covar[2, i])!r} s{i}.nii.gz")
    (tmp_path / "design.txt").write_text("\n".join(rows) + "\n")
    (tmp_path / "seed.dat").write_text(" ".join(str(v) for v in SEED3) + "\n")
    return d, L, inner, inside, wm, covar


def read_map(path, inner, inside):
    vol = B.nifti_read(str(path))
    return vol, vol[inside], inner[inside]


def test_bincar_end_to_end_matches_library(tmp_path):
    print("case BinCar c1 N=40 NTypes=2 NCov=3 WM, 10200/200", flush=True)
    d, L, inner, inside, wm, covar = study(tmp_path)
    r = subprocess.run([BINCAR, "2", "3", "1", "design.txt", "1", "wm.nii", "10200", "200"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-500:], flush=True)
    assert r.returncode == 0
    assert "TOTSAV = 10000" in r.stdout and "iter = 10200" in r.stdout
    m = re.search(r"DE = (\S+) ED = (\S+) PD = (\S+) DIC = (\S+)", (tmp_path / "DIC.dat").read_text())
    DE, ED, PD, DIC = (float(v) for v in m.groups())
    assert np.isfinite([DE, ED, PD]).all() and abs(DIC - (DE + 2 * PD)) < 1e-3 and abs(PD - (ED - DE)) < 1e-3
    assert (tmp_path / "seed.dat").read_text().split() != [str(v) for v in SEED3]

    # the same chain through the library
    seed = list(SEED3)
    key = B.devseed(seed)
    N, V = covar.shape[1], L.totvox
    S = Sampler(L, covar, d["data"], wm=wm, seed=key, xtz_mode=2, z_store=2)
    S.sweep(201)
    S.accum_reset(ntypes=2, nbatches=50, batch_len=200)
    S.sweep_accumulate(10000, 1)
    post = S.accum_get(0)
    names = ["grpA", "grpB", "age"]
    for c, nm in enumerate(names):
        vol, vin, rank = read_map(tmp_path / f"spatCoef_{nm}.nii.gz", inner, inside)
        assert np.all(vol[~inside] == 0)
        assert np.array_equal(vin, post["mean"][c][rank])
        _, var_in, _ = read_map(tmp_path / f"spatCoef_{nm}.Var.nii.gz", inner, inside)
        assert np.array_equal(var_in, post["var"][c][rank]) and np.all(var_in > 0)
        _, sc_in, _ = read_map(tmp_path / f"standCoef_{nm}.nii.gz", inner, inside)
        assert np.allclose(sc_in, post["mean"][c][rank] / np.sqrt(post["var"][c][rank]), rtol=1e-12)
    for t, nm in enumerate(names[:2]):
        _, p_in, rank = read_map(tmp_path / f"prb_{nm}.nii.gz", inner, inside)
        assert np.array_equal(p_in, post["prb"][t][rank]) and np.all((p_in >= 0) & (p_in <= 1))
    _, bwm_in, rank = read_map(tmp_path / "bWM.nii.gz", inner, inside)
    assert np.array_equal(bwm_in, post["bwm"][rank])

    # empirical maps (read_data.cpp:418-487)
    y = d["data"].astype(np.float64)
    tot = np.zeros(V)
    for t, nm in enumerate(names[:2]):
        sel = covar[t] != 0
        s = y[sel].sum(0)
        tot += s
        _, e_in, rank = read_map(tmp_path / f"empir_prb_{nm}.nii.gz", inner, inside)
        assert np.array_equal(e_in, (s / sel.sum())[rank])
    assert np.array_equal(read_map(tmp_path / "total_lesion_cnt.nii.gz", inner, inside)[1], tot[rank])
    assert np.array_equal(read_map(tmp_path / "total_empir_prb.nii.gz", inner, inside)[1], (tot / N)[rank])

    # Qhat: N rows of id, NTypes probabilities, predicted and true subtype
    q = np.loadtxt(tmp_path / "Qhat.dat")
    assert q.shape == (N, 2 + 2 + 1)
    assert np.array_equal(q[:, 0], np.arange(N)) and np.allclose(q[:, 1:3].sum(1), 1.0, atol=2e-6)
    assert np.array_equal(q[:, 4], (np.arange(N) >= N // 2).astype(float))
    assert np.loadtxt(tmp_path / "Qhat2.dat").shape == (N, 5)


def test_bincar_usage_and_input_errors(tmp_path):
    r = subprocess.run([BINCAR], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "Usage:" in r.stdout
    d, L, inner, inside, wm, covar = study(tmp_path, nsubs=6)
    # WM image must be UINT8 (read_data.cpp:503-507)
    make_nifti(str(tmp_path / "images" / "wm64.nii"), np.zeros(inner.shape), 64)
    r = subprocess.run([BINCAR, "2", "3", "1", "design.txt", "1", "wm64.nii", "1000", "500"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "Must be unsigned char (NIFTI_TYPE_UINT8)" in r.stdout
    r = subprocess.run([BINCAR, "2", "3", "1", "design.txt", "nomask.nii", "0", "1000", "500"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "Cannot open" in r.stdout


def phi_as(x):
    """A&S 26.2.17 (ProbSDNorm, covar.cpp:327-348) in FP64."""
    t = 1.0 / (1.0 + 0.2316419 * np.abs(x))
    poly = t * (0.319381530 + t * (-0.356563782 + t * (1.781477937 + t * (-1.821255978 + t * 1.330274429))))
    tail = 0.39894228 * np.exp(-0.5 * x * x) * poly
    return np.where(x >= 0, 1.0 - tail, tail)


def test_loglik_and_coef_means_against_fp64():
    print("case c1 loglik / coef means, 2 chains", flush=True)
    d = synth.generate("c1", nsubs=30, ncovar=4)
    nx, ny, nz = d["grid"]
    L = Layout(d["mask"], nx, ny, nz)
    rng = np.random.default_rng(2)
    wm = rng.uniform(0, 1, L.totvox).astype(np.float32)
    S = Sampler(L, d["covar"], d["data"], wm=wm, seed=3, nchains=2, xtz_mode=2, z_store=2)
    S.sweep(20)
    e = L.export()
    hi, hs = e["host_idx"], e["host_idx_sc"]
    ids = np.flatnonzero(hi >= 0)
    col = np.empty(L.totvox, np.int64)
    col[hi[ids]] = hs[ids]
    xs = np.linspace(-9, 9, 20001)
    assert np.max(np.abs(phi_as(xs) - sp.ndtr(xs))) < 7.5e-8
    X = d["covar"].astype(np.float64)
    y = d["data"].astype(bool)
    for chain in range(2):
        beta = S.get_spat_coef(chain)[:, col].astype(np.float64)
        b = S.get_wm_coef(chain)
        assert np.allclose(S.coef_means(chain), beta.mean(1), rtol=1e-12, atol=1e-14)
        for T in (1, 2):
            ll = S.loglik(T, chain)
            rest = b * wm.astype(np.float64)[None, :] + X[T:].T @ beta[T:]
            for t in range(T):
                p = phi_as(beta[t][None, :] + rest)
                want = np.where(y, np.log(p + 1e-35), np.log(1 - p + 1e-35)).sum(1)
                assert np.allclose(ll[:, t], want, rtol=2e-5), (chain, T, t)
    # closed form: b = 0 and all coefficients 0 give p = 1/2 everywhere
    S2 = Sampler(L, d["covar"], d["data"], seed=3)
    assert np.allclose(S2.loglik(2), L.totvox * np.log(0.5), rtol=1e-6)


def test_bincar_checkpoint_and_restart_equal_one_run(tmp_path):
    """A run stopped after iteration 4321 (-checkpoint 1000 -stop 4321) and
    continued with -restart writes every output file the uninterrupted run
    writes, with the same contents: maps, Qhat, DIC and seed.dat."""
    print("case BinCar c1 N=40 NTypes=2 NCov=3 WM 10200/200: one run vs stop at 4321 + restart", flush=True)
    one, two = tmp_path / "one", tmp_path / "two"
    one.mkdir()
    two.mkdir()
    study(one)
    study(two)
    args = [BINCAR, "2", "3", "1", "design.txt", "1", "wm.nii", "10200", "200"]
    r = subprocess.run(args, cwd=one, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-800:]
    r = subprocess.run(args + ["-checkpoint", "1000", "-stop", "4321"], cwd=two, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-800:]
    assert "stopped after iteration 4321" in r.stdout
    assert not (two / "DIC.dat").exists()
    for f in ("last_state.bsg", "last_run.dat", "last_parms.dat", "last_spatCoef.dat"):
        assert (two / f).exists(), f
    npar = len((two / "last_parms.dat").read_text().split())
    assert npar == 1 + 3 + 9  # b, the plane means, Lambda (save_iter, mcmc.cpp:1265-1286)
    r = subprocess.run(args + ["-restart"], cwd=two, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-800:]
    assert "RESTART at iteration 4322" in r.stdout and "TOTSAV = 10000" in r.stdout
    outs = sorted(p.name for p in one.iterdir() if p.is_file() and p.suffix in (".gz", ".dat"))
    assert {"DIC.dat", "Qhat.dat", "seed.dat", "spatCoef_age.nii.gz", "bWM.nii.gz"} <= set(outs)
    for name in outs:
        a, b = one / name, two / name
        if name.endswith(".nii.gz"):
            assert np.array_equal(B.nifti_read(str(a)), B.nifti_read(str(b))), name
        else:
            assert a.read_text() == b.read_text(), name
    # a restart needs the same run lengths
    r = subprocess.run(args[:-2] + ["10400", "200", "-restart"], cwd=two, capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "MaxIter" in r.stdout
