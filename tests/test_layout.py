"""Layout (main.cu:189-394, 471-502): the library's C ABI against the oracle,
invariants that pin the colouring, and the bundled mask's counts. CPU only:
bsg_layout_* never touches the GPU."""
import os

import numpy as np
import pytest

import oracle
from bsglmm_amd import BsgError, Layout, synth

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KEYS = ["host_idx", "host_idx_sc", "vox0", "vox1", "nbrs0", "nbrs1", "nbr6"]


def random_mask(rng, shape, p):
    m = rng.random(shape) < p
    return np.pad(m.astype(np.uint8), 1)


def masks():
    rng = np.random.default_rng(11)
    out = []
    for name in ("c1", "c2"):
        g, spec = synth.CONFIGS[name]["grid"], synth.CONFIGS[name]["mask"]
        m, _ = synth.make_mask(g, spec)
        out.append((name, synth.pad_mask(m), g))
    for k, (shape, p) in enumerate([((5, 6, 7), 0.5), ((3, 1, 4), 0.9), ((1, 1, 1), 1.0), ((9, 4, 2), 0.2)]):
        out.append((f"random{k}", random_mask(rng, shape, p), shape[::-1]))
    full = np.pad(np.ones((4, 3, 5), np.uint8), 1)
    out.append(("full", full, (5, 3, 4)))
    return out


@pytest.mark.parametrize("name,mask,grid", masks(), ids=[m[0] for m in masks()])
def test_library_layout_equals_oracle(name, mask, grid):
    nx, ny, nz = grid
    ref = oracle.layout(mask, nx, ny, nz)
    L = Layout(mask, nx, ny, nz)
    assert (L.totvox, L.totvoxp, L.n0, L.n1) == (ref["totvox"], ref["totvoxp"], ref["n0"], ref["n1"])
    got = L.export()
    for k in KEYS:
        assert np.array_equal(got[k], ref[k]), k


@pytest.mark.parametrize("name,mask,grid", masks(), ids=[m[0] for m in masks()])
def test_layout_invariants(name, mask, grid):
    nx, ny, nz = grid
    L = Layout(mask, nx, ny, nz)
    e = L.export()
    sx, sxy = nx + 2, (nx + 2) * (ny + 2)
    flat = mask.reshape(-1)
    ids = np.flatnonzero(flat)
    # mask rank in raster order
    assert np.array_equal(e["host_idx"][ids], np.arange(len(ids)))
    assert np.all(e["host_idx"][flat == 0] == -1)
    # colour = parity of i+j+k (1-based padded coordinates): colour 0 iff odd
    for c, vox in ((0, e["vox0"]), (1, e["vox1"])):
        k, r = np.divmod(vox, sxy)
        j, i = np.divmod(r, sx)
        assert np.all(((i + j + k) % 2 == 1) == (c == 0))
        assert np.all(np.diff(vox) > 0)  # raster order within a colour
    assert np.array_equal(np.sort(np.concatenate([e["vox0"], e["vox1"]])), ids)
    # neighbours always have the other colour; counts add up to twice the pairs
    offs = [1, sx, sxy]
    pairs = sum(int(np.sum(flat[ids + o] != 0)) for o in offs)
    assert int(e["nbrs0"].sum()) + int(e["nbrs1"].sum()) == 2 * pairs
    n0 = L.n0
    nb = e["nbr6"].reshape(-1, 6)
    colour = np.r_[np.zeros(n0, int), np.ones(L.totvox - n0, int)]
    for d in range(6):
        k = nb[:, d]
        ok = k >= 0
        assert np.all(colour[k[ok]] != colour[ok])
    assert np.array_equal((nb >= 0).sum(1), np.r_[e["nbrs0"], e["nbrs1"]])
    # every in-mask voxel and each 6-neighbour has a dilated-mask rank
    for o in (-sxy, -sx, -1, 0, 1, sx, sxy):
        assert np.all(e["host_idx_sc"][ids + o] >= 0)
    assert L.totvoxp == int(np.sum(e["host_idx_sc"] >= 0))


def test_bundled_mask_counts():
    """bsglmm/images/mask.nii.gz (bit-packed fixture): the counts an independent
    restatement of main.cu:189-394 gave for it."""
    f = np.load(os.path.join(GOLD, "bundled_mask.npz"))
    nx, ny, nz = (int(v) for v in f["grid"])
    mask = np.unpackbits(f["bits"])[: int(f["nsize"])].astype(np.uint8)
    L = Layout(mask, nx, ny, nz)
    assert (L.totvox, L.totvoxp, L.n0, L.n1) == (274596, 294322, 137280, 137316)
    e = L.export()
    hist = np.bincount(np.r_[e["nbrs0"], e["nbrs1"]], minlength=7)
    assert dict(enumerate(hist.tolist())) == {0: 0, 1: 0, 2: 0, 3: 4184, 4: 5465, 5: 9448, 6: 255499}
    assert int(np.sum(hist * np.arange(7))) // 2 == 807323


def test_layout_rejects_bad_masks():
    with pytest.raises(BsgError):
        Layout(np.zeros(5 * 5 * 5, np.uint8), 3, 3, 3)  # empty mask
    m = np.zeros((5, 5, 5), np.uint8)
    m[0, 2, 2] = 1  # on the padding border
    with pytest.raises(BsgError):
        Layout(m, 3, 3, 3)
    with pytest.raises(ValueError):
        Layout(np.zeros(10, np.uint8), 3, 3, 3)


def test_synthetic_configurations():
    for name, V in (("c1", 7056), ("c2", 230906)):
        d = synth.generate(name)
        assert d["data"].shape == (synth.CONFIGS[name]["nsubs"], V)
        prev = d["data"].mean()
        assert 0.01 < prev < 0.12, prev
        assert d["covar"].shape == (synth.CONFIGS[name]["ncovar"], synth.CONFIGS[name]["nsubs"])
        assert np.all(d["covar"][0] == 1)
