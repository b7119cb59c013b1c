"""Synthetic lesion data for the benchmark and test configurations.

The reference ships no data set (bsglmm/images holds only a mask and a WM
prior), so every run here uses data drawn from the model itself:

  * mask: an ellipsoid, or an ellipsoidal shell, on a NIfTI-ordered grid
    (x fastest), padded by one voxel as read_nifti1_mask builds it
    (bsglmm/read_data.cpp:227-240);
  * design X: intercept, standardised age, sex ~ Bernoulli(0.5) and further
    covariates per configuration, stored covariate-major as
    bsglmm/read_data.cpp:299-302 stores it;
  * true coefficient maps: iid N(0,1) fields smoothed with a Gaussian of the
    given FWHM, standardised over the mask and scaled;
  * Z = x_i' beta_j + N(0,1), y = 1[Z > 0], stored uint8[N][TOTVOX] in mask
    rank order as bsglmm/read_data.cpp:385 stores it.

The generator is numpy's Philox bit generator keyed by (0x42534731, config).
"""
import numpy as np

try:  # scipy is part of the image; keep the import local to what needs it
    from scipy.ndimage import gaussian_filter
except ImportError:  # pragma: no cover
    gaussian_filter = None

FWHM_TO_SIGMA = 1.0 / (2.0 * np.sqrt(2.0 * np.log(2.0)))

CONFIGS = {
    # name: grid (x, y, z), mask, N, P, fwhm, sd per covariate
    "c1": dict(config_id=1, grid=(32, 32, 32), mask=("ellipsoid", (15.5, 15.5, 15.5), (14, 12, 10)),
               nsubs=50, ncovar=3, fwhm=4.0, sd=[0.5, 0.5, 0.5]),
    "c2": dict(config_id=2, grid=(91, 109, 91),
               mask=("shell", (45, 54, 45), (37, 47, 35), (17, 21, 16), (22, 27, 21)),
               nsubs=250, ncovar=4, fwhm=8.0, sd=[0.5, 0.3, 0.3, 0.3]),
    "c3": dict(config_id=3, grid=(91, 109, 91),
               mask=("shell", (45, 54, 45), (37, 47, 35), (17, 21, 16), (22, 27, 21)),
               nsubs=1000, ncovar=8, fwhm=8.0, sd=[0.5, 0.3, 0.3, 0.3, 0.2, 0.2, 0.2, 0.2]),
}

# Uneven batch of 12 analyses: ellipsoid masks on the full grid whose
# semi-axes run linearly from (16, 20, 15) to (42, 54, 42), with subject
# counts and covariate counts that differ per analysis (P_k = 2 + k mod 7).
C5_NSUBS = [40, 80, 120, 250, 400, 600, 800, 1000, 60, 150, 300, 500]
for _k in range(12):
    _f = _k / 11.0
    _axes = tuple(a + _f * (b - a) for a, b in zip((16, 20, 15), (42, 54, 42)))
    CONFIGS[f"c5_{_k}"] = dict(config_id=50 + _k, grid=(91, 109, 91), mask=("ellipsoid", (45, 54, 45), _axes),
                               nsubs=C5_NSUBS[_k], ncovar=2 + _k % 7, fwhm=8.0, sd=[0.5, 0.3, 0.3, 0.3])


def _ellipsoid(grid, centre, axes):
    nx, ny, nz = grid
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ((x - centre[0]) / axes[0]) ** 2 + ((y - centre[1]) / axes[1]) ** 2 + \
        ((z - centre[2]) / axes[2]) ** 2


def make_mask(grid, spec):
    """Boolean mask of shape (z, y, x), plus an optional sub-region."""
    if spec[0] == "ellipsoid":
        return _ellipsoid(grid, spec[1], spec[2]) <= 1.0, None
    _, centre, outer, inner, sub = spec
    m = (_ellipsoid(grid, centre, outer) <= 1.0) & (_ellipsoid(grid, centre, inner) > 1.0)
    return m, (_ellipsoid(grid, centre, sub) <= 1.0)


def pad_mask(mask3d):
    """Padded uint8 grid as read_nifti1_mask builds it: one zero voxel on every side."""
    return np.ascontiguousarray(np.pad(mask3d.astype(np.uint8), 1))


def design(rng, nsubs, ncovar):
    """X as float32[P][N]: intercept, age, sex, subtype (p=0.4), then N(0,1) extras."""
    cols = [np.ones(nsubs)]
    age = rng.standard_normal(nsubs)
    cols.append((age - age.mean()) / age.std())
    cols.append((rng.random(nsubs) < 0.5).astype(float))
    cols.append((rng.random(nsubs) < 0.4).astype(float))
    while len(cols) < ncovar:
        cols.append(rng.standard_normal(nsubs))
    return np.ascontiguousarray(np.array(cols[:ncovar], dtype=np.float32))


def generate(name="c1", nsubs=None, ncovar=None, chunk=32):
    """Data set of a named configuration as a dict of numpy arrays.

    Keys: grid (nrow, ncol, ndepth), mask (padded uint8), mask3d (bool, z y x),
    covar float32[P][N], data uint8[N][TOTVOX], beta_true float32[P][TOTVOX].
    """
    cfg = dict(CONFIGS[name])
    n = nsubs or cfg["nsubs"]
    p = ncovar or cfg["ncovar"]
    rng = np.random.Generator(np.random.Philox(key=(0x42534731 << 32) | cfg["config_id"]))
    mask3d, sub = make_mask(cfg["grid"], cfg["mask"])
    V = int(mask3d.sum())
    X = design(rng, n, p)
    sds = list(cfg["sd"]) + [0.2] * max(0, p - len(cfg["sd"]))
    sigma = cfg["fwhm"] * FWHM_TO_SIGMA
    beta = np.empty((p, V), dtype=np.float64)
    for c in range(p):
        field = rng.standard_normal(mask3d.shape)
        if gaussian_filter is not None:
            field = gaussian_filter(field, sigma, mode="nearest")
        vals = field[mask3d]
        vals = (vals - vals.mean()) / vals.std()
        beta[c] = vals * sds[c]
    beta[0] += -2.0
    if sub is not None:
        beta[0] += np.where(sub[mask3d], 1.0, 0.0)
    data = np.empty((n, V), dtype=np.uint8)
    Xd = X.astype(np.float64)
    for i0 in range(0, n, chunk):
        i1 = min(n, i0 + chunk)
        z = Xd[:, i0:i1].T @ beta + rng.standard_normal((i1 - i0, V))
        data[i0:i1] = z > 0.0
    return dict(name=name, grid=tuple(cfg["grid"]), mask=pad_mask(mask3d), mask3d=mask3d,
                covar=X, data=data, beta_true=beta.astype(np.float32))
