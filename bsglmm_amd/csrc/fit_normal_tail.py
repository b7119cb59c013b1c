"""Fits the FP32 normal-tail approximations used by k_latent (bsg_kernels.hip).

  Q(t)    = P(N(0,1) > t), t >= 0:
            Q(t) = r exp(-t^2/2 + g(r)),  r = 1/(1 + t/(2 sqrt 2)),
            g a degree-7 polynomial, least squares on Chebyshev nodes in r over
            [0.2, 1] (t <= 11.3). The kernel needs Q to FP32 accuracy only for
            t < 8 (past that its own tail path takes over, or Q only has to be
            far below 2^-24); below r = 0.2 g stays within [-1.96, -1.74], so
            Q keeps underflowing as it should.
  Qinv(p) for p in (1e-38, 1/2]:
            s = sqrt(-2 ln p),  Qinv(p) = n(s)/d(s), n of degree 4, d of degree 3
            (d(0) = 1), fitted by Lawson-reweighted nonlinear least squares on
            the error (approx - exact) / max(1, exact).

Run: python bsglmm_amd/csrc/fit_normal_tail.py   (prints the coefficient lists,
then the same constants rescaled as the kernel uses them)
Max errors: the fitted functions 6.3e-7 (Q, relative, t <= 8) and 1.4e-7 (Qinv,
relative with floor 1); in FP32 evaluation Q 6e-6 relative (from rounding
t^2/2 near t = 8, the same at degrees 7 to 9), Qinv 6e-7. Degrees 9 and 4/4
gave the same FP32 errors with three more FMAs per draw.
"""
import numpy as np
import scipy.optimize as so
import scipy.special as sp

C_Q = 1.0 / (2.0 * np.sqrt(2.0))


def fit_q(deg=7, rmin=0.2, nodes=6000):
    r = rmin + (1 - rmin) * 0.5 * (1 + np.cos(np.pi * (np.arange(nodes) + 0.5) / nodes))
    t = (1 / r - 1) / C_Q
    g = sp.log_ndtr(-t) - np.log(r) + t * t / 2
    return np.polynomial.polynomial.polyfit(r, g, deg)


def _rational(c, s, m):
    num = np.polynomial.polynomial.polyval(s, c[:m + 1])
    den = np.polynomial.polynomial.polyval(s, np.concatenate([[1.0], c[m + 1:]]))
    return num / den


def fit_qinv(m=4, n=3, pmin=1e-38, npts=20000):
    s = np.linspace(np.sqrt(-2 * np.log(0.5)), np.sqrt(-2 * np.log(pmin)), npts)
    x = -sp.ndtri_exp(-0.5 * s * s)
    sc = np.maximum(1.0, x)
    A = np.concatenate([np.vstack([s ** k for k in range(m + 1)]).T,
                        -np.vstack([x * s ** k for k in range(1, n + 1)]).T], axis=1)
    c, *_ = np.linalg.lstsq(A / sc[:, None], x / sc, rcond=None)

    def err(cc, w):
        return np.sqrt(w) * (_rational(cc, s, m) - x) / sc

    w = np.ones_like(s)
    c = so.least_squares(err, c, args=(w,), xtol=1e-15, ftol=1e-15, gtol=1e-15, max_nfev=20000).x
    for _ in range(30):
        r = np.abs(_rational(c, s, m) - x) / sc
        w = w * np.sqrt(r)
        w = w / w.max() + 1e-3
        c = so.least_squares(err, c, args=(w,), xtol=1e-15, ftol=1e-15, gtol=1e-15, max_nfev=2000).x
    return c


def kernel_constants():
    """The constants as the kernel uses them: g scaled by log2(e) (the kernel
    evaluates 2^(g log2 e - t^2 log2(e)/2) with v_exp_f32), and the rational in
    s2 = sqrt(-log2 p) = s / sqrt(2 ln 2) (v_log_f32 is log2)."""
    q = fit_q() * np.log2(np.e)
    qi = fit_qinv()
    k = np.sqrt(2.0 * np.log(2.0))
    num = qi[:5] * k ** np.arange(5)
    den = np.concatenate([[1.0], qi[5:]]) * k ** np.arange(len(qi) - 4)
    return q, num, den


if __name__ == "__main__":
    q = fit_q()
    qi = fit_qinv()
    print("Q coefficients (r^0 .. r^7):", ", ".join(f"{float(np.float32(v))!r}f" for v in q))
    print("Qinv numerator (s^0 .. s^4):", ", ".join(f"{float(np.float32(v))!r}f" for v in qi[:5]))
    print("Qinv denominator (s^1 .. s^3):", ", ".join(f"{float(np.float32(v))!r}f" for v in qi[5:]))
    gq, num, den = kernel_constants()
    print("kernel: g*log2e (r^0 .. r^7):", ", ".join(f"{float(np.float32(v))!r}f" for v in gq))
    print("kernel: numerator in sqrt(-log2 p) (s^0 .. s^4):", ", ".join(f"{float(np.float32(v))!r}f" for v in num))
    print("kernel: denominator (s^0 .. s^3):", ", ".join(f"{float(np.float32(v))!r}f" for v in den))
