"""Derive the polynomial coefficients used by trax_math.h.

The device and the host oracle share one implementation of the four
transcendental built-ins the triplet-seeding kernels call (atan2, acos, atan,
asinh; log is an internal helper).  Each one reduces its argument and then
evaluates  r = s + s*t*P(t)  with t = s*s, so only the correction polynomial P
has to be approximated.  P is fitted here as a weighted minimax polynomial
(Lawson's iteratively re-weighted least squares on a dense Chebyshev grid),
with the weight chosen so that the error that is minimised is the *relative*
error of the final result.  The coefficients are rounded to float32 and the
whole float32 evaluation (plain multiplies and adds, no FMA, exactly as the
C code does it) is then checked against float64 numpy; the achieved max ulp
error is printed.  Re-run:  python trax_amd/csrc/fit_math_coeffs.py
"""
import numpy as np


def lawson_fit(g, w, lo, hi, deg, iters=400, npts=6000):
    x = 0.5 * (lo + hi) + 0.5 * (hi - lo) * np.cos(np.linspace(np.pi, 0.0, npts))
    A = np.vander(x, deg + 1, increasing=True) * w(x)[:, None]
    b = g(x) * w(x)
    lw = np.full(npts, 1.0 / npts)
    c = None
    for _ in range(iters):
        sw = np.sqrt(lw)
        c, *_ = np.linalg.lstsq(A * sw[:, None], b * sw, rcond=None)
        err = np.abs(A @ c - b)
        lw = lw * (err + 1e-300)
        lw /= lw.sum()
    return c


def horner32(c, t):
    acc = np.full_like(t, np.float32(c[-1]))
    for k in range(len(c) - 2, -1, -1):
        acc = (acc * t).astype(np.float32) + np.float32(c[k])
        acc = acc.astype(np.float32)
    return acc


def ulp_err(approx32, exact64):
    approx32 = approx32.astype(np.float32)
    ref32 = exact64.astype(np.float32)
    ulp = np.spacing(np.abs(ref32)).astype(np.float64)
    return np.max(np.abs(approx32.astype(np.float64) - exact64) / ulp)


def core(s, c):
    """s + s*t*P(t), float32, plain ops."""
    s = s.astype(np.float32)
    t = (s * s).astype(np.float32)
    p = horner32(c, t)
    return (s + ((s * t).astype(np.float32) * p).astype(np.float32)).astype(np.float32)


def fit_atan(deg):
    g = lambda z: (np.arctan(np.sqrt(z)) / np.sqrt(z) - 1.0) / z
    w = lambda z: z / (np.arctan(np.sqrt(z)) / np.sqrt(z))
    c = lawson_fit(g, w, 1e-12, 1.0, deg).astype(np.float32)
    a = np.linspace(0.0, 1.0, 2_000_001, dtype=np.float32)[1:]
    return c, ulp_err(core(a, c), np.arctan(a.astype(np.float64)))


def fit_asin(deg):
    g = lambda z: (np.arcsin(np.sqrt(z)) / np.sqrt(z) - 1.0) / z
    w = lambda z: z / (np.arcsin(np.sqrt(z)) / np.sqrt(z))
    c = lawson_fit(g, w, 1e-12, 0.25, deg).astype(np.float32)
    a = np.linspace(0.0, 0.5, 2_000_001, dtype=np.float32)[1:]
    return c, ulp_err(core(a, c), np.arcsin(a.astype(np.float64)))


def fit_atanh(deg):
    smax = (np.sqrt(2.0) - 1.0) / (np.sqrt(2.0) + 1.0)
    g = lambda z: (np.arctanh(np.sqrt(z)) / np.sqrt(z) - 1.0) / z
    w = lambda z: z / (np.arctanh(np.sqrt(z)) / np.sqrt(z))
    c = lawson_fit(g, w, 1e-14, smax * smax * 1.0001, deg).astype(np.float32)
    a = np.linspace(0.0, smax, 2_000_001, dtype=np.float32)[1:]
    return c, ulp_err(core(a, c), np.arctanh(a.astype(np.float64)))


def fmt(name, c):
    body = ",\n    ".join(f"{float(v):.9e}f" for v in c)
    return f"// {name}: P(t) coefficients, constant term first\n{name} = {{\n    {body}\n}};"


if __name__ == "__main__":
    for fn, name, degs in ((fit_atan, "atan", range(6, 10)),
                           (fit_asin, "asin", range(3, 7)),
                           (fit_atanh, "atanh", range(2, 5))):
        for d in degs:
            c, e = fn(d)
            print(f"{name} deg {d}: max err {e:.3f} ulp (core only)")
            print(fmt(f"k_{name}", c))
