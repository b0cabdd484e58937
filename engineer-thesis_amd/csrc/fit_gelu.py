"""Derives the coefficients of the device GELU (frw_mlp.h, gelu_f):

    GELU(z) = max(z, 0) - t * P(x),  t = min(|z|, 4.5),  x = 1 - t / 4.5,
    P(x) = c1 x + c2 x^2 + ... + c9 x^9  ~  Phi(-t)

(Phi the standard normal CDF, so z Phi(z) = max(z, 0) - |z| Phi(-|z|)).
P vanishes at t = 4.5, where |z| Phi(-|z|) < 1.6e-5, so no clamp of the
result is needed beyond it.  The c_k minimise max over t in [0, 4.5] of
max(t, 1) |P(x) - Phi(-t)| (a linear programme).  The device evaluates
P(x) / x by Horner with FMA, then fma(-(t x), P(x) / x, max(z, 0)); in FP32
that is within 1.7e-5 of the exact GELU everywhere (degree 10 is no closer:
the FP32 evaluation's rounding sets the floor; degree 8 is ten times
further).
Run: python3 fit_gelu.py"""
import numpy as np
from scipy.optimize import linprog
from scipy.special import erf, erfc

T, DEG = 4.5, 9


def fit():
    x = np.linspace(0.0, 1.0, 6000)
    t = (1.0 - x) * T
    V = np.vander(x, DEG + 1, increasing=True)[:, 1:]
    w = np.maximum(t, 1.0)
    f = 0.5 * erfc(t / np.sqrt(2.0))
    A = np.vstack([np.hstack([w[:, None] * V, -np.ones((len(x), 1))]),
                   np.hstack([-w[:, None] * V, -np.ones((len(x), 1))])])
    b = np.concatenate([w * f, -w * f])
    res = linprog(np.r_[np.zeros(DEG), 1.0], A_ub=A, b_ub=b, bounds=[(None, None)] * DEG + [(0, None)],
                  method="highs")
    return res.x[:DEG]


def fma32(a, b, c):
    """FP32 fused multiply-add: the exact a b + c (FP64 holds the product of
    two FP32 values exactly) rounded once to FP32."""
    return (np.asarray(a, np.float64) * np.asarray(b, np.float64) + np.asarray(c, np.float64)).astype(np.float32)


def device_form(z, c):
    z = np.asarray(z, dtype=np.float32)
    t = np.minimum(np.abs(z), np.float32(T))
    x = fma32(t, np.float32(-1.0 / T), np.float32(1.0))
    q = (t * x).astype(np.float32)
    p = np.float32(c[-1])
    for k in range(DEG - 2, -1, -1):
        p = fma32(p, x, np.float32(c[k]))
    return fma32(-q, p, np.maximum(z, np.float32(0.0)))


if __name__ == "__main__":
    c = fit()
    z = np.linspace(-12.0, 12.0, 400001)
    err = np.abs(device_form(z, c).astype(np.float64) - z * 0.5 * (1.0 + erf(z / np.sqrt(2.0))))
    print("coefficients c1..c9:", ", ".join(f"{v:.9e}f" for v in c))
    print(f"max |error| {err.max():.2e}")
