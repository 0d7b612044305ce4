"""Derive the polynomial coefficients used by pt_detmath.h.

atan(u) = u + u^3 * P(u^2) on |u| <= tan(pi/8), and
asin(z) = z + z^3 * Q(z^2) on 0 <= z <= 1/2,
fitted by weighted least squares (relative error) on Chebyshev nodes in
float64, then rounded to float32. Run: python gen_detmath_coeffs.py
"""
import numpy as np


def fit(func, lo, hi, degree, nodes=4000):
    k = np.arange(nodes)
    x = 0.5 * (lo + hi) + 0.5 * (hi - lo) * np.cos(np.pi * (k + 0.5) / nodes)
    x = x[x > 1e-6]
    t = x * x
    target = (func(x) - x) / (x ** 3)
    # weight by x^3 / func(x) so the fit minimises relative error of func
    w = x ** 3 / func(x)
    a = np.vstack([t ** i for i in range(degree + 1)]).T
    coef, *_ = np.linalg.lstsq(a * w[:, None], target * w, rcond=None)
    return coef


def check(func, coef, lo, hi):
    c32 = coef.astype(np.float32)
    x = np.linspace(max(lo, 1e-4), hi, 200001).astype(np.float32)
    xd = x.astype(np.float64)
    t = xd * xd
    p = np.zeros_like(t)
    for c in c32[::-1].astype(np.float64):
        p = p * t + c
    approx = xd + xd ** 3 * p
    exact = func(xd)
    return np.max(np.abs(approx - exact) / np.abs(exact))


if __name__ == "__main__":
    tp8 = np.tan(np.pi / 8)
    for name, func, lo, hi, deg in (("ATAN", np.arctan, 0.0, tp8, 7),
                                    ("ASIN", np.arcsin, 0.0, 0.5, 7)):
        coef = fit(func, lo, hi, deg)
        print(name, "max rel err (float32 coeffs, float64 eval):", check(func, coef, lo, hi))
        for i, c in enumerate(coef.astype(np.float32)):
            print(f"  {name}_C{i} = {float(c)!r}f  ({np.float32(c).view(np.uint32):#010x})")
