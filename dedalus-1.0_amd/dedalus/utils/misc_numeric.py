"""Small numerical helpers (dedalus/utils/misc_numeric.py): quadrature,
interpolation and root finding on host arrays."""

import numpy as np


def integrate_quad(f, x0, x1, **kw):
    from scipy.integrate import quad
    return quad(f, x0, x1, **kw)[0]


def integrate_simp(y, x):
    from scipy.integrate import simpson
    return simpson(y, x=x)


def interp_linear(x, xp, fp):
    return np.interp(x, xp, fp)


def find_zero(f, a, b, tol=1e-12, maxiter=200):
    """Bisection root of f on [a, b]."""
    fa = f(a)
    for _ in range(maxiter):
        m = 0.5 * (a + b)
        fm = f(m)
        if fa * fm <= 0:
            b = m
        else:
            a, fa = m, fm
        if abs(b - a) < tol:
            break
    return 0.5 * (a + b)
