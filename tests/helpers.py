"""Shared helpers for the parity tests (seeded inputs, error norms)."""

import numpy as np

from oracle import numpy_path as npth


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    if nb == 0:
        return float(np.linalg.norm(a.ravel()))
    return float(np.linalg.norm((a - b).ravel()) / nb)


def random_x(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


def random_k(shape, seed, length=None, slope=0.0):
    """Dealiased, Hermitian k array (reference layout) of a random real field.

    slope > 0 damps high wavenumbers as |k|^-slope so products stay smooth.
    """
    length = length or (2 * np.pi,) * len(shape)
    c = npth.Component(shape, length)
    c['xspace'] = random_x(shape, seed)
    k = c['kspace'].copy()
    if slope:
        kk = np.sqrt(c.k2(no_zero=True))
        k *= kk ** (-slope)
    return k


def kfields(kind, shape, seed, length=None, slope=1.5, amp=1.0):
    nd = len(shape)
    out = {'u': [amp * random_k(shape, seed + i, length, slope) for i in range(nd)]}
    if kind == 'boussinesq':
        out['T'] = [amp * random_k(shape, seed + 10, length, slope)]
    if kind == 'mhd':
        out['B'] = [amp * random_k(shape, seed + 20 + i, length, slope) for i in range(nd)]
    return out
