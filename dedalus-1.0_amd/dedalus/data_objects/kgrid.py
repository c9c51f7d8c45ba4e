"""Host-side wavenumber grids in the reference k layout.

Follows FourierRepresentation._setup_k (dedalus/data_objects/representations.py:204-233):
k-space is stored transposed -- (kx, ky) in 2D, (ky, kz, kx) in 3D -- with the
real-to-complex half axis on kx; k_i = fftfreq(n)*2*kny_i with kny = pi N / L and
the Nyquist entry of an even axis stored as +kny.  The 2/3 rule zeroes modes with
|k_i| >= (2/3) kny_i on any axis (representations.py:384-397).
"""

import numpy as np


def swap01(seq):
    """Swap entries 0 and 1 (dedalus/utils/parallelism.py:155-173)."""
    out = list(seq)
    out[0], out[1] = out[1], out[0]
    return out


def ktrans(ndim):
    if ndim == 2:
        return {'x': 0, 0: 'x', 'y': 1, 1: 'y'}
    return {'x': 2, 2: 'x', 'y': 0, 0: 'y', 'z': 1, 1: 'z'}


def kshape(xshape):
    ks = list(int(n) for n in xshape)
    ks[-1] = ks[-1] // 2 + 1
    return tuple(swap01(ks))


def wavenumbers(xshape, length):
    """(k dict of broadcastable arrays, kny array in k order, k shape)."""
    xshape = tuple(int(n) for n in xshape)
    length = np.asarray(length, dtype=np.float64)
    ndim = len(xshape)
    kt = ktrans(ndim)
    ksh = kshape(xshape)
    kny = np.array(swap01(np.pi * np.array(xshape, dtype=np.float64) / length))
    xs = swap01(xshape)
    rcomp = kt['x']
    k = {}
    for i, ksize in enumerate(ksh):
        n = xs[i]
        if i == rcomp:
            ki = np.fft.fftfreq(n)[:ksize] * 2. * kny[i]
            if n % 2 == 0:
                ki[-1] *= -1.
        else:
            ki = np.fft.fftfreq(ksize) * 2. * kny[i]
            if n % 2 == 0:
                ki[ksize // 2] *= -1.
        shp = [1] * ndim
        shp[i] = ksize
        k[kt[i]] = ki.reshape(shp)
    return k, kny, ksh


def dealias_mask(xshape, length, rule='2/3'):
    """True where a dealiasing rule removes a mode (reference k layout):
    '2/3' (representations.py:384-397), '2/3 spherical' (:410-416: |k| >=
    (2/3) min(kny)) or 'None' (zero_nyquist, :442-455: the Nyquist planes)."""
    k, kny, ksh = wavenumbers(xshape, length)
    kt = ktrans(len(xshape))
    names = ('x', 'y', 'z')[:len(xshape)]
    if rule == '2/3 spherical':
        k2 = np.zeros(ksh)
        for kk in k.values():
            k2 = k2 + kk ** 2
        return np.sqrt(k2) >= 2 / 3. * np.min(kny)
    mask = np.zeros(ksh, dtype=bool)
    for name in names:
        if rule == 'None':
            mask = mask | (np.abs(k[name]) == kny[kt[name]])
        else:
            mask = mask | (np.abs(k[name]) >= 2. / 3. * kny[kt[name]])
    return mask


def k2(xshape, length):
    k, _, ksh = wavenumbers(xshape, length)
    out = np.zeros(ksh)
    for kk in k.values():
        out = out + kk ** 2
    return out
