"""Seeded initial conditions of the benchmark workloads, built in k space.

Every generator returns k-space arrays in the reference layout ((kx, ky) in
2D, (ky, kz, kx) in 3D; see dedalus/data_objects/representations.py:82-93),
dealiased by the 2/3 rule and Hermitian on the kx = 0 plane, so the same
arrays can be fed to any implementation of the step.  Nothing here
transforms: random phases are drawn directly per mode, which is the
Rogallo-style construction of init_cond.py:281-341 without the x-space
detour.

Workloads
---------
taylor_green_64      2D hydro 64^2, u = sin x cos y, v = -cos x sin y, nu = 1e-3,
                     RK2mid, dt = 1e-3, 200 steps
hydro3d_128          3D hydro 128^3, E(k) ~ k^-5/3 on 1 <= |k| <= 8, <|u|^2> = 1,
                     seed 1, nu = 2e-3, dt = 5e-4, 100 steps
mhd3d_256            3D MHD 256^3, E(k) ~ k^4 exp(-2 (k/4)^2), <|u|^2> = <|B|^2> = 1,
                     seeds 2 (u) and 3 (B), nu = eta = 1e-3, dt = 2.5e-4, 50 steps
boussinesq2d_512     2D Boussinesq 512^2, u = 0, T white noise with rms 1e-3 after
                     dealiasing (seed 4), g = alpha_t = beta = 1 (N = 1),
                     direction y, nu = kappa = 1e-3, dt = 2.5e-4, 50 steps
"""

import numpy as np

from dedalus.data_objects.kgrid import dealias_mask, ktrans, wavenumbers


def _kvec(shape, length):
    k, kny, ksh = wavenumbers(shape, length)
    nd = len(shape)
    comps = [np.broadcast_to(k[n], ksh) for n in ('x', 'y', 'z')[:nd]]
    return comps, ksh


def hermitian_kx0(arr, shape):
    """Make the kx = 0 plane Hermitian in place: the mirror of (fy[, fz]) is set
    to the conjugate of the entry with the canonical (positive) index, and the
    mean mode is zeroed."""
    nd = len(shape)
    if nd == 2:
        ny = shape[0]
        line = arr[0, :]
        for fy in range(1, ny):
            my = (-fy) % ny
            if fy < my:
                line[my] = np.conj(line[fy])
            elif fy == my:
                line[fy] = line[fy].real
        line[0] = 0.
    else:
        nz, ny = shape[0], shape[1]
        plane = arr[:, :, 0]  # (ky, kz)
        fy = np.arange(ny)[:, None]
        fz = np.arange(nz)[None, :]
        my = (-fy) % ny
        mz = (-fz) % nz
        lin = fy * nz + fz
        mlin = my * nz + mz
        canon = np.broadcast_to(lin < mlin, plane.shape)
        selfm = np.broadcast_to(lin == mlin, plane.shape)
        src = plane.copy()
        mirrored = np.conj(src[my % ny, mz % nz])
        plane[~canon & ~selfm] = mirrored[~canon & ~selfm]
        plane[selfm] = plane[selfm].real
        plane[0, 0] = 0.
    return arr


def mean_square(fields, shape):
    """<|f|^2> over x space from reference-layout k arrays (Parseval with
    Hermitian weights: 1 on the kx = 0 plane, 2 elsewhere)."""
    nd = len(shape)
    tot = 0.
    for a in fields:
        p = np.abs(a) ** 2
        if nd == 2:
            tot += p[0, :].sum() + 2 * p[1:, :].sum()
        else:
            tot += p[..., 0].sum() + 2 * p[..., 1:].sum()
    return float(tot)


def random_solenoidal(shape, length, seed, spectrum, rms=1.0):
    """Random-phase solenoidal vector field with shell spectrum E(|k|).

    Per mode: a complex Gaussian vector is projected perpendicular to k,
    normalised, and scaled to |u_k| = sqrt(E(|k|) / (4 pi |k|^2)) in 3D or
    sqrt(E(|k|) / (2 pi |k|)) in 2D (the shell densities of turb_new,
    init_cond.py:295-299); then Hermitian on kx = 0, dealiased and scaled
    so that <|u|^2> = rms^2.
    """
    nd = len(shape)
    kv, ksh = _kvec(shape, length)
    kk2 = sum(c ** 2 for c in kv)
    kk = np.sqrt(kk2)
    rng = np.random.default_rng(seed)
    xi = [rng.standard_normal(ksh) + 1j * rng.standard_normal(ksh) for _ in range(nd)]
    safe = np.where(kk2 == 0, 1., kk2)
    kdot = sum(c * x for c, x in zip(kv, xi))
    xi = [x - c * kdot / safe for c, x in zip(kv, xi)]
    nrm = np.sqrt(sum(np.abs(x) ** 2 for x in xi))
    nrm = np.where(nrm == 0, 1., nrm)
    with np.errstate(divide='ignore', invalid='ignore'):
        E = np.where(kk > 0, spectrum(np.where(kk > 0, kk, 1.)), 0.)
        dens = (4 * np.pi * kk2) if nd == 3 else (2 * np.pi * kk)
        amp = np.where(kk > 0, np.sqrt(np.abs(E) / np.where(dens == 0, 1., dens)), 0.)
    mask = dealias_mask(shape, length)
    out = []
    for x in xi:
        a = amp * x / nrm
        a[mask] = 0.
        hermitian_kx0(a, shape)
        out.append(a)
    # the Hermitian fix keeps k.u = 0 (the mirror of a solenoidal mode is solenoidal)
    ms = mean_square(out, shape)
    s = rms / np.sqrt(ms) if ms > 0 else 0.
    return [a * s for a in out]


def scalar_noise(shape, length, seed, rms):
    """Dealiased white noise with the given rms, mean zero."""
    ksh = wavenumbers(shape, length)[2]
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(ksh) + 1j * rng.standard_normal(ksh)
    a[dealias_mask(shape, length)] = 0.
    hermitian_kx0(a, shape)
    ms = mean_square([a], shape)
    return a * (rms / np.sqrt(ms))


def taylor_green_2d(shape=(64, 64)):
    """u = sin x cos y, v = -cos x sin y on [0, 2pi)^2: the kx = +1 coefficients
    u(1, +-1) = -i/4, v(1, 1) = +i/4, v(1, -1) = -i/4 (reference layout (kx, ky))."""
    ksh = wavenumbers(shape, (2 * np.pi, 2 * np.pi))[2]
    u = np.zeros(ksh, dtype=np.complex128)
    v = np.zeros(ksh, dtype=np.complex128)
    u[1, 1] = -0.25j
    u[1, -1] = -0.25j
    v[1, 1] = 0.25j
    v[1, -1] = -0.25j
    return [u, v]


# ---------------------------------------------------------------------------
# workload descriptions
# ---------------------------------------------------------------------------

def _box(nd):
    return (2 * np.pi,) * nd


def workload(name, shape=None):
    """Dict with physics kind, shape, length, params, dt, nsteps and k fields.

    ``shape`` overrides the grid (same recipe, different resolution) for
    parity tests at sizes the oracle finishes quickly.
    """
    if name == 'taylor_green_64':
        shape = shape or (64, 64)
        return dict(name=name, kind='hydro', shape=shape, length=_box(2),
                    params=dict(nu=1e-3), dt=1e-3, nsteps=200, integrator='RK2mid',
                    fields={'u': taylor_green_2d(shape)})
    if name == 'hydro3d_128':
        shape = shape or (128, 128, 128)
        L = _box(3)
        spec = lambda k: np.where((k >= 1) & (k <= 8), k ** (-5. / 3.), 0.)
        return dict(name=name, kind='hydro', shape=shape, length=L, params=dict(nu=2e-3),
                    dt=5e-4, nsteps=100, integrator='RK2mid',
                    fields={'u': random_solenoidal(shape, L, 1, spec)})
    if name == 'mhd3d_256':
        shape = shape or (256, 256, 256)
        L = _box(3)
        spec = lambda k: k ** 4 * np.exp(-2 * (k / 4.) ** 2)
        return dict(name=name, kind='mhd', shape=shape, length=L,
                    params=dict(nu=1e-3, eta=1e-3, rho0=1.0), dt=2.5e-4, nsteps=50,
                    integrator='RK2mid',
                    fields={'u': random_solenoidal(shape, L, 2, spec),
                            'B': random_solenoidal(shape, L, 3, spec)})
    if name == 'boussinesq2d_512':
        shape = shape or (512, 512)
        L = _box(2)
        ksh = wavenumbers(shape, L)[2]
        zero = np.zeros(ksh, dtype=np.complex128)
        return dict(name=name, kind='boussinesq', shape=shape, length=L,
                    params=dict(nu=1e-3, kappa=1e-3, g=1., alpha_t=1., beta=1.,
                                boussinesq_direction='y'),
                    dt=2.5e-4, nsteps=50, integrator='RK2mid',
                    fields={'u': [zero, zero.copy()], 'T': [scalar_noise(shape, L, 4, 1e-3)]})
    raise KeyError(name)


WORKLOADS = ('taylor_green_64', 'hydro3d_128', 'mhd3d_256', 'boussinesq2d_512')
