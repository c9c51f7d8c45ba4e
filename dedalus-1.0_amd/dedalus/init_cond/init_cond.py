"""Initial-condition helpers with the reference's call signatures.

Follow dedalus/init_cond/init_cond.py (taylor_green 33-85, sin_k / cos_k
136-142, alfven 145-243, turb_new 281-341, MIT_vortices 383-399,
vorticity_wave 401-438, add_gaussian_white_noise 440-464, constant
466-490) and turb_spectra.py (mcwilliams_spec).  They are setup code: they
write k-space entries through the representation interface and use the
device transforms where the reference transforms.  The seeded generators
of the benchmark workloads are in ``dedalus.init_cond.seeded``.
"""

import numpy as np

from dedalus.utils.logger import mylog


def taylor_green(data):
    """Taylor-Green vortex entries as the reference writes them (init_cond.py:33-85).

    In 2D the reference writes rows 1 and -1 of the (kx, ky) array; row -1
    is the kx Nyquist row, which the first dealias clears, leaving
    u = cos x sin y, v = -sin x cos y.  ``seeded.taylor_green_2d`` builds
    u = sin x cos y, v = -cos x sin y instead.
    """
    mylog.info("Initializing Taylor Green Vortex.")
    ndim = data['u'].ndim
    if ndim == 2:
        ux = data['u']['x'].data
        ux[1, 1] = -1j / 4.
        ux[-1, -1] = 1j / 4.
        ux[-1, 1] = -1j / 4.
        ux[1, -1] = 1j / 4.
        uy = data['u']['y'].data
        uy[1, 1] = 1j / 4.
        uy[-1, -1] = -1j / 4.
        uy[-1, 1] = -1j / 4.
        uy[1, -1] = 1j / 4.
    else:
        signs = {'x': [(1, 1, 1, -1), (1, 1, -1, 1), (-1, 1, 1, -1), (-1, 1, -1, 1),
                       (1, -1, 1, -1), (1, -1, -1, 1), (-1, -1, 1, -1), (-1, -1, -1, 1)],
                 'y': [(1, 1, 1, 1), (1, 1, -1, 1), (-1, 1, 1, -1), (-1, 1, -1, -1),
                       (1, -1, 1, 1), (1, -1, -1, 1), (-1, -1, 1, -1), (-1, -1, -1, -1)]}
        for comp, entries in signs.items():
            rep = data['u'][comp]
            for a, b, c, s in entries:
                if rep.find_mode((a, b, c)):
                    rep['kspace'][a, b, c] = s * 1j / 8.


def sin_k(f, kindex, ampl=1.):
    """Sine mode at kindex (init_cond.py:136-138)."""
    f[tuple(kindex)] = ampl * 1j / 2.
    f[tuple(-1 * np.array(kindex))] = f[tuple(kindex)].conjugate()


def cos_k(f, kindex, ampl=1.):
    """Cosine mode at kindex (init_cond.py:140-142)."""
    f[tuple(kindex)] = ampl / 2.
    f[tuple(-1 * np.array(kindex))] = f[tuple(kindex)].conjugate()


def _psi_to_u(data, aux):
    """u = d_y psi, v = -d_x psi with psi = w / k^2 (init_cond.py:129-133)."""
    aux['psi']['kspace'] = aux['w']['kspace'] / aux['w'].k2(no_zero=True)
    data['u']['x']['kspace'] = aux['psi'].deriv('y')
    data['u']['y']['kspace'] = -aux['psi'].deriv('x')


def MIT_vortices(data):
    """Three Gaussian vortices of the MIT 18.336 example (init_cond.py:383-399)."""
    y, x = data['u']['x'].xspace_grid()
    aux = data.clone()
    aux.add_field('w', 'ScalarField')
    aux.add_field('psi', 'ScalarField')
    aux['w']['xspace'] = (np.exp(-((x - np.pi) ** 2 + (y - np.pi + np.pi / 4) ** 2) / 0.2)
                          + np.exp(-((x - np.pi) ** 2 + (y - np.pi - np.pi / 4) ** 2) / 0.2)
                          - 0.5 * np.exp(-((x - np.pi - np.pi / 4) ** 2
                                           + (y - np.pi - np.pi / 4) ** 2) / 0.4))
    _psi_to_u(data, aux)


def vorticity_wave(data, mode, w_amp):
    """Single vorticity wave of complex amplitude w_amp (init_cond.py:401-438)."""
    aux = data.clone()
    aux.add_field('w', 'ScalarField')
    aux.add_field('psi', 'ScalarField')
    i0 = data['u']['x'].find_mode(mode)
    if i0:
        aux['w']['kspace'][i0] = w_amp / 2.
    i1 = data['u']['x'].find_mode(-1 * np.array(mode))
    if i1:
        aux['w']['kspace'][i1] = np.conj(w_amp) / 2.
    if i0 or i1:
        _psi_to_u(data, aux)


def alfven(data, k=(1, 0, 0), B0mag=5.0, u1mag=5e-6, p_vec=(0., 1., 0.)):
    """Alfven wave on a uniform field along x, 3D only (init_cond.py:145-243)."""
    N = len(k)
    if N != 3:
        raise ValueError('Only setup for 3d')
    B0 = np.array([1., 0., 0.])[:N] * B0mag
    p_vec = np.array(p_vec)
    rep = data['u']['x']
    if 0 in rep.k['z'] and 0 in rep.k['y'] and 0 in rep.k['x']:
        for i in range(data['B'].ndim):
            data['B'][i]['kspace'][(0,) * N] = B0[i]
    cA = B0mag / np.sqrt(4 * np.pi * data.parameters['rho0'])
    omega = np.abs(cA * np.dot(k, B0) / B0mag)
    u1 = p_vec[3 - N:] * u1mag
    B1 = (np.dot(k, u1) * B0 - np.dot(k, B0) * u1) / omega
    for sgn in (1, -1):
        kk = [sgn * v for v in k]
        if kk[2] in rep.k['z'] and kk[1] in rep.k['y'] and kk[0] in rep.k['x']:
            kiz = np.array(np.where(rep.k['z'] == kk[2]))
            kiy = np.array(np.where(rep.k['y'] == kk[1]))
            kix = np.array(np.where(rep.k['x'] == kk[0]))
            kindex = tuple(kiz + kiy + kix)
            for i in range(data['u'].ndim):
                data['u'][i]['kspace']
                data['B'][i]['kspace']
                data['u'][i].data[kindex] = sgn * u1[i] * 1j / 2.
                data['B'][i].data[kindex] = sgn * B1[i] * 1j / 2.


def mcwilliams_spec(k, k0=6., E0=1.):
    """McWilliams (1990) energy spectrum shape (turb_spectra.py)."""
    spec = k ** 6. / (k + 2. * k0) ** 18.
    spec *= E0
    return spec


def turb_new(data, spec, tot_en=0.5, **kwargs):
    """Random-phase solenoidal field with shell spectrum spec (init_cond.py:281-341)."""
    from dedalus.data_objects.kgrid import dealias_mask
    from dedalus.init_cond.seeded import hermitian_kx0, mean_square
    rep = data['u'][0]
    kk = np.sqrt(rep.k2())
    sp = spec(kk, **kwargs)
    kk[kk == 0] = 1.
    nd = data.ndim
    amp = np.sqrt(2. * (sp / (2. * np.pi * kk) if nd == 2 else sp / (4 * np.pi * kk ** 2)))
    shape = tuple(int(n) for n in rep.global_shape['xspace'])
    mask = dealias_mask(shape, rep.length)
    k = [np.broadcast_to(rep.k[n], kk.shape) for n in ('x', 'y', 'z')[:nd]]
    xi = [np.random.standard_normal(kk.shape) + 1j * np.random.standard_normal(kk.shape)
          for _ in range(nd)]
    kdot = sum(a * b for a, b in zip(k, xi))
    k2 = rep.k2(no_zero=True)
    xi = [x - c * kdot / k2 for c, x in zip(k, xi)]
    nrm = np.sqrt(sum(np.abs(x) ** 2 for x in xi))
    nrm[nrm == 0] = 1.
    comps = []
    for x in xi:
        a = amp * x / nrm
        a[mask] = 0.
        a.flat[0] = 0.
        comps.append(hermitian_kx0(a, shape))
    ms = mean_square(comps, shape)
    s = np.sqrt(2. * tot_en / ms) if ms > 0 else 0.
    for i, a in enumerate(comps):
        data['u'][i]['kspace'] = a * s


def add_gaussian_white_noise(comp, std):
    """Random-phase noise of x-space std ``std`` added to a component (init_cond.py:440-464)."""
    kd = comp['kspace']
    phase = 2 * np.pi * np.random.random(kd.shape)
    amp = std / np.sqrt(comp.nmodes - 1)
    noise = amp * np.exp(1j * phase)
    zero_index = comp.find_mode([0., ] * comp.ndim)
    if zero_index:
        noise[zero_index] = 0.
    comp['kspace'] = kd + noise
    comp.enforce_hermitian()


def constant(data, comp, value):
    """Add a constant to the mean mode of a field or component (init_cond.py:466-490)."""
    if len(comp) == 1:
        rep = data[comp]
    elif len(comp) == 2:
        rep = data[comp[0]][comp[1]]
    else:
        raise ValueError("constant: comp %s invalid. must be 1 or 2 characters." % comp)
    zero_index = rep.find_mode([0., ] * rep.ndim)
    if not zero_index:
        return
    kd = rep['kspace']
    kd[zero_index] += value + 0j


def kida_vortex(data, a, chi=None, smooth=False):
    """Kida vortex: needs the shearing box, not part of the MI355X path."""
    raise NotImplementedError("kida_vortex needs a shearing-box representation")


def turb(ux, uy, spec, tot_en=0.5, **kwargs):
    """Superseded 2D generator of the reference; use turb_new."""
    raise NotImplementedError("turb is superseded by turb_new")
