"""Numpy restatement of dedalus-1.0's numpy path (oracle; test infrastructure).

Restates, with the reference's semantics and dataflow:

* ``FourierRepresentation`` numpy path  -- dedalus/data_objects/representations.py:67-440
* vector / scalar fields                 -- dedalus/data_objects/fields.py:59-245
* state container                        -- dedalus/data_objects/state_data.py:45-145
* physics operators and RHS              -- dedalus/physics/physics.py:134-819
* ETD kernels (2D and 3D variants)       -- dedalus/time_stepping/forward_step_cy_{2,3}d.pyx
* RK2mid / RK2trap                       -- dedalus/time_stepping/time_step.py:224-392

The side effects that decide rounding are kept on purpose: dealias on both
sides of every transform (in place on k data), ``XcrossY`` transforming the
*data* fields to x space (so the integrator re-transforms them), the shared
per-component derivative buffer, integrating factors taken from deriv1 only.
"""

from collections import OrderedDict

import numpy as np

# ---------------------------------------------------------------------------
# representation (one field component)
# ---------------------------------------------------------------------------


# The dealiasing method every new Component takes, as the reference's
# FourierRepresentation takes decfg's FFT.dealiasing at construction
# (representations.py:117, 129): '2/3' (also '2/3 cython'), '2/3 spherical'
# or 'None' (None and 0 too).
DEALIASING = '2/3'


def _swap01(seq):
    """Swap entries 0 and 1 (dedalus/utils/parallelism.py:155-173)."""
    out = list(seq)
    out[0], out[1] = out[1], out[0]
    return out


class Component:
    """One periodic field component, numpy FFT backend.

    x space holds float64 of shape (y, x) / (z, y, x); k space holds
    complex128 of the transposed half-spectrum shape (kx, ky) / (ky, kz, kx)
    (representations.py:82-93, 173-176).  The forward transform divides by
    N = prod(shape); the backward transform is unnormalised
    (representations.py:327-333).
    """

    def __init__(self, shape, length):
        self.xshape = tuple(int(n) for n in shape)
        self.length = np.asarray(length, dtype=np.float64)
        self.ndim = len(self.xshape)
        if self.ndim not in (2, 3):
            raise ValueError("Must use either 2 or 3 dimensions.")
        if self.xshape[-1] % 2:
            # rev_np calls irfftn without s=, which only round-trips an even
            # last axis (representations.py:331-333)
            raise ValueError("x extent must be even on the numpy path")
        self.npts = int(np.prod(self.xshape))
        kshape = list(self.xshape)
        kshape[-1] = kshape[-1] // 2 + 1
        self.kshape = tuple(_swap01(kshape))          # representations.py:174-176
        if self.ndim == 2:
            self.ktrans = {'x': 0, 0: 'x', 'y': 1, 1: 'y'}   # representations.py:120-122
        else:
            self.ktrans = {'x': 2, 2: 'x', 'y': 0, 0: 'y', 'z': 1, 1: 'z'}  # :123-125
        self.kdata = np.zeros(self.kshape, dtype=np.complex128)   # :187-189
        self.xdata = np.zeros(self.xshape, dtype=np.float64)
        self.deriv_data = np.zeros_like(self.kdata)               # :196
        self._setup_k()
        self.space = 'kspace'
        self.integrating_factor = None
        self.fwd_count = 0
        self.rev_count = 0
        self.set_dealiasing(DEALIASING)

    def set_dealiasing(self, dealiasing):
        """representations.py:359-382: the rule dealias() applies."""
        if dealiasing in ('2/3', '2/3 cython'):
            self.rule = '2/3'
        elif dealiasing == '2/3 spherical':
            self.rule = '2/3 spherical'
        elif dealiasing in ['None', None, 0]:
            self.rule = 'None'
        else:
            raise NotImplementedError("Specified dealiasing method not implemented.")

    # --- wavenumbers (representations.py:204-233) -------------------------
    def _setup_k(self):
        self.kny = np.array(_swap01(np.pi * np.array(self.xshape, dtype=np.float64)
                                    / self.length))
        xs = _swap01(self.xshape)
        rcomp = self.ktrans['x']
        self.k = {}
        for i, ksize in enumerate(self.kshape):
            n = xs[i]
            if i == rcomp:
                ki = np.fft.fftfreq(n)[:ksize] * 2. * self.kny[i]
                if n % 2 == 0:
                    ki[-1] *= -1.
            else:
                ki = np.fft.fftfreq(ksize) * 2. * self.kny[i]
                if n % 2 == 0:
                    ki[ksize // 2] *= -1.
            shp = [1] * self.ndim
            shp[i] = ksize
            self.k[self.ktrans[i]] = ki.reshape(shp)

    def k2(self, no_zero=False, set_zero=1.):
        """|k|^2 over the k array (representations.py:427-440)."""
        out = np.zeros(self.kshape)
        for kk in self.k.values():
            out = out + kk ** 2
        if no_zero:
            out[out == 0] = set_zero
        return out

    def dealias_mask(self):
        """True where the component's rule zeroes a mode: the 2/3 rule
        (representations.py:384-397), the spherical 2/3 rule (:410-416) or
        the Nyquist planes of zero_nyquist (:442-455)."""
        if self.rule == '2/3 spherical':
            return np.sqrt(self.k2()) >= 2 / 3. * np.min(self.kny)
        names = ['x', 'y'] if self.ndim == 2 else ['x', 'y', 'z']
        mask = np.zeros(self.kshape, dtype=bool)
        for nm in names:
            i = self.ktrans[nm]
            if self.rule == 'None':
                mask = mask | (np.abs(self.k[nm]) == self.kny[i])
            else:
                mask = mask | (np.abs(self.k[nm]) >= 2. / 3. * self.kny[i])
        return mask

    # --- lazy space state machine (representations.py:138-166, 235-245) ---
    @property
    def data(self):
        return self.kdata if self.space == 'kspace' else self.xdata

    def __getitem__(self, space):
        self.require_space(space)
        return self.data

    def __setitem__(self, space, value):
        if space == 'xspace':
            buf = self.xdata
        elif space == 'kspace':
            buf = self.kdata
        else:
            raise KeyError("space must be either xspace or kspace.")
        if isinstance(value, (float, complex, int)) and not isinstance(value, np.ndarray):
            buf[...] = value
        else:
            buf[...] = np.asarray(value)[tuple(slice(n) for n in buf.shape)]
        self.space = space

    def require_space(self, space):
        if self.space == space:
            return
        if space == 'xspace':
            self.backward()
        elif space == 'kspace':
            self.forward()
        else:
            raise ValueError("space must be either xspace or kspace.")

    # --- transforms (representations.py:327-357) ---------------------------
    def _perm(self):
        return [1, 0, 2][:self.ndim]

    def forward(self):
        if self.space == 'kspace':
            raise ValueError("Forward transform cannot be called from kspace.")
        self.kdata[...] = np.transpose(np.fft.rfftn(self.xdata / self.npts), self._perm())
        self.space = 'kspace'
        self.dealias()
        self.fwd_count += 1

    def backward(self):
        if self.space == 'xspace':
            raise ValueError("Backward transform cannot be called from xspace.")
        self.dealias()
        self.xdata[...] = np.fft.irfftn(np.transpose(self.kdata, self._perm())) * self.npts
        self.space = 'xspace'
        self.rev_count += 1

    def dealias(self):
        """The component's rule, in place on k data (representations.py:384-455)."""
        self.require_space('kspace')
        self.kdata[self.dealias_mask()] = 0.

    def deriv(self, dim):
        """i k_dim * data into the shared derivative buffer (representations.py:419-425)."""
        self.require_space('kspace')
        np.multiply(self.kdata, 1j * self.k[dim], self.deriv_data)
        return self.deriv_data

    def zero_all(self, space='kspace'):
        self[space] = 0.


# ---------------------------------------------------------------------------
# fields and state
# ---------------------------------------------------------------------------


class Field:
    """Vector (ncomp = ndim) or scalar (ncomp = 1) field (fields.py:59-245)."""

    def __init__(self, shape, length, ncomp):
        self.ncomp = ncomp
        self.ndim = len(shape)
        self.comps = [Component(shape, length) for _ in range(ncomp)]
        if ncomp == 1:
            self.ctrans = {'': 0, 0: ''}
        elif ncomp == 2:
            self.ctrans = {'x': 0, 0: 'x', 'y': 1, 1: 'y'}
        else:
            self.ctrans = {'x': 0, 0: 'x', 'y': 1, 1: 'y', 'z': 2, 2: 'z'}

    def __getitem__(self, item):
        if self.ncomp == 1 and item in ('kspace', 'xspace'):
            return self.comps[0][item]                     # fields.py:206-216
        if isinstance(item, str):
            item = self.ctrans[item]
        return self.comps[item]

    def __setitem__(self, item, value):
        if self.ncomp == 1 and item in ('kspace', 'xspace'):
            self.comps[0][item] = value                     # fields.py:218-228
            return
        raise TypeError("only scalar fields accept space assignment")

    def __iter__(self):
        return iter(enumerate(self.comps))

    def zero_all(self, space='kspace'):
        for c in self.comps:
            c[space] = 0.

    # scalar-field conveniences (fields.py:230-245)
    def deriv(self, dim):
        return self.comps[0].deriv(dim)

    def k2(self, **kw):
        return self.comps[0].k2(**kw)


class State:
    """Ordered fields plus time (state_data.py:45-145)."""

    def __init__(self, shape, length, field_list, time=0.):
        self.shape = tuple(shape)
        self.length = tuple(length)
        self.ndim = len(shape)
        self.time = time
        self.fields = OrderedDict()
        for name, kind in field_list:
            ncomp = self.ndim if kind == 'VectorField' else 1
            self.fields[name] = Field(shape, length, ncomp)

    def __getitem__(self, name):
        return self.fields[name]

    def __iter__(self):
        return iter(self.fields.items())

    def set_time(self, t):
        self.time = t


# ---------------------------------------------------------------------------
# physics (physics.py)
# ---------------------------------------------------------------------------

_TRANS = {0: 'x', 1: 'y', 2: 'z'}


class Physics:
    """Operators and RHS assembly of dedalus/physics/physics.py."""

    kind = None

    def __init__(self, shape, length=None):
        self.shape = tuple(shape)
        self.ndim = len(shape)
        self.length = tuple(length) if length else (2 * np.pi,) * self.ndim  # physics.py:69-72
        self.dims = range(self.ndim)
        self.parameters = {'viscosity_order': 1, 'nu': 0., 'shear_rate': 0.,
                           'Omega': None}                                     # physics.py:462-465
        self.field_list = [('u', 'VectorField')]
        self.aux_list = [('mathscalar', 'ScalarField'), ('mathvector', 'VectorField')]
        self.forcing_functions = {}
        self.aux = None
        self._k2 = None
        self._first_rhs = True

    def create_fields(self, t=0.):
        return State(self.shape, self.length, self.field_list, t)

    # --- operators --------------------------------------------------------
    def divX(self, X, out):
        """Divergence in k space (physics.py:180-195)."""
        out.comps[0]['kspace'] = 0.
        for i in range(X.ncomp):
            out.comps[0].kdata += X[i].deriv(_TRANS[i])

    def laplace_solve(self, X, out):
        """laplace(out) = X, k = 0 divided by 1 (physics.py:407-416)."""
        if self._k2 is None:
            self._k2 = out.k2(no_zero=True)
        out['kspace'] = X['kspace']
        out.comps[0].kdata /= self._k2
        out.comps[0].kdata *= -1.

    def XgradY(self, X, Y, stmp, vtmp, output):
        """(X.grad)Y pseudospectrally, result left in x space (physics.py:197-228)."""
        output.zero_all('xspace')
        for ci, comp in X:
            vtmp[ci]['kspace'] = comp['kspace']
        for ci, comp in Y:
            for i in self.dims:
                stmp['kspace'] = comp.deriv(_TRANS[i])
                sx = stmp['xspace']
                sx *= vtmp[i]['xspace']
                stmp['xspace'] = sx
                ox = output[ci]['xspace']
                ox += stmp['xspace']
                output[ci]['xspace'] = ox

    def XcrossY(self, X, Y, output):
        """X x Y in x space (physics.py:309-354); transforms X and Y in place."""
        if self.ndim == 2:
            if X.ncomp == 1:
                Xz = X['xspace']
            else:
                Xx = X['x']['xspace']
                Xy = X['y']['xspace']
        else:
            Xx = X['x']['xspace']
            Xy = X['y']['xspace']
            Xz = X['z']['xspace']
        Yx = Y['x']['xspace']
        Yy = Y['y']['xspace']
        if self.ndim == 3:
            Yz = Y['z']['xspace']
        if self.ndim == 2:
            if X.ncomp == 1:
                output['x']['xspace'] = -Xz * Yy
                output['y']['xspace'] = Xz * Yx
            else:
                output['xspace'] = Xx * Yy - Xy * Yx
        else:
            output['x']['xspace'] = Xy * Yz - Xz * Yy
            output['y']['xspace'] = Xz * Yx - Xx * Yz
            output['z']['xspace'] = Xx * Yy - Xy * Yx

    def curlX(self, X, output):
        """Curl in k space (physics.py:374-405)."""
        n = X.ncomp
        if n == 3:
            output[0]['kspace'] = X[2].deriv(_TRANS[1]) - X[1].deriv(_TRANS[2])
            output[1]['kspace'] = X[0].deriv(_TRANS[2]) - X[2].deriv(_TRANS[0])
            output[2]['kspace'] = X[1].deriv(_TRANS[0]) - X[0].deriv(_TRANS[1])
        elif n == 2:
            output['kspace'] = X[1].deriv(_TRANS[0]) - X[0].deriv(_TRANS[1])
        else:
            output[0]['kspace'] = X[0].deriv(_TRANS[1])
            output[1]['kspace'] = -X[0].deriv(_TRANS[0])

    def pressure_projection(self, data, deriv):
        """Solenoidal projection of deriv['u'] (physics.py:588-599)."""
        ms = self.aux['mathscalar']
        self.divX(deriv['u'], ms)
        self.laplace_solve(ms, ms)
        for i in self.dims:
            c = deriv['u'][i]
            c.require_space('kspace')
            c.kdata -= ms.deriv(_TRANS[i])

    # --- RHS --------------------------------------------------------------
    def _setup_integrating_factors(self, deriv):
        """physics.py:504-522 (velocity part)."""
        nu = self.parameters['nu']
        vo = self.parameters['viscosity_order']
        for _, comp in deriv['u']:
            comp.integrating_factor = None if nu == 0. else nu * comp.k2() ** vo

    def _rhs_prologue(self, data, deriv):
        """physics.py:134-149 plus the one-time IF setup of physics.py:537-540."""
        if self._first_rhs:
            self._setup_integrating_factors(deriv)
            self._first_rhs = False
        if self.aux is None:
            self.aux = State(self.shape, self.length, self.aux_list)
        for _, f in deriv:
            f.zero_all()
        for _, f in self.aux:
            f.zero_all('kspace')
        deriv.set_time(data.time)
        self.aux.set_time(data.time)

    def _hydro_rhs(self, data, deriv):
        """IncompressibleHydro.RHS without the final projection (physics.py:527-572)."""
        self._rhs_prologue(data, deriv)
        ms = self.aux['mathscalar']
        mv = self.aux['mathvector']
        self.XgradY(data['u'], data['u'], ms, mv, deriv['u'])
        for i in self.dims:
            c = deriv['u'][i]
            c.require_space('kspace')
            c.kdata *= -1.
        if 'VelocityForcing' in self.forcing_functions:
            for i in self.dims:
                c = deriv['u'][i]
                c.require_space('kspace')
                c.kdata += self.forcing_functions['VelocityForcing'](data, i)


class IncompressibleHydro(Physics):
    """physics.py:419-610."""

    kind = 'hydro'

    def RHS(self, data, deriv):
        self._hydro_rhs(data, deriv)
        # the projection runs only for exactly this class (physics.py:575):
        # a subclass that inherits this RHS gets the unprojected tendency
        if type(self) is IncompressibleHydro:
            self.pressure_projection(data, deriv)


class BoussinesqHydro(Physics):
    """physics.py:612-721."""

    kind = 'boussinesq'

    def __init__(self, shape, length=None, direction='z'):
        Physics.__init__(self, shape, length)
        self.field_list.append(('T', 'ScalarField'))
        self.parameters.update({'kappa': 0., 'g': 1., 'alpha_t': 1., 'beta': 1.,
                                'boussinesq_direction': direction})

    def _setup_integrating_factors(self, deriv):
        Physics._setup_integrating_factors(self, deriv)
        kappa = self.parameters['kappa']
        vo = self.parameters['viscosity_order']
        c = deriv['T'][0]
        c.integrating_factor = None if kappa == 0. else kappa * c.k2() ** vo   # physics.py:648-659

    def RHS(self, data, deriv):
        self._hydro_rhs(data, deriv)
        p = self.parameters
        d = p['boussinesq_direction']
        ms = self.aux['mathscalar']
        mv = self.aux['mathvector']
        # buoyancy (physics.py:690-694)
        ms['kspace'] = data['T']['kspace']
        ms.comps[0].kdata *= p['g']
        ms.comps[0].kdata *= p['alpha_t']
        c = deriv['u'][d]
        c.require_space('kspace')
        c.kdata += ms['kspace']
        if type(self) is BoussinesqHydro:                               # :697-698
            self.pressure_projection(data, deriv)
        # temperature advection (:700-703)
        self.XgradY(data['u'], data['T'], ms, mv, deriv['T'])
        deriv['T'].comps[0].require_space('kspace')
        deriv['T'].comps[0].kdata *= -1.
        # stratification (:705-708)
        ms['kspace'] = data['u'][d]['kspace']
        ms.comps[0].kdata *= p['beta']
        deriv['T'].comps[0].kdata -= ms['kspace']
        if 'ThermalForcing' in self.forcing_functions:                  # :710-712
            deriv['T'].comps[0].kdata += self.forcing_functions['ThermalForcing'](data)


class IncompressibleMHD(Physics):
    """physics.py:724-836."""

    kind = 'mhd'

    def __init__(self, shape, length=None):
        Physics.__init__(self, shape, length)
        self.field_list.append(('B', 'VectorField'))
        self.aux_list.append(('mathvector2', 'VectorField'))
        self.parameters.update({'rho0': 1., 'eta': 0.})

    def _setup_integrating_factors(self, deriv):
        Physics._setup_integrating_factors(self, deriv)
        eta = self.parameters['eta']
        vo = self.parameters['viscosity_order']
        for _, c in deriv['B']:
            c.integrating_factor = None if eta == 0. else eta * c.k2() ** vo  # physics.py:756-768

    def RHS(self, data, deriv):
        self._hydro_rhs(data, deriv)
        ms = self.aux['mathscalar']
        mv = self.aux['mathvector']
        mv2 = self.aux['mathvector2']
        fpr = 4 * np.pi * self.parameters['rho0']                       # physics.py:790-791
        # Lorentz force (:794-802)
        if self.ndim == 2:
            self.curlX(data['B'], ms)
            self.XcrossY(ms, data['B'], mv2)
        else:
            self.curlX(data['B'], mv)
            self.XcrossY(mv, data['B'], mv2)
        for i in self.dims:
            c = deriv['u'][i]
            c.require_space('kspace')
            c.kdata += mv2[i]['kspace'] / fpr
        if type(self) is IncompressibleMHD:                             # :804-806
            self.pressure_projection(data, deriv)
        # induction (:808-815)
        if self.ndim == 2:
            self.XcrossY(data['u'], data['B'], ms)
            self.curlX(ms, deriv['B'])
        else:
            self.XcrossY(data['u'], data['B'], mv)
            self.curlX(mv, deriv['B'])


# ---------------------------------------------------------------------------
# ETD kernels (forward_step_cy_2d.pyx / forward_step_cy_3d.pyx)
# ---------------------------------------------------------------------------


def f_taylor(k, Z):
    """sum_{j=k}^{14} Z^(j-k) / j!  (forward_step_cy_2d.pyx:123-148)."""
    Z = np.asarray(Z, dtype=np.float64)
    factj = 1.
    total = np.zeros_like(Z)
    zp = np.ones_like(Z)
    if k > 2:
        for j in range(1, k):
            factj *= j
    for j in range(k, 15):
        if j > 1:
            factj *= j
        total += zp / factj
        zp = zp * Z
    return total


def phi_functions(Z, ndim):
    """(f0, f1, f2) with the reference's branches.

    |Z| < 0.5: Taylor series to degree 14; f0 by Taylor in 2D
    (forward_step_cy_2d.pyx:54-56) and by exp in 3D
    (forward_step_cy_3d.pyx:54-56).  Otherwise f0 = exp(Z),
    f1 = (f0 - 1)/Z, f2 = (f1 - 1)/Z.  Entries with Z == 0 are returned as
    NaN here; callers take the Euler branch for them.
    """
    Z = np.asarray(Z, dtype=np.float64)
    small = (Z < 0.5) & (Z > -0.5)
    big = ~small
    f0 = np.empty_like(Z)
    f1 = np.empty_like(Z)
    f2 = np.empty_like(Z)
    zs = Z[small]
    f0[small] = f_taylor(0, zs) if ndim == 2 else np.exp(zs)
    f1[small] = f_taylor(1, zs)
    f2[small] = f_taylor(2, zs)
    zb = Z[big]
    eb = np.exp(zb)
    f0[big] = eb
    f1[big] = (eb - 1.) / zb
    f2[big] = (f1[big] - 1.) / zb
    return f0, f1, f2


def euler(start, out, deriv, dt):
    """forward_step_cy_2d.pyx:21-33."""
    out[...] = start + dt * deriv


def etd1(start, out, deriv, intfactor, dt, ndim):
    """forward_step_cy_2d.pyx:36-60 / _3d.pyx:34-60 (intfactor passed negated)."""
    Z = intfactor * dt
    f0, f1, _ = phi_functions(Z, ndim)
    zero = (Z == 0.)
    res = start * f0 + deriv * f1 * dt
    out[...] = np.where(zero, start + dt * deriv, res)


def etd2rk1(start, out, d1, d2, intfactor, dt, ndim):
    """forward_step_cy_2d.pyx:63-90 / _3d.pyx:63-92."""
    Z = intfactor * dt
    _, _, f2 = phi_functions(Z, ndim)
    zero = (Z == 0.)
    res = start + (d2 - d1) * f2 * dt
    out[...] = np.where(zero, start + dt / 2. * (d2 - d1), res)


def etd2rk2(start, out, d1, d2, intfactor, dt, ndim):
    """forward_step_cy_2d.pyx:93-120 / _3d.pyx:95-124."""
    Z = intfactor * dt
    f0, f1, f2 = phi_functions(Z, ndim)
    zero = (Z == 0.)
    res = start * f0 + (d2 - d1) * 2. * f2 * dt + d1 * f1 * dt
    out[...] = np.where(zero, start + dt * d2, res)


# ---------------------------------------------------------------------------
# integrators (time_step.py)
# ---------------------------------------------------------------------------


class RK2mid:
    """ETD midpoint RK2 (time_step.py:224-309)."""

    def __init__(self, physics):
        self.RHS = physics
        self.data2 = physics.create_fields()
        self.deriv1 = physics.create_fields()
        self.deriv2 = physics.create_fields()
        self.time = 0.
        self.iteration = 0

    def do_advance(self, data, dt):
        nd = self.RHS.ndim
        self.RHS.RHS(data, self.deriv1)
        for f in data.fields:
            for c in range(data[f].ncomp):
                IF = self.deriv1[f][c].integrating_factor
                y = data[f][c]['kspace']
                a = self.data2[f][c]['kspace']
                d1 = self.deriv1[f][c]['kspace']
                if IF is None:
                    euler(y, a, d1, dt / 2.)
                else:
                    etd1(y, a, d1, -IF, dt / 2., nd)
        self.data2.set_time(data.time + dt / 2.)
        self.RHS.RHS(self.data2, self.deriv2)
        for f in data.fields:
            for c in range(data[f].ncomp):
                IF = self.deriv1[f][c].integrating_factor
                y = data[f][c]['kspace']
                d1 = self.deriv1[f][c]['kspace']
                d2 = self.deriv2[f][c]['kspace']
                if IF is None:
                    euler(y, y, d2, dt)
                else:
                    etd2rk2(y, y, d1, d2, -IF, dt, nd)
        data.set_time(data.time + dt)
        self.time += dt
        self.iteration += 1


class RK2trap:
    """ETD trapezoidal RK2 (time_step.py:312-392)."""

    def __init__(self, physics):
        self.RHS = physics
        self.deriv1 = physics.create_fields()
        self.deriv2 = physics.create_fields()
        self.time = 0.
        self.iteration = 0

    def do_advance(self, data, dt):
        nd = self.RHS.ndim
        self.RHS.RHS(data, self.deriv1)
        for f in data.fields:
            for c in range(data[f].ncomp):
                IF = self.deriv1[f][c].integrating_factor
                y = data[f][c]['kspace']
                d1 = self.deriv1[f][c]['kspace']
                if IF is None:
                    euler(y, y, d1, dt)
                else:
                    etd1(y, y, d1, -IF, dt, nd)
        data.set_time(data.time + dt)
        self.RHS.RHS(data, self.deriv2)
        for f in data.fields:
            for c in range(data[f].ncomp):
                IF = self.deriv1[f][c].integrating_factor
                y = data[f][c]['kspace']
                d1 = self.deriv1[f][c]['kspace']
                d2 = self.deriv2[f][c]['kspace']
                if IF is None:
                    euler(y, y, d2 - d1, dt / 2.)
                else:
                    etd2rk1(y, y, d1, d2, -IF, dt, nd)
        self.time += dt
        self.iteration += 1


# ---------------------------------------------------------------------------
# convenience wrappers used by the tests and the CPU baseline
# ---------------------------------------------------------------------------

PHYSICS = {'hydro': IncompressibleHydro, 'boussinesq': BoussinesqHydro,
           'mhd': IncompressibleMHD}


def make_physics(kind, shape, length=None, **params):
    if kind == 'boussinesq':
        direction = params.pop('boussinesq_direction', 'z')
        ph = BoussinesqHydro(shape, length, direction)
    else:
        ph = PHYSICS[kind](shape, length)
    ph.parameters.update(params)
    return ph


def load_state(physics, kfields):
    """New state with k-space arrays set from {name: [comp arrays]}."""
    data = physics.create_fields()
    for name, comps in kfields.items():
        for i, arr in enumerate(comps):
            data[name][i]['kspace'] = np.array(arr, dtype=np.complex128)
    return data


def read_state(data):
    """{name: [k-space comp arrays]} (transforms to k space as needed)."""
    return {name: [c['kspace'].copy() for _, c in f] for name, f in data}


def run(kind, shape, kfields, dt, nsteps, integrator='RK2mid', length=None, forcing=None,
        **params):
    """Integrate nsteps from kfields and return the final k-space fields.

    forcing: optional list of reference-layout k arrays, one per velocity
    component, returned by a state-independent VelocityForcing function
    (physics.py:524-525, 569-572)."""
    ph = make_physics(kind, shape, length, **params)
    if forcing is not None:
        fk = [np.array(f) for f in forcing]
        ph.forcing_functions['VelocityForcing'] = lambda data, i: fk[i]
    data = load_state(ph, kfields)
    ti = (RK2mid if integrator == 'RK2mid' else RK2trap)(ph)
    for _ in range(nsteps):
        ti.do_advance(data, dt)
    return read_state(data)


def rhs(kind, shape, kfields, length=None, **params):
    """One RHS evaluation; returns {name: [F comp arrays]}."""
    ph = make_physics(kind, shape, length, **params)
    data = load_state(ph, kfields)
    deriv = ph.create_fields()
    ph.RHS(data, deriv)
    return read_state(deriv)
