"""Representation classes for transformable fields, on the MI355X.

``FourierRepresentation`` keeps the interface of the reference class
(dedalus/data_objects/representations.py:57-556): ``rep['kspace']`` /
``rep['xspace']`` return the data in that space, transforming lazily;
assignment copies into the existing buffer; ``deriv``, ``k2``, ``k``,
``kny``, ``find_mode``, ``xspace_grid`` behave as there.

Where the data lives
--------------------
The canonical copy is on the device: k space as the library's compact
array of the modes kept by the dealiasing rule, x space as float64 (z, y, x).
Host arrays in the reference layout -- ``kdata`` (complex128, (kx, ky) /
(ky, kz, kx)) and ``xdata`` -- are mirrors, filled on first access.  An
array handed out by indexing or by ``.data`` / ``.kdata`` / ``.xdata``
counts as possibly modified: the next device operation on that space
uploads it first.  Arrays obtained before a device operation are
snapshots; index again afterwards to see the new values.

Modes the dealiasing rule removes are not stored on the device (for the
spherical rule, the compact cube holds the ball and the shell between them is
zeroed where the reference dealiases), so values a caller writes into removed
modes survive only until the next device operation --
the reference clears them at its next dealias (representations.py:344,
353); the ETD update alone would carry them along, inert, without ever
feeding them into x space.
"""

import numpy as np

from dedalus import _hip
from dedalus.config import decfg
from dedalus.data_objects import kgrid
from dedalus.utils.logger import mylog
from dedalus.utils.parallelism import swap_indices
from dedalus.utils.timer import timer

_OPCODE = {'x': 0, 'y': 1, 'z': 2}


# FourierRepresentation.set_dealiasing's methods (representations.py:359-382)
# and the device rule of each; the name of the rule for kgrid.dealias_mask
_RULES = {'2/3': (_hip.DEALIAS_23, '2/3'), '2/3 cython': (_hip.DEALIAS_23, '2/3'),
          '2/3 spherical': (_hip.DEALIAS_SPHERICAL, '2/3 spherical'),
          'None': (_hip.DEALIAS_NONE, 'None')}


def dealias_rule(dealiasing):
    """(device rule, rule name) of a dealiasing method."""
    if dealiasing is None or (isinstance(dealiasing, int) and not isinstance(dealiasing, bool)
                              and dealiasing == 0):
        dealiasing = 'None'
    if dealiasing not in _RULES:
        raise NotImplementedError("Specified dealiasing method not implemented.")
    return _RULES[dealiasing]


class Representation:
    """Representation of a field component (representations.py:47-55)."""

    def __init__(self, sd, shape, length):
        pass


class IntegratingFactor:
    """Lazy coeff * |k|^(2 order) on a component (physics.py:504-522).

    The device kernels recompute it from the wavenumbers; the array is only
    built when host code uses the value (``np.asarray(IF)``, ``-IF``).
    """

    def __init__(self, coeff, order, comp):
        self.coeff = float(coeff)
        self.order = int(order)
        self._comp = comp

    def __array__(self, dtype=None, copy=None):
        a = self.coeff * self._comp.k2() ** self.order
        return a if dtype is None else a.astype(dtype)

    def __neg__(self):
        return -np.asarray(self)

    def __repr__(self):
        return "IntegratingFactor(coeff=%g, order=%d)" % (self.coeff, self.order)


class FourierRepresentation(Representation):
    """A field component periodic in every direction, transformed on the GPU."""

    timer = timer
    _static_k = True

    def __init__(self, sd, shape, length):
        self.sd = sd
        self.global_shape = {'xspace': np.array(shape)}
        self.length = np.asarray(length, dtype=np.float64)
        self.ndim = len(shape)
        if self.ndim not in (2, 3):
            raise ValueError("Must use either 2 or 3 dimensions.")
        if len(shape) != len(length):
            raise ValueError("Shape and Length must have same dimensions.")
        self.dtype = {'kspace': 'complex128', 'xspace': 'float64'}
        self._eps = {'kspace': np.finfo(self.dtype['kspace']).eps,
                     'xspace': np.finfo(self.dtype['xspace']).eps}
        method = decfg.get('FFT', 'method')
        dealiasing = decfg.get('FFT', 'dealiasing')
        self.ktrans = kgrid.ktrans(self.ndim)
        if self.ndim == 2:
            self.xtrans = {'x': 1, 1: 'x', 'y': 0, 0: 'y'}
        else:
            self.xtrans = {'x': 2, 2: 'x', 'y': 1, 1: 'y', 'z': 0, 0: 'z'}
        self._rule = dealias_rule(dealiasing)
        self.set_fft(method)
        self.set_dealiasing(dealiasing)
        self._setup_k()
        self._curr_space = 'kspace'
        self.fwd_count = 0
        self.rev_count = 0

    # ------------------------------------------------------------------
    # memory and coherence
    # ------------------------------------------------------------------
    @timer
    def _allocate_memory(self, method):
        kshape = self.global_shape['xspace'].copy()
        kshape[-1] = kshape[-1] // 2 + 1
        self.global_shape['kspace'] = np.array(swap_indices(kshape))
        self.local_shape = {'kspace': self.global_shape['kspace'].copy(),
                            'xspace': self.global_shape['xspace'].copy()}
        self.offset = {'xspace': 0, 'kspace': 0}
        _hip.init(decfg.getint('hip', 'device'))
        prec = decfg.get('hip', 'precision')
        if prec not in ('float64', 'float32'):
            raise ValueError("[hip] precision must be float64 or float32")
        self._plan = _hip.get_plan(tuple(int(n) for n in self.global_shape['xspace']),
                                   tuple(float(x) for x in self.length),
                                   _hip.F64 if prec == 'float64' else _hip.F32,
                                   self._rule[0])
        self._kc = self._plan.alloc_k()
        self._xd = None
        self._dd = None
        self._host = {'kspace': None, 'xspace': None}
        self._hderiv = None
        self._newest = {'kspace': 'dev', 'xspace': 'dev'}
        self._lent = set()

    def _xbuf(self):
        if self._xd is None:
            self._xd = self._plan.alloc_x()
        return self._xd

    def _host_array(self, space):
        a = self._host[space]
        if a is None:
            shp = tuple(int(n) for n in self.global_shape[space])
            a = np.zeros(shp, dtype=self.dtype[space])
            self._host[space] = a
        return a

    def _to_host(self, space):
        """Host mirror of a space, downloaded if the device copy is newer; lent."""
        arr = self._host_array(space)
        if self._newest[space] == 'dev':
            if space == 'kspace':
                self._plan.download_k(self._kc, out=arr)
            else:
                self._plan.download_x(self._xbuf(), out=arr)
            self._newest[space] = 'both'
        self._lent.add(space)
        return arr

    def _to_dev(self, space):
        """Device buffer of a space, uploading a newer or lent host mirror."""
        if space in self._lent or self._newest[space] == 'host':
            arr = self._host_array(space)
            if space == 'kspace':
                self._plan.upload_k(self._kc, arr)
                # the upload drops the modes the 2/3 rule removes: the host
                # mirror is re-read from the device on next access
                self._newest['kspace'] = 'dev'
            else:
                self._plan.upload_x(self._xbuf(), arr)
                self._newest['xspace'] = 'both'
            self._lent.discard(space)
        return self._kc if space == 'kspace' else self._xbuf()

    def _mark(self, space):
        """The device buffer of ``space`` was just written; it is current."""
        self._newest[space] = 'dev'
        self._lent.discard(space)
        self._curr_space = space

    def _kdev(self):
        """Device k buffer, transforming to k space first if needed."""
        self.require_space('kspace')
        return self._to_dev('kspace')

    def _xdev(self):
        """Device x buffer, transforming to x space first if needed."""
        self.require_space('xspace')
        return self._to_dev('xspace')

    def _kout(self):
        """Device k buffer about to be overwritten (no transform, no upload)."""
        self._lent.discard('kspace')
        return self._kc

    def _xout(self):
        self._lent.discard('xspace')
        return self._xbuf()

    # ------------------------------------------------------------------
    # reference interface
    # ------------------------------------------------------------------
    @property
    def data(self):
        return self._to_host(self._curr_space)

    @property
    def kdata(self):
        return self._to_host('kspace')

    @property
    def xdata(self):
        return self._to_host('xspace')

    @property
    def deriv_data(self):
        if self._hderiv is None:
            self._hderiv = np.zeros(tuple(self.global_shape['kspace']), dtype=np.complex128)
        return self._hderiv

    def __getitem__(self, space):
        """Return data in specified space, transforming as necessary."""
        self.require_space(space)
        return self._to_host(space)

    def __setitem__(self, space, data):
        """Copy into the existing buffer of ``space`` (representations.py:144-166)."""
        if space not in ('xspace', 'kspace'):
            raise KeyError("space must be either xspace or kspace.")
        if isinstance(data, (float, complex, int)) and not isinstance(data, bool):
            if data == 0:
                buf = self._kc if space == 'kspace' else self._xbuf()
                self._plan.zero(buf)
                self._mark(space)
                return
            arr = self._host_array(space)
            arr[...] = data
        else:
            arr = self._host_array(space)
            src = np.asarray(data)
            arr[...] = src[tuple(slice(n) for n in arr.shape)]
        self._newest[space] = 'host'
        self._lent.discard(space)
        self._curr_space = space

    def _setup_k(self):
        """Wavenumber arrays (representations.py:204-233)."""
        xs = tuple(int(n) for n in self.global_shape['xspace'])
        self.dk = np.array(swap_indices(2 * np.pi / self.length))
        self.k, self.kny, _ = kgrid.wavenumbers(xs, self.length)

    def require_space(self, space):
        if self._curr_space == space:
            return
        if space == 'xspace':
            self.backward()
        elif space == 'kspace':
            self.forward()
        else:
            raise ValueError("space must be either xspace or kspace.")

    def find_mode(self, mode, exact=False):
        """Index of the mode closest to ``mode`` (k order), or None (representations.py:247-288)."""
        dk = 2 * np.pi / np.array(swap_indices(self.length))
        test = np.ones(tuple(self.local_shape['kspace']), dtype=bool)
        for name, kn in self.k.items():
            i = self.ktrans[name]
            if exact:
                itest = (kn == mode[i])
            else:
                itest = (kn <= mode[i] + dk[i] / 2.) & (kn > mode[i] - dk[i] / 2.)
            test = test & itest
        index = list(zip(*test.nonzero()))
        if len(index) == 0:
            return None
        if len(index) == 1:
            return index[0]
        raise ValueError("Multiple modes tested true. This shouldn't happen.")

    def set_fft(self, method):
        mylog.debug("Setting FFT method to %s." % method)
        if method == 'fftw':
            # the reference's default method name selects this package's transforms
            method = 'hip'
        if method != 'hip':
            raise NotImplementedError(
                "FFT method %r is not provided: this package transforms on the MI355X "
                "(FFT.method = 'hip')" % method)
        self._allocate_memory(method)
        self.fft = self.fwd_hip
        self.ifft = self.rev_hip

    @timer
    def fwd_hip(self):
        self._plan.fwd(self._to_dev('xspace'), self._kout())

    @timer
    def rev_hip(self):
        self._plan.bwd(self._to_dev('kspace'), self._xout())

    def forward(self):
        """xspace -> kspace, 1/N normalised, dealiased (representations.py:335-345)."""
        if self._curr_space == 'kspace':
            raise ValueError("Forward transform cannot be called from kspace.")
        self.fft()
        self._mark('kspace')
        self.fwd_count += 1

    def backward(self):
        """kspace -> xspace of the dealiased data (representations.py:347-357)."""
        if self._curr_space == 'xspace':
            raise ValueError("Backward transform cannot be called from xspace.")
        self.ifft()
        self._mark('xspace')
        self.rev_count += 1

    def set_dealiasing(self, dealiasing):
        """Assign the dealiasing method (representations.py:359-382).  The
        device plan follows the rule; changing it moves the data through the
        host onto a plan of the new rule."""
        mylog.debug("Setting dealiasing method to %s." % dealiasing)
        rule = dealias_rule(dealiasing)
        n = self.global_shape['xspace']
        if rule[0] == _hip.DEALIAS_23:
            self.dealias = self.dealias_23
            self.nmodes = np.prod(2 * np.ceil(n / 3. - 1) + 1)
        elif rule[0] == _hip.DEALIAS_SPHERICAL:
            self.dealias = self.dealias_23_spherical
            self.nmodes = None
        else:
            self.dealias = self.zero_nyquist
            self.nmodes = np.prod(2 * np.ceil(n / 2. - 1) + 1)
        if rule[0] != self._rule[0]:
            space = self._curr_space
            arr = self._to_host(space).copy()
            self._rule = rule
            self._allocate_memory('hip')
            self[space] = arr
        self._rule = rule

    def dealias_23(self):
        """Orszag 2/3 rule: on the device the removed modes are never stored."""
        self.require_space('kspace')
        self._to_dev('kspace')

    dealias_23_cython = dealias_23

    def dealias_23_spherical(self):
        """Spherical 2/3 rule (representations.py:410-416), in place on the device."""
        self.require_space('kspace')
        self._plan.dealias(self._to_dev('kspace'))
        self._mark('kspace')

    def dealias_mask(self):
        """True where this component's dealiasing rule removes a mode."""
        return kgrid.dealias_mask(tuple(int(n) for n in self.global_shape['xspace']), self.length,
                                  self._rule[1])

    def deriv(self, dim):
        """i k_dim * data, into the component's derivative buffer (representations.py:419-425)."""
        self.require_space('kspace')
        if self._dd is None:
            self._dd = self._plan.alloc_k()
        self._plan.kcombine(self._dd, [(self._to_dev('kspace'), 1.0, _OPCODE[dim])])
        return self._plan.download_k(self._dd, out=self.deriv_data)

    def k2(self, no_zero=False, set_zero=1.):
        """|k|^2 over the k array (representations.py:427-440)."""
        k2 = np.zeros(tuple(self.local_shape['kspace']))
        for k in self.k.values():
            k2 = k2 + k ** 2
        if no_zero:
            k2[k2 == 0] = set_zero
        return k2

    def zero_nyquist(self):
        """Zero the Nyquist planes (representations.py:442-455)."""
        self.require_space('kspace')
        arr = self._to_host('kspace')
        names = ('x', 'y') if self.ndim == 2 else ('x', 'y', 'z')
        mask = np.zeros(arr.shape, dtype=bool)
        for nm in names:
            mask = mask | (np.abs(self.k[nm]) == self.kny[self.ktrans[nm]])
        arr[mask] = 0.

    def enforce_hermitian(self):
        """Hermitian symmetry on the kx = 0 plane (representations.py:457-503)."""
        self.zero_nyquist()
        arr = self._to_host('kspace')
        if self.ndim == 2:
            arr[0, 0] = arr[0, 0].real
            ny = arr.shape[1]
            nyh = ny // 2
            grab = arr[0, 1:nyh + 1].copy()
            arr[0, ny - nyh:] = grab[::-1].conj()
        else:
            plane = arr[:, :, 0]
            nyy, nyz = np.array(plane.shape) // 2
            grab = plane[0, 1:nyz + 1].copy()
            plane[0, -nyz:] = grab[::-1].conj()
            grab = plane[1:nyy + 1, 0].copy()
            plane[-nyy:, 0] = grab[::-1].conj()
            grab = plane[1:nyy + 1, 1:].copy()
            plane[-nyy:, 1:] = grab[::-1, ::-1].conj()

    def zero_under_eps(self):
        self.require_space('kspace')
        arr = self._to_host('kspace')
        arr[np.abs(arr) < self._eps['kspace']] = 0.

    def save(self, group, name):
        """Store the current space's data into a dict (snapshot writer)."""
        group[name] = np.array(self.data)
        group[name + '.space'] = np.array(self._curr_space)

    def xspace_grid(self, open=False):
        """x-space grid [z, y, x] / [y, x] (representations.py:528-548)."""
        ref = np.ogrid if open else np.mgrid
        grid = ref[tuple(slice(int(n)) for n in self.local_shape['xspace'])]
        if open:
            grid = [np.asarray(g, dtype=np.float64) for g in grid]
        else:
            grid = np.asarray(grid, dtype=np.float64)
        grid[0] += self.offset['xspace']
        dx = self.dx()
        for i in range(self.ndim):
            grid[i] *= dx[i]
        return grid

    def dx(self):
        """Grid spacing [z, y, x] / [y, x] (representations.py:550-557)."""
        return self.length / self.global_shape['xspace']


class FourierShearRepresentation(FourierRepresentation):
    """Shearing-box representation (representations.py:558-740): not provided."""

    _static_k = False

    def __init__(self, *args, **kwargs):
        raise NotImplementedError("shearing-box representations are not part of the MI355X path")


class ChebyshevRepresentation(Representation):
    """Empty in the reference (representations.py:743)."""

    def __init__(self, *args, **kwargs):
        raise NotImplementedError("Chebyshev representation is a stub in dedalus-1.0")
