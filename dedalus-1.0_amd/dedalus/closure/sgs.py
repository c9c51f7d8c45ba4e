"""Coarse-graining and the sub-grid tendency, on the device.

For a fine 2D velocity u on N x N and the coarse grid M x M (N = f M):

    ubar   = B(u)                       box average over f x f cells, then the
                                        coarse forward transform (2/3 rule)
    tau    = B(NL_N(u)) - NL_M(ubar)    sub-grid tendency of the coarse velocity

NL is the nonlinear right-hand side of IncompressibleHydro.RHS
(dedalus/physics/physics.py:527-572 with no forcing: -P[u.grad u], dealiased
and projected; viscosity is in the integrating factor, not in NL).  tau is
exactly what a coarse run lacks: d ubar/dt = NL_M(ubar) + tau - nu k^2 ubar
+ B(f) when the fine run obeys du/dt = NL_N(u) - nu k^2 u + f, up to the
box filter's commutation error with the viscous term.

Network features are the coarse (u, v, omega) in x space; targets are tau's
two components in x space.
"""

import ctypes

import numpy as np

from .. import _hip
from ..init_cond.seeded import random_solenoidal


def band_forcing(shape, length, seed, kmin=3.5, kmax=4.5, rms=0.1):
    """State-independent solenoidal forcing on kmin <= |k| < kmax with random
    phases fixed by seed; x-space rms = rms (reference-layout k arrays)."""
    spec = lambda k: np.where((k >= kmin) & (k < kmax), 1.0, 0.0)
    return random_solenoidal(shape, length, seed, spec, rms=rms)


def initial_velocity(shape, length, seed, rms=1.0):
    """Random-phase solenoidal velocity with E(k) ~ k^-3 on 2 <= |k| <= 10."""
    spec = lambda k: np.where((k >= 2) & (k <= 10), k ** -3.0, 0.0)
    return random_solenoidal(shape, length, seed, spec, rms=rms)


def _stream_of(stream):
    return ctypes.c_void_p(int(stream) if stream else None)


class CoarseGrainer:
    """Fine plan N x N -> coarse plan M x M (float64), all launches on one
    stream; with ``batch`` both plans hold that many members."""

    def __init__(self, fine_shape, factor, length=(2 * np.pi, 2 * np.pi), stream=None, batch=None):
        self.factor = int(factor)
        self.fine_shape = tuple(int(n) for n in fine_shape)
        if len(self.fine_shape) != 2 or any(n % self.factor for n in self.fine_shape):
            raise ValueError("2D grid with extents divisible by the factor required")
        self.coarse_shape = tuple(n // self.factor for n in self.fine_shape)
        self.length = tuple(length)
        self.batch = int(batch) if batch else None
        self.fine = _hip.Plan(self.fine_shape, self.length, batch=self.batch)
        self.coarse = _hip.Plan(self.coarse_shape, self.length, batch=self.batch)
        self.stream = int(stream) if stream else None
        for p in (self.fine, self.coarse):
            p.set_stream(self.stream)
        self.params = _hip.make_params(_hip.HYDRO)  # NL only: no viscosity in the RHS
        F, C = self.fine, self.coarse
        self._xf = F.alloc_x()
        self._df = [F.alloc_k() for _ in range(2)]
        self._xc = [C.alloc_x() for _ in range(3)]
        self._db = [C.alloc_k() for _ in range(2)]
        self._dc = [C.alloc_k() for _ in range(2)]
        self._wk = C.alloc_k()

    def _box(self, src, dst):
        # stacked members (B, ny, nx) are one (B ny, nx) image: ny % factor == 0,
        # so no averaging block straddles two members
        ny, nx = self.fine_shape
        ny *= self.batch or 1
        _hip._check(_hip._lib.ddx_box_filter(_stream_of(self.stream), ny, nx, self.factor, 1,
                                             ctypes.c_void_p(int(src)), ctypes.c_void_p(int(dst))))

    def filter_k(self, fine_k, coarse_k):
        """Coarse compact k of the box-filtered fine field (one component)."""
        self.fine.bwd(fine_k, self._xf)
        self._box(self._xf, self._xc[0])
        self.coarse.fwd(self._xc[0], coarse_k)

    def pair(self, u_fine, ubar, tau):
        """ubar (2 coarse k buffers) and tau (2 coarse k buffers) from the fine velocity."""
        for c in range(2):
            self.filter_k(u_fine[c], ubar[c])
        self.fine.rhs(self.params, u_fine, self._df)
        for c in range(2):
            self.filter_k(self._df[c], self._db[c])
        self.coarse.rhs(self.params, ubar, self._dc)
        for c in range(2):
            self.coarse.kcombine(tau[c], [(self._db[c], 1.0, -1), (self._dc[c], -1.0, -1)])

    def vorticity_k(self, ubar, out):
        """omega = d_x v - d_y u (coarse compact k)."""
        vorticity_k(self.coarse, ubar, out)

    def features_x(self, ubar, xbufs):
        """x-space u, v, omega of the coarse velocity into three coarse x buffers."""
        features_x(self.coarse, ubar, xbufs, self._wk)


def vorticity_k(plan, u, out):
    """omega = d_x v - d_y u on any 2D plan (batched plans: every member)."""
    plan.kcombine(out, [(u[1], 1.0, 0), (u[0], -1.0, 1)])


# the features' recipes: u, v and omega = d_x v - d_y u
FEATURE_RECIPES = [[(0, -1, 1.0)], [(1, -1, 1.0)], [(1, 0, 1.0), (0, 1, -1.0)]]


def features_x(plan, u, xbufs, work_k=None):
    """x-space u, v, omega of a 2D velocity into three x buffers of the plan:
    one backward transform call that forms omega in its first column pass
    (work_k is unused; omega needs no k-space array of its own)."""
    plan.bwd_combine([u[0], u[1]], FEATURE_RECIPES, xbufs[:3])


def shell_spectrum(kref_comps, shape, length):
    """Hermitian-weighted shell sum E(k) = 1/2 sum_{k-1/2 <= |k'| < k+1/2} w |u_k'|^2
    over reference-layout (kx, ky) arrays; w = 1 on kx = 0, 2 elsewhere, so
    sum_k E(k) = <|u|^2> / 2 (Parseval).  Returns (k = 0..kmax, E)."""
    from ..data_objects.kgrid import wavenumbers
    k, kny, ksh = wavenumbers(shape, length)
    kx = np.broadcast_to(k['x'], ksh)
    ky = np.broadcast_to(k['y'], ksh)
    # integer shells for the 2 pi box (k in units of 2 pi / L)
    scale = length[0] / (2 * np.pi)
    kmag = np.sqrt(kx ** 2 + ky ** 2) * scale
    w = np.full(ksh, 2.0)
    w[0, :] = 1.0
    p = sum(np.abs(a) ** 2 for a in kref_comps) * w * 0.5
    shell = np.floor(kmag + 0.5).astype(int)
    nmax = int(shell.max()) + 1
    E = np.bincount(shell.ravel(), weights=p.ravel(), minlength=nmax)
    return np.arange(nmax), E
