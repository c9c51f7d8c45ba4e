"""Physics: right-hand sides of the fluid equations, on the MI355X.

Same classes, parameters and operator methods as dedalus/physics/physics.py:42-836.

* ``IncompressibleHydro.RHS``, ``BoussinesqHydro.RHS`` and
  ``IncompressibleMHD.RHS`` run the fused device right-hand side
  (``ddx_rhs``): spectral prologue, pruned inverse transforms, x-space
  products, forward transforms, projection / curl / buoyancy epilogue.  The
  velocity nonlinearity is evaluated as u x (curl u) -- equal to
  -(u.grad)u up to a gradient, which the pressure projection removes --
  so the step needs 6 inverse + 3 forward transforms per 3D hydro RHS
  (18 for MHD) instead of the reference's 15 (36).
* The operator methods (``gradX``, ``divX``, ``XgradY``, ``XcrossY``,
  ``curlX``, ``laplace_solve`` ...) keep the reference's dataflow and
  transform sequence, each step a device kernel, for physics classes built
  on them.
* Python forcing callables are called on the host with the state, as in
  the reference; their k-space result is uploaded and added on the device.
* Shear, rotation and the passive tracer are not part of the MI355X path
  and raise ``NotImplementedError``.
"""

import numpy as np

from dedalus import _hip
from dedalus.config import decfg
from dedalus.data_objects.api import AuxEquation, StateData, create_field_classes
from dedalus.data_objects.representations import IntegratingFactor
from dedalus.utils.logger import mylog
from dedalus.utils.timer import timer


def _reconstruct_object(*args, **kwargs):
    new_args = [args[1]['shape'], args[1]['_representation'], args[1]['length']]
    obj = args[0](*new_args)
    obj.__dict__.update(args[1])
    return obj


def _plan(field):
    return field[0]._plan if hasattr(field, 'ncomp') else field._plan


class Physics:
    """Base class: fields, parameters and k/x-space operators (physics.py:42-416)."""

    _phys_id = None

    def __init__(self, shape, representation, length=None):
        self.shape = shape
        self._representation = representation
        self.length = length if length else (2 * np.pi,) * len(self.shape)
        self.ndim = len(self.shape)
        self.dims = range(self.ndim)
        self._field_list = []
        self._aux_field_list = []
        self.parameters = {}
        self.aux_eqns = {}
        self.forcing_functions = {}
        self._forcing_function_names = {}
        self._field_classes = create_field_classes(self._representation, self.shape, self.length)
        self._is_finalized = False
        self._tracer = decfg.getboolean('physics', 'use_tracer')
        self.k2 = None
        self._trans = {0: 'x', 1: 'y', 2: 'z'}
        self._force_bufs = None

    def __getitem__(self, item):
        value = self.parameters.get(item, None)
        if value is None:
            raise KeyError(item)
        return value

    def _finalize(self):
        self._setup_aux_fields(0., self._aux_field_list)
        self._is_finalized = True

    def create_fields(self, time, field_list=None):
        if field_list is None:
            field_list = self._field_list
        return StateData(time, self.shape, self.length, self._field_classes,
                         field_list=field_list, params=self.parameters)

    def _setup_aux_fields(self, t, aux_field_list=None):
        self.aux_fields = self.create_fields(t, aux_field_list)

    def _setup_aux_eqns(self, aux_eqns, RHS, ics, kwarglists):
        for f, r, ic, kwargs in zip(aux_eqns, RHS, ics, kwarglists):
            self.aux_eqns[f] = AuxEquation(r, kwargs, ic)

    def RHS(self, data, deriv, zero_outputs=True):
        """Finalize once, zero deriv and aux fields, synchronise times (physics.py:134-149)."""
        if not self._is_finalized:
            self._finalize()
        if zero_outputs:
            for fname, field in deriv:
                field.zero_all()
            for fname, field in self.aux_fields:
                field.zero_all(space='kspace')
        deriv.set_time(data.time)
        self.aux_fields.set_time(data.time)

    def compute_dt(self, data):
        self.dtlist = []
        self.set_dtlist(data)
        return min(self.dtlist)

    # ------------------------------------------------------------------
    # operators (device kernels, reference dataflow)
    # ------------------------------------------------------------------
    def gradX(self, X, output):
        """gradX[N*i + j] = dX_i/dx_j in k space (physics.py:160-178)."""
        N = self.ndim
        plan = _plan(X)
        for i in range(X.ncomp):
            src = X[i]._kdev()
            for j in self.dims:
                out = output[N * i + j]
                plan.kcombine(out._kout(), [(src, 1.0, j)])
                out._mark('kspace')

    def divX(self, X, output):
        """div X in k space (physics.py:180-195)."""
        plan = _plan(X)
        plan.kcombine(output._kout(), [(X[i]._kdev(), 1.0, i) for i in range(X.ncomp)])
        output._mark('kspace')

    def XgradY(self, X, Y, stmp, vtmp, output):
        """(X.grad)Y, left in x space (physics.py:197-228)."""
        plan = _plan(X)
        for ci, c in output:
            plan.zero(c._xout())
            c._mark('xspace')
        for ci, comp in X:
            plan.copy(vtmp[ci]._kout(), comp._kdev())
            vtmp[ci]._mark('kspace')
        for ci, comp in Y:
            for i in self.dims:
                plan.kcombine(stmp._kout(), [(comp._kdev(), 1.0, i)])
                stmp._mark('kspace')
                sx = stmp._xdev()
                plan.xcombine(sx, [(sx, vtmp[i]._xdev(), 1.0)])
                stmp._mark('xspace')
                ox = output[ci]._xdev()
                plan.xcombine(ox, [(sx, None, 1.0)], accumulate=True)
                output[ci]._mark('xspace')

    def XlistgradY(self, Xlist, Y, stmp, vtmp, outlist):
        """(X.grad)Y for every X of Xlist (physics.py:230-263)."""
        plan = _plan(Y)
        for outfield in outlist:
            for ci, c in outfield:
                plan.zero(c._xout())
                c._mark('xspace')
        for i, X in enumerate(Xlist):
            for j in range(X.ncomp):
                plan.copy(vtmp[i][j]._kout(), X[j]._kdev())
                vtmp[i][j]._mark('kspace')
                vtmp[i][j]._xdev()
        for i in range(Y.ncomp):
            for j in self.dims:
                plan.kcombine(stmp._kout(), [(Y[i]._kdev(), 1.0, j)])
                stmp._mark('kspace')
                sx = stmp._xdev()
                for k, X in enumerate(vtmp):
                    o = outlist[k][i]
                    plan.xcombine(o._xdev(), [(X[j]._xdev(), sx, 1.0)], accumulate=True)
                    o._mark('xspace')

    def XconstcrossY(self, X, Y, output):
        """Constant X cross Y in k space (physics.py:266-307)."""
        plan = _plan(Y)
        Yx = Y['x']._kdev()
        Yy = Y['y']._kdev()
        if self.ndim == 2:
            if np.isscalar(X):
                Xz = float(X)
                plan.kcombine(output['x']._kout(), [(Yy, -Xz, -1)])
                output['x']._mark('kspace')
                plan.kcombine(output['y']._kout(), [(Yx, Xz, -1)])
                output['y']._mark('kspace')
            else:
                Xy, Xx = (float(v) for v in X)
                plan.kcombine(output._kout(), [(Yy, Xx, -1), (Yx, -Xy, -1)])
                output._mark('kspace')
        else:
            Xz, Xy, Xx = (float(v) for v in X)
            Yz = Y['z']._kdev()
            for name, terms in (('x', [(Yz, Xy, -1), (Yy, -Xz, -1)]),
                                ('y', [(Yx, Xz, -1), (Yz, -Xx, -1)]),
                                ('z', [(Yy, Xx, -1), (Yx, -Xy, -1)])):
                plan.kcombine(output[name]._kout(), terms)
                output[name]._mark('kspace')

    def XcrossY(self, X, Y, output):
        """X cross Y in x space (physics.py:309-354); X and Y move to x space."""
        plan = _plan(Y)
        if self.ndim == 2:
            if X.ncomp == 1:
                Xz = X._xdev()
            else:
                Xx = X['x']._xdev()
                Xy = X['y']._xdev()
        else:
            Xx = X['x']._xdev()
            Xy = X['y']._xdev()
            Xz = X['z']._xdev()
        Yx = Y['x']._xdev()
        Yy = Y['y']._xdev()
        if self.ndim == 3:
            Yz = Y['z']._xdev()
        if self.ndim == 2:
            if X.ncomp == 1:
                plan.xcombine(output['x']._xout(), [(Xz, Yy, -1.0)])
                output['x']._mark('xspace')
                plan.xcombine(output['y']._xout(), [(Xz, Yx, 1.0)])
                output['y']._mark('xspace')
            else:
                plan.xcombine(output._xout(), [(Xx, Yy, 1.0), (Xy, Yx, -1.0)])
                output._mark('xspace')
        else:
            for name, terms in (('x', [(Xy, Yz, 1.0), (Xz, Yy, -1.0)]),
                                ('y', [(Xz, Yx, 1.0), (Xx, Yz, -1.0)]),
                                ('z', [(Xx, Yy, 1.0), (Xy, Yx, -1.0)])):
                plan.xcombine(output[name]._xout(), terms)
                output[name]._mark('xspace')

    def XdotY(self, X, Y, output, space):
        """X.Y (physics.py:356-372); x space on the device."""
        if X.ncomp != Y.ncomp:
            raise ValueError('Vectors not the same size')
        if space != 'xspace':
            raise NotImplementedError("XdotY is provided in x space")
        plan = _plan(X)
        terms = [(X[i]._xdev(), Y[i]._xdev(), 1.0) for i in range(X.ncomp)]
        plan.xcombine(output._xout(), terms)
        output._mark('xspace')

    def curlX(self, X, output):
        """Curl in k space (physics.py:374-405)."""
        plan = _plan(X)
        N = X.ncomp
        if N == 3:
            X0, X1, X2 = X[0]._kdev(), X[1]._kdev(), X[2]._kdev()
            for o, terms in ((0, [(X2, 1.0, 1), (X1, -1.0, 2)]),
                             (1, [(X0, 1.0, 2), (X2, -1.0, 0)]),
                             (2, [(X1, 1.0, 0), (X0, -1.0, 1)])):
                plan.kcombine(output[o]._kout(), terms)
                output[o]._mark('kspace')
        elif N == 2:
            plan.kcombine(output._kout(), [(X[1]._kdev(), 1.0, 0), (X[0]._kdev(), -1.0, 1)])
            output._mark('kspace')
        elif N == 1:
            X0 = X[0]._kdev()
            plan.kcombine(output[0]._kout(), [(X0, 1.0, 1)])
            output[0]._mark('kspace')
            plan.kcombine(output[1]._kout(), [(X0, -1.0, 0)])
            output[1]._mark('kspace')

    @timer
    def laplace_solve(self, X, output):
        """laplace(output) = X, k = 0 divided by 1 (physics.py:407-416)."""
        plan = _plan(X)
        plan.kcombine(output._kout(), [(X._kdev(), -1.0, 3)])
        output._mark('kspace')

    # ------------------------------------------------------------------
    # fused device right-hand side
    # ------------------------------------------------------------------
    def _bdir(self):
        return 0

    def _params(self, coeffs=None):
        """ddx_params from the parameters; coeffs overrides (nu, kappa, eta, order)."""
        p = self.parameters
        nu = p.get('nu', 0.)
        kappa = p.get('kappa', 0.)
        eta = p.get('eta', 0.)
        order = p.get('viscosity_order', 1)
        if coeffs is not None:
            nu, kappa, eta, order = coeffs
        return _hip.make_params(self._phys_id, nu=nu, viscosity_order=order, kappa=kappa,
                                eta=eta, g=p.get('g', 1.), alpha_t=p.get('alpha_t', 1.),
                                beta=p.get('beta', 1.), bdir=self._bdir(),
                                rho0=p.get('rho0', 1.))

    def _state_comps(self, data):
        return [c for name, _ in self._field_list for i, c in data[name]]

    def _forcing(self, data):
        """Device buffers of the forcing terms, aligned with the state components."""
        funcs = self.forcing_functions
        if not funcs:
            return None
        comps = self._state_comps(data)
        plan = comps[0]._plan
        if self._force_bufs is None:
            self._force_bufs = [None] * len(comps)
        out = [None] * len(comps)
        kshape = plan.kref_shape
        entries = []
        if 'VelocityForcing' in funcs:
            entries += [(i, funcs['VelocityForcing'](data, i)) for i in self.dims]
        if 'ThermalForcing' in funcs:
            entries.append((self.ndim, funcs['ThermalForcing'](data)))
        for slot, val in entries:
            if self._force_bufs[slot] is None:
                self._force_bufs[slot] = plan.alloc_k()
            plan.upload_k(self._force_bufs[slot],
                          np.broadcast_to(np.asarray(val, dtype=np.complex128), kshape))
            out[slot] = self._force_bufs[slot]
        return out

    def _check_supported(self):
        if self.parameters.get('shear_rate', 0.):
            raise NotImplementedError("linear shear is not part of the MI355X path")
        if self.parameters.get('Omega', None) is not None:
            raise NotImplementedError("rotation is not part of the MI355X path")
        if self._tracer:
            raise NotImplementedError("the passive tracer is not part of the MI355X path")

    def _fused_rhs(self, data, deriv):
        self._check_supported()
        state = [c._kdev() for c in self._state_comps(data)]
        dcomps = self._state_comps(deriv)
        outs = [c._kout() for c in dcomps]
        forcing = self._forcing(data)
        dcomps[0]._plan.rhs(self._params(), state, outs, forcing)
        for c in dcomps:
            c._mark('kspace')


class IncompressibleHydro(Physics):
    """Homogeneous incompressible hydrodynamics (physics.py:419-610)."""

    _phys_id = _hip.HYDRO

    def __init__(self, *args, **kwargs):
        Physics.__init__(self, *args, **kwargs)
        self._field_list.append(('u', 'VectorField'))
        self._aux_field_list.append(('mathscalar', 'ScalarField'))
        self._aux_field_list.append(('mathvector', 'VectorField'))
        self.parameters['viscosity_order'] = 1
        self.parameters['nu'] = 0.
        self.parameters['shear_rate'] = 0.
        self.parameters['Omega'] = None
        if self._tracer:
            self._field_list.append(('c', 'ScalarField'))
            self.parameters['c_diff'] = 0.
        self._first_rhs = True

    def _finalize(self):
        Physics._finalize(self)
        self._shear = self.parameters['shear_rate'] != 0.
        self._rotation = self.parameters['Omega'] is not None

    def _setup_integrating_factors(self, deriv):
        nu = self.parameters['nu']
        vo = self.parameters['viscosity_order']
        for cindex, comp in deriv['u']:
            comp.integrating_factor = None if nu == 0. else IntegratingFactor(nu, vo, comp)

    def _coeffs(self, deriv):
        """(nu, kappa, eta, order) as captured by the integrating factors of deriv."""
        def coeff(field):
            IF = deriv[field][0].integrating_factor
            return (0., None) if IF is None else (IF.coeff, IF.order)
        nu, o1 = coeff('u')
        kappa, o2 = coeff('T') if 'T' in deriv.fields else (0., None)
        eta, o3 = coeff('B') if 'B' in deriv.fields else (0., None)
        orders = {o for o in (o1, o2, o3) if o is not None}
        if len(orders) > 1:
            raise NotImplementedError("one viscosity order for all fields")
        order = orders.pop() if orders else self.parameters.get('viscosity_order', 1)
        return nu, kappa, eta, order

    def set_velocity_forcing(self, func):
        self.forcing_functions['VelocityForcing'] = func

    def RHS(self, data, deriv):
        """RHS(u) = -P[(u.grad)u] (+ forcing), physics.py:527-586.

        The three built-in classes run the fused device right-hand side.  An
        object of any other class (a user subclass that inherits or calls
        this RHS) runs the reference's operator sequence, in which the
        pressure projection happens only when the object's class is exactly
        the one whose RHS projects (physics.py:575, 697, 805)."""
        if self._first_rhs:
            self._setup_integrating_factors(deriv)
            self._first_rhs = False
        # the fused right-hand side forms the nonlinear terms in rotational
        # form: equal to the reference's only for 2/3-dealiased products
        if type(self) in FUSED_CLASSES and _plan(data['u']).dealias_rule != _hip.DEALIAS_NONE:
            Physics.RHS(self, data, deriv, zero_outputs=False)
            self._fused_rhs(data, deriv)
        else:
            self._operator_rhs(data, deriv)

    def _operator_rhs(self, data, deriv):
        """The reference's RHS bodies operator by operator (physics.py:543-586,
        676-712, 782-815), each operator a device kernel."""
        self._check_supported()
        Physics.RHS(self, data, deriv)
        ms = self.aux_fields['mathscalar']
        mv = self.aux_fields['mathvector']
        plan = _plan(data['u'])
        forcing = self._forcing(data)

        def scaled(out, terms, accumulate=False):
            plan.kcombine(out._kdev() if accumulate else out._kout(), terms, accumulate=accumulate)
            out._mark('kspace')

        # inertial term (physics.py:554-556), forcing (:570-572)
        self.XgradY(data['u'], data['u'], ms, mv, deriv['u'])
        for i in self.dims:
            c = deriv['u'][i]
            scaled(c, [(c._kdev(), -1.0, -1)])
            if forcing is not None and forcing[i] is not None:
                scaled(c, [(forcing[i], 1.0, -1)], accumulate=True)
        if type(self) is IncompressibleHydro:
            self.pressure_projection(data, deriv)
        if isinstance(self, BoussinesqHydro):
            p = self.parameters
            d = self._bdir()
            # buoyancy (:690-694), projection (:697-698)
            scaled(deriv['u'][d], [(data['T'][0]._kdev(), p['g'] * p['alpha_t'], -1)], accumulate=True)
            if type(self) is BoussinesqHydro:
                self.pressure_projection(data, deriv)
            # temperature advection and stratification (:700-708), forcing (:710-712)
            T = deriv['T'][0]
            self.XgradY(data['u'], data['T'], ms, mv, deriv['T'])
            scaled(T, [(T._kdev(), -1.0, -1), (data['u'][d]._kdev(), -p['beta'], -1)])
            if forcing is not None and forcing[self.ndim] is not None:
                scaled(T, [(forcing[self.ndim], 1.0, -1)], accumulate=True)
        if isinstance(self, IncompressibleMHD):
            mv2 = self.aux_fields['mathvector2']
            fpr = 4 * np.pi * self.parameters['rho0']
            # Lorentz force (:794-802), projection (:804-806)
            cur = ms if self.ndim == 2 else mv
            self.curlX(data['B'], cur)
            self.XcrossY(cur, data['B'], mv2)
            for i in self.dims:
                scaled(deriv['u'][i], [(mv2[i]._kdev(), 1.0 / fpr, -1)], accumulate=True)
            if type(self) is IncompressibleMHD:
                self.pressure_projection(data, deriv)
            # induction (:808-815)
            e = ms if self.ndim == 2 else mv
            self.XcrossY(data['u'], data['B'], e)
            self.curlX(e, deriv['B'])

    def pressure_projection(self, data, deriv):
        """Solenoidal projection of deriv['u'] (physics.py:588-599)."""
        if self.parameters.get('shear_rate', 0.):
            raise NotImplementedError("linear shear is not part of the MI355X path")
        bufs = [c._kdev() for i, c in deriv['u']]
        deriv['u'][0]._plan.project(bufs)
        for i, c in deriv['u']:
            c._mark('kspace')

    def max_abs_vel(self, data):
        return np.sqrt(data['u'].max_square())

    def set_dtlist(self, data):
        dx = data['u']['x'].dx().min()
        self.dtlist.append(dx / self.max_abs_vel(data))


class BoussinesqHydro(IncompressibleHydro):
    """Boussinesq hydrodynamics (physics.py:612-721)."""

    _phys_id = _hip.BOUSSINESQ

    def __init__(self, *args, **kwargs):
        IncompressibleHydro.__init__(self, *args, **kwargs)
        self._field_list.append(('T', 'ScalarField'))
        self.parameters['kappa'] = 0.
        self.parameters['g'] = 1.
        self.parameters['alpha_t'] = 1.
        self.parameters['beta'] = 1.
        self.parameters['boussinesq_direction'] = decfg.get('physics', 'boussinesq_direction')

    def _bdir(self):
        d = self.parameters['boussinesq_direction']
        table = {'x': 0, 'y': 1} if self.ndim == 2 else {'x': 0, 'y': 1, 'z': 2}
        return table[d]

    def _setup_integrating_factors(self, deriv):
        IncompressibleHydro._setup_integrating_factors(self, deriv)
        kappa = self.parameters['kappa']
        vo = self.parameters['viscosity_order']
        c = deriv['T'][0]
        c.integrating_factor = None if kappa == 0. else IntegratingFactor(kappa, vo, c)

    def set_thermal_forcing(self, func):
        self.forcing_functions['ThermalForcing'] = func

    def RHS(self, data, deriv):
        """RHS(u) += g alpha_t T; RHS(T) = -u.grad T - beta u_dir (physics.py:664-712)."""
        IncompressibleHydro.RHS(self, data, deriv)

    def set_dtlist(self, data):
        IncompressibleHydro.set_dtlist(self, data)
        dx = data['u']['x'].dx().min()
        N = np.sqrt(np.abs(self.parameters['alpha_t'] * self.parameters['g'] * self.parameters['beta']))
        self.dtlist.append(dx / N)


class IncompressibleMHD(IncompressibleHydro):
    """Homogeneous incompressible MHD (physics.py:724-836)."""

    _phys_id = _hip.MHD

    def __init__(self, *args, **kwargs):
        IncompressibleHydro.__init__(self, *args, **kwargs)
        self._field_list.append(('B', 'VectorField'))
        self._aux_field_list.append(('mathvector2', 'VectorField'))
        self.parameters['rho0'] = 1.
        self.parameters['eta'] = 0.

    def _setup_integrating_factors(self, deriv):
        IncompressibleHydro._setup_integrating_factors(self, deriv)
        eta = self.parameters['eta']
        vo = self.parameters['viscosity_order']
        for cindex, comp in deriv['B']:
            comp.integrating_factor = None if eta == 0. else IntegratingFactor(eta, vo, comp)

    def RHS(self, data, deriv):
        """RHS(u) += curl(B) x B / (4 pi rho0); RHS(B) = curl(u x B) (physics.py:770-819)."""
        IncompressibleHydro.RHS(self, data, deriv)

    def max_alfven_speed(self, data):
        fpr = 4 * np.pi * self.parameters['rho0']
        return np.sqrt(data['B'].max_square() / fpr)

    def set_dtlist(self, data):
        IncompressibleHydro.set_dtlist(self, data)
        dx = data['u']['x'].dx().min()
        self.dtlist.append(dx / self.max_alfven_speed(data))


FUSED_CLASSES = (IncompressibleHydro, BoussinesqHydro, IncompressibleMHD)
