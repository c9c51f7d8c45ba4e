"""Time integrators (dedalus/time_stepping/time_step.py:48-506), on the MI355X.

``RK2mid`` and ``RK2trap`` keep the reference's storage (data2 / deriv1 /
deriv2), integrating-factor semantics (taken from deriv1, set at the first
right-hand side) and statistics.  A step of one of the built-in physics
classes without Python forcing is one call of the fused device step
(``ddx_step``: both right-hand sides and both ETD sweeps); otherwise each
stage calls ``RHS.RHS`` and the device ETD kernel (``ddx_update``) per
component, as the reference calls its Cython kernels.  ``advance`` returns
after the device has finished the step, so ``timer.timers['advance']`` is
wall time of completed work.
"""

import inspect
import json
import os
import time

import numpy as np

from dedalus import _hip
from dedalus.data_objects.representations import IntegratingFactor
from dedalus.physics.physics import FUSED_CLASSES
from dedalus.utils.logger import mylog
from dedalus.utils.parallelism import com_sys
from dedalus.utils.timer import timer

hg_version = "mi355x"


class TimeStepBase:
    """Stopping controls, statistics, snapshots (time_step.py:48-179)."""

    timer = timer

    def __init__(self, RHS, CFL=0.1, int_factor=None):
        self.RHS = RHS
        self.CFL = CFL
        self.int_factor = int_factor
        self.sim_stop_time = 1000
        self.wall_stop_time = 60. * 60.
        self.stop_iteration = 1000
        self.save_cadence = 100
        self.max_save_period = 100.
        self.time = 0
        self.iteration = 0
        self._nsnap = 0
        self._tlastsnap = 0.
        self.dt_old = np.finfo('d').max / 10.
        self._start_time = time.time()

    @property
    def ok(self):
        if self.iteration >= self.stop_iteration:
            if com_sys.myproc == 0:
                mylog.info("Timestepping complete: stop iteration reached.")
            return False
        if self.time >= self.sim_stop_time:
            if com_sys.myproc == 0:
                mylog.info("Timestepping complete: simulation stop time reached.")
            return False
        if (time.time() - self._start_time) >= self.wall_stop_time:
            if com_sys.myproc == 0:
                mylog.info("Timestepping complete: wall stop time reached.")
            return False
        return True

    @timer
    def advance(self, data, dt=None):
        if (self.iteration % self.save_cadence) == 0 or \
                (self.time - self._tlastsnap >= self.max_save_period):
            self.snapshot(data)
        if dt is None:
            dt = self.cfl_dt(data)
        self.do_advance(data, dt)
        mylog.info("step %i" % self.iteration)

    def do_advance(self, data, dt):
        raise NotImplementedError("do_advance must be provided by subclass.")

    @timer
    def snapshot(self, data):
        """snap_%05i/: forcing sources, a JSON description and the fields (npz).

        Layout of time_step.py:112-151 with the fields in an npz
        (``fields/<name>/<comp>`` plus ``.space``) instead of HDF5, and a
        JSON description of the run instead of pickled objects.
        """
        pathname = "snap_%05i" % self._nsnap
        os.makedirs(pathname, exist_ok=True)
        source = ""
        for k, forcer in self.RHS.forcing_functions.items():
            if forcer:
                try:
                    source += inspect.getsource(forcer) + "\n"
                except (OSError, TypeError):
                    pass
                self.RHS._forcing_function_names[k] = getattr(forcer, '__name__', str(forcer))
        with open(os.path.join(pathname, 'forcing_functions.py'), 'w') as f:
            f.write(source)
        params = {}
        for k, v in self.RHS.parameters.items():
            params[k] = v.tolist() if isinstance(v, np.ndarray) else v
        # what restart() needs to rebuild the objects the reference pickles
        # (restart.py:40-69): classes, parameters, forcing-function names and
        # the integrator's clock and stopping controls
        integ = {'CFL': self.CFL, 'sim_stop_time': self.sim_stop_time,
                 'wall_stop_time': self.wall_stop_time, 'stop_iteration': self.stop_iteration,
                 'save_cadence': self.save_cadence, 'max_save_period': self.max_save_period,
                 'nsnap': self._nsnap + 1, 'tlastsnap': self.time,
                 'dt_old': float(self.dt_old)}
        meta = {'physics': self.RHS.__class__.__name__, 'shape': list(self.RHS.shape),
                'length': list(self.RHS.length), 'parameters': params,
                'forcing_function_names': dict(self.RHS._forcing_function_names),
                'integrator': self.__class__.__name__, 'time': self.time,
                'iteration': self.iteration, 'integrator_state': integ,
                'hg_version': hg_version}
        with open(os.path.join(pathname, 'dedalus_obj_%04i.json' % com_sys.myproc), 'w') as f:
            json.dump(meta, f, default=str)
        group = {'time': np.array(self.time)}
        data.snapshot(group)
        np.savez(os.path.join(pathname, "data.cpu%04i.npz" % com_sys.myproc), **group)
        self._nsnap += 1
        self._tlastsnap = self.time

    def final_stats(self):
        self.timer.print_stats()
        if com_sys.myproc == 0:
            total_wtime = self.timer.timers.get('advance', 0.)
            print("total advance wall time: %10.5e sec" % total_wtime)
            print("%10.5e sec/step " % (total_wtime / max(self.iteration, 1)))
            print()
            print("Simulation complete. Status: awesome")

    def finalize(self, data):
        self.snapshot(data)
        self.final_stats()

    def cfl_dt(self, data):
        dt = self.CFL * self.RHS.compute_dt(data)
        if dt > 1.05 * self.dt_old:
            dt = 1.05 * self.dt_old
        self.dt_old = dt
        mylog.info("dt = %10.5e" % dt)
        return dt


def _if_args(IF, plan):
    """How a sweep applies an integrating factor on the device: None for
    None, ('coeff', coeff, order) for coeff*|k|^(2 order), which the kernels
    recompute from the wavenumbers, and ('array', buffer) for any other array
    (a per-mode factor in the reference k layout, uploaded in compact form)."""
    if IF is None:
        return None
    if isinstance(IF, IntegratingFactor):
        return ('coeff', IF.coeff, IF.order)
    return ('array', plan.upload_real_k(np.asarray(IF, dtype=np.float64)))


def _sweep(plan, kind, IF, y, out, F1, F2, dt):
    """One ETD sweep of a component with its integrating factor (kind is
    the ETD form; Euler when IF is None, as time_step.py:290-292 do)."""
    a = _if_args(IF, plan)
    if a is None:
        if kind == _hip.ETD2RK1:
            plan.update(_hip.ETD2RK1, y, out, F1, F2, 0., 1, dt)
        else:
            plan.update(_hip.EULER, y, out, F2 if kind == _hip.ETD2RK2 else F1, None, 0., 1, dt)
    elif a[0] == 'coeff':
        plan.update(kind, y, out, F1, F2, a[1], a[2], dt)
    else:
        plan.update_if(kind, y, out, F1, F2, a[1], dt)
        plan.sync()  # the uploaded factor is freed on return


class RKBase(TimeStepBase):
    """Base class for the Runge-Kutta integrators (time_step.py:182-221)."""

    def _fusable(self):
        ph = self.RHS
        # exactly a built-in class: a subclass may rely on the reference's
        # class-dependent projection (physics.py:575, 697, 805); and every
        # integrating factor of the form the fused step recomputes
        return (type(ph) in FUSED_CLASSES and not ph.forcing_functions
                and ph._phys_id is not None
                and self._comps(self.deriv1)[0]._plan.dealias_rule != _hip.DEALIAS_NONE
                and all(isinstance(getattr(c, 'integrating_factor', None), (IntegratingFactor, type(None)))
                        for c in self._comps(self.deriv1)))

    def _comps(self, sd):
        return self.RHS._state_comps(sd)

    def _plan(self, data):
        return self._comps(data)[0]._plan

    def _prepare_fused(self, data):
        ph = self.RHS
        if ph._first_rhs:
            ph._setup_integrating_factors(self.deriv1)
            ph._first_rhs = False
        if not ph._is_finalized:
            ph._finalize()
        ph._check_supported()
        return ph._params(ph._coeffs(self.deriv1))


class RK2mid(RKBase):
    """ETD midpoint RK2 (time_step.py:224-309)."""

    def __init__(self, *arg, **kwargs):
        TimeStepBase.__init__(self, *arg, **kwargs)
        self.data2 = self.RHS.create_fields(0.)
        self.deriv1 = self.RHS.create_fields(0.)
        self.deriv2 = self.RHS.create_fields(0.)

    def do_advance(self, data, dt):
        if self._fusable():
            prm = self._prepare_fused(data)
            state = [c._kdev() for c in self._comps(data)]
            groups = [self._comps(sd) for sd in (self.data2, self.deriv1, self.deriv2)]
            bufs = [[c._kout() for c in g] for g in groups]
            plan = self._plan(data)
            plan.step(_hip.RK2MID, prm, state, bufs[0], bufs[1], bufs[2], 1, dt)
            for g in groups:
                for c in g:
                    c._mark('kspace')
            for c in self._comps(data):
                c._mark('kspace')
            self.deriv1.set_time(data.time)
            self.data2.set_time(data.time + dt / 2.)
            self.deriv2.set_time(data.time + dt / 2.)
        else:
            plan = self._plan(data)
            self.RHS.RHS(data, self.deriv1)
            self._sweep_stage1(data, dt)
            self.data2.set_time(data.time + dt / 2.)
            self.RHS.RHS(self.data2, self.deriv2)
            self._sweep_stage2(data, dt)
        data.set_time(data.time + dt)
        self.time += dt
        self.iteration += 1
        self._plan(data).sync()

    def _sweep_stage1(self, data, dt):
        plan = self._plan(data)
        for y, a2, f1 in zip(self._comps(data), self._comps(self.data2), self._comps(self.deriv1)):
            _sweep(plan, _hip.ETD1, f1.integrating_factor, y._kdev(), a2._kout(), f1._kdev(), None,
                   dt / 2.)
            a2._mark('kspace')

    def _sweep_stage2(self, data, dt):
        plan = self._plan(data)
        for y, f1, f2 in zip(self._comps(data), self._comps(self.deriv1), self._comps(self.deriv2)):
            yb = y._kdev()
            _sweep(plan, _hip.ETD2RK2, f1.integrating_factor, yb, yb, f1._kdev(), f2._kdev(), dt)
            y._mark('kspace')


class RK2trap(RKBase):
    """ETD trapezoidal RK2 (time_step.py:312-392)."""

    def __init__(self, *arg, **kwargs):
        TimeStepBase.__init__(self, *arg, **kwargs)
        self.deriv1 = self.RHS.create_fields(0.)
        self.deriv2 = self.RHS.create_fields(0.)

    def do_advance(self, data, dt):
        if self._fusable():
            prm = self._prepare_fused(data)
            state = [c._kdev() for c in self._comps(data)]
            groups = [self._comps(sd) for sd in (self.deriv1, self.deriv2)]
            bufs = [[c._kout() for c in g] for g in groups]
            plan = self._plan(data)
            plan.step(_hip.RK2TRAP, prm, state, None, bufs[0], bufs[1], 1, dt)
            for g in groups:
                for c in g:
                    c._mark('kspace')
            for c in self._comps(data):
                c._mark('kspace')
            self.deriv1.set_time(data.time)
            self.deriv2.set_time(data.time + dt)
            data.set_time(data.time + dt)
        else:
            plan = self._plan(data)
            self.RHS.RHS(data, self.deriv1)
            for y, f1 in zip(self._comps(data), self._comps(self.deriv1)):
                yb = y._kdev()
                _sweep(plan, _hip.ETD1, f1.integrating_factor, yb, yb, f1._kdev(), None, dt)
                y._mark('kspace')
            data.set_time(data.time + dt)
            self.RHS.RHS(data, self.deriv2)
            for y, f1, f2 in zip(self._comps(data), self._comps(self.deriv1),
                                 self._comps(self.deriv2)):
                yb = y._kdev()
                # Euler on F2 - F1 with dt/2 is ETD2RK1 at Z = 0
                _sweep(plan, _hip.ETD2RK1, f1.integrating_factor, yb, yb, f1._kdev(), f2._kdev(), dt)
                y._mark('kspace')
        self.time += dt
        self.iteration += 1
        self._plan(data).sync()


class RK4(RKBase):
    """Classical RK4: broken in dedalus-1.0 (time_step.py:209, 214 call undefined methods)."""

    def __init__(self, *args, **kwargs):
        raise NotImplementedError("RK4 is broken in dedalus-1.0 and not provided")


class CrankNicholsonVisc(RKBase):
    """Broken in dedalus-1.0 (time_step.py:495 calls RHS.RHS with one argument)."""

    def __init__(self, *args, **kwargs):
        raise NotImplementedError("CrankNicholsonVisc is broken in dedalus-1.0 and not provided")
