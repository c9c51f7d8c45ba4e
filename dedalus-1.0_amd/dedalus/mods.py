"""``from dedalus.mods import *``: the reference's public names (dedalus/mods.py:24-101)."""

import os
import signal

import dedalus.startup_tasks as __startup_tasks  # noqa: F401
from dedalus.config import decfg
from dedalus.utils.logger import mylog
from dedalus.analysis.api import (AnalysisSet, VolumeAverageSet, Snapshot, TrackMode,
                                  VolumeAverage, PowerSpectrum)
from dedalus.data_objects.api import (FourierRepresentation, FourierShearRepresentation,
                                      ChebyshevRepresentation)
from dedalus.init_cond.api import (taylor_green, sin_k, cos_k, turb, turb_new, mcwilliams_spec,
                                   MIT_vortices, vorticity_wave, alfven, kida_vortex,
                                   add_gaussian_white_noise, constant)
from dedalus.physics.api import IncompressibleHydro, BoussinesqHydro, IncompressibleMHD
from dedalus.time_stepping.api import RK2mid, RK2trap, RK4, CrankNicholsonVisc
from dedalus.utils.api import (restart, Timer, integrate_quad, integrate_simp, interp_linear,
                               find_zero, load_all, com_sys, swap_indices, identify_version)
from dedalus.funcs import signal_print_traceback

try:
    signal.signal(signal.SIGHUP, signal_print_traceback)
except (ValueError, AttributeError):
    pass

if decfg.getboolean("utils", "loadplugins"):
    _fn = decfg.get("utils", "pluginfilename")
    if not os.path.isfile(_fn):
        _fn = os.path.expanduser("~/.dedalus/%s" % _fn)
    if os.path.isfile(_fn):
        mylog.info("Loading plugins from %s", _fn)
        with open(_fn) as _f:
            exec(compile(_f.read(), _fn, 'exec'))
