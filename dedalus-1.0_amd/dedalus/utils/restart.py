"""Restart from a snapshot directory (dedalus/utils/restart.py:33-99).

A snapshot (``TimeStepBase.snapshot``) holds the fields in an npz whose keys
follow the reference's HDF5 layout (``fields/<name>/<comp>`` plus a
``.space`` entry), a JSON description of the run in place of the pickled
(RHS, data, ti) objects, and the source of the forcing functions.
``restart`` rebuilds the three objects from them, rebinds the forcing
functions by name from the snapshot's ``forcing_functions.py`` (as
restart.py:46-59 does) and loads the fields (restart.py:71-99).
"""

import glob
import importlib.util
import inspect
import json
import os
import time

import numpy as np

from dedalus.utils.parallelism import com_sys


def load_snapshot(snap_dir):
    """(description dict, {'fields/<name>/<comp>': (array, space)}, time)."""
    meta = {}
    js = os.path.join(snap_dir, 'dedalus_obj_%04i.json' % com_sys.myproc)
    if not os.path.exists(js):
        cands = sorted(glob.glob(os.path.join(snap_dir, 'dedalus_obj_*.json')))
        js = cands[0] if cands else None
    if js:
        with open(js) as f:
            meta = json.load(f)
    data_file = os.path.join(snap_dir, 'data.cpu%04i.npz' % com_sys.myproc)
    if not os.path.exists(data_file):
        files = sorted(glob.glob(os.path.join(snap_dir, 'data.cpu*.npz')))
        if not files:
            raise IOError("no snapshot data in %s" % snap_dir)
        data_file = files[0]
    out = {}
    with np.load(data_file, allow_pickle=False) as f:
        t = float(f['time'])
        for key in f.files:
            if key.startswith('fields/') and key.count('/') == 2 and '.' not in key.split('/')[-1]:
                out[key] = (f[key], str(f[key + '.space']))
    return meta, out, t


def restart_into(data, snap_dir):
    """Fill the fields of ``data`` from a snapshot and set its time
    (load_all_data, restart.py:71-99: each component in the space it was
    saved in)."""
    meta, fields, t = load_snapshot(snap_dir)
    for name in data.fields:
        for comp in range(data[name].ncomp):
            key = 'fields/%s/%i' % (name, comp)
            if key not in fields:
                raise KeyError("Data File field %s component %i is missing or corrupt" % (name, comp))
            arr, space = fields[key]
            data[name][comp][space] = arr
    data.set_time(t)
    return meta


def _forcing_module(snap_dir):
    fn = os.path.join(snap_dir, 'forcing_functions.py')
    if not os.path.exists(fn):
        return None
    spec = importlib.util.spec_from_file_location('forcing_functions', fn)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def restart(snap_dir):
    """Restart a run from the snapshot in snap_dir (restart.py:33-69).

    Returns (RHS, data, ti): the physics object with the saved parameters
    and its forcing functions rebound by name, the state at the saved time,
    and the integrator with the saved clock, snapshot counter and stopping
    controls.  The integrating factors are set up again at the first
    right-hand side from the saved parameters, as on the original run."""
    from dedalus.data_objects.representations import FourierRepresentation
    from dedalus.physics import physics as phys_mod
    from dedalus.time_stepping import time_step as ts_mod

    meta, _, t = load_snapshot(snap_dir)
    if not meta:
        raise IOError("no run description (dedalus_obj_*.json) in %s" % snap_dir)
    cls = getattr(phys_mod, meta['physics'], None)
    if cls is None or not isinstance(cls, type):
        raise ValueError("unknown physics class %r in the snapshot" % meta['physics'])
    RHS = cls(tuple(meta['shape']), FourierRepresentation, tuple(meta['length']))
    for k, v in meta['parameters'].items():
        RHS.parameters[k] = np.array(v) if isinstance(v, list) else v
    RHS._setup_aux_fields(t, RHS._aux_field_list)

    names = meta.get('forcing_function_names', {})
    if names:
        mod = _forcing_module(snap_dir)
        for k, fname in names.items():
            if mod is not None:
                for ffn, ffunc in inspect.getmembers(mod):
                    if ffn == fname:
                        RHS.forcing_functions[k] = ffunc
            if k not in RHS.forcing_functions:
                raise ValueError("Failed to load driver function for %s. Snapshot may be corrupted." % k)
            RHS._forcing_function_names[k] = fname

    data = RHS.create_fields(t)
    restart_into(data, snap_dir)

    tcls = getattr(ts_mod, meta['integrator'], None)
    if tcls is None or not isinstance(tcls, type):
        raise ValueError("unknown integrator %r in the snapshot" % meta['integrator'])
    ti = tcls(RHS)
    st = meta.get('integrator_state', {})
    ti.time = meta['time']
    ti.iteration = meta['iteration']
    for attr in ('CFL', 'sim_stop_time', 'wall_stop_time', 'stop_iteration', 'save_cadence',
                 'max_save_period'):
        if attr in st:
            setattr(ti, attr, st[attr])
    ti._nsnap = st.get('nsnap', ti._nsnap)
    ti._tlastsnap = st.get('tlastsnap', ti._tlastsnap)
    ti.dt_old = st.get('dt_old', ti.dt_old)
    ti._start_time = time.time()
    return RHS, data, ti
