"""Restart from a snapshot directory (dedalus/utils/restart.py:33-99).

Snapshots here hold the fields (npz) and a JSON description instead of
pickled objects, so a restart loads the fields into an existing state.
"""

import glob
import json
import os

import numpy as np


def load_snapshot(snap_dir):
    """(description dict, {'fields/<name>/<comp>': (array, space)}, time)."""
    meta = {}
    js = sorted(glob.glob(os.path.join(snap_dir, 'dedalus_obj_*.json')))
    if js:
        with open(js[0]) as f:
            meta = json.load(f)
    files = sorted(glob.glob(os.path.join(snap_dir, 'data.cpu*.npz')))
    if not files:
        raise IOError("no snapshot data in %s" % snap_dir)
    out = {}
    with np.load(files[0], allow_pickle=False) as f:
        t = float(f['time'])
        for key in f.files:
            if key.startswith('fields/') and key.count('/') == 2 and '.' not in key.split('/')[-1]:
                out[key] = (f[key], str(f[key + '.space']))
    return meta, out, t


def restart_into(data, snap_dir):
    """Fill the fields of ``data`` from a snapshot and set its time."""
    meta, fields, t = load_snapshot(snap_dir)
    for key, (arr, space) in fields.items():
        _, name, comp = key.split('/')
        if name in data.fields:
            data[name][int(comp)][space] = arr
    data.set_time(t)
    return meta


def restart(snap_dir):
    raise NotImplementedError("snapshots hold fields and a JSON description; rebuild the "
                              "physics and state, then call restart_into(data, snap_dir)")
