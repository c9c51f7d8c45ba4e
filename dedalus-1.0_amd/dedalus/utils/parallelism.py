"""Communication system and reductions (dedalus/utils/parallelism.py:29-173).

This package runs one process on one GPU, so ``com_sys`` is the reference's
serial fallback (nproc = 1, comm = None) and the reductions reduce locally.
"""

import numpy as np


class CommunicationSystem:
    comm = None
    MPI = None

    def __init__(self):
        self.myproc = 0
        self.nproc = 1


com_sys = CommunicationSystem()


def reduce_mean(data):
    return data.mean()


def reduce_sum(data, reduce_all=False):
    return data.sum()


def reduce_min(data, reduce_all=False):
    return data.min()


def reduce_max(data, reduce_all=False):
    return data.max()


def swap_indices(arr):
    """Swap entries 0 and 1 of an array or list (parallelism.py:155-173)."""
    if isinstance(arr, np.ndarray):
        out = arr.copy()
    elif isinstance(arr, (list, tuple)):
        out = list(arr)
    else:
        raise NotImplementedError("swap_indices only implemented for numpy arrays and lists.")
    a = out[1]
    out[1] = out[0]
    out[0] = a
    return out


def load_all(field, snap_dir):
    """Single-process snapshot loader (parallelism.py:54-100 without MPI concat)."""
    import glob
    import os
    files = sorted(glob.glob(os.path.join(snap_dir, "data.cpu*.npz")))
    if not files:
        raise IOError("no snapshot data in %s" % snap_dir)
    if not field.startswith('fields/'):
        field = 'fields/' + field.lstrip('/')
    with np.load(files[0], allow_pickle=False) as f:
        return f[field], str(f[field + '.space'])
