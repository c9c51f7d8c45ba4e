"""Volume-averaged diagnostics (dedalus/analysis/volume_average.py:25-300).

k-space averages use Parseval with Hermitian weights (1 on the kx = 0
plane, 2 elsewhere, volume_average.py:71-106); the energy and enstrophy
sums run as device reductions.
"""

import numpy as np

from dedalus.utils.parallelism import com_sys, reduce_max, reduce_mean, reduce_sum


class VolumeAverageSet:

    known_analysis = {}

    def __init__(self, data, filename='time_series.dat'):
        self.data = data
        self.filename = filename
        self.tasks = []
        self.scratch = self.data.clone()
        self.scratch.add_field('scalar', 'ScalarField')
        self.scratch.add_field('vector', 'VectorField')
        if com_sys.myproc == 0:
            self.outfile = open(self.filename, 'a')
            self.outfile.write("# Dedalus Volume Average\n")
            self.outfile.write("# Column 0: time\n")

    def add(self, name, fmt, options=None):
        self.tasks.append((self.known_analysis[name], fmt, options or {}))
        if com_sys.myproc == 0:
            self.outfile.write("# Column %i: %s\n" % (len(self.tasks), name))

    def run(self):
        line = ["%10.5f" % self.data.time]
        for f, fmt, kwargs in self.tasks:
            line.append(fmt % f(self.data, self.scratch, **kwargs))
        if com_sys.myproc == 0:
            self.outfile.write("\t".join(line) + "\n")
            self.outfile.flush()

    @classmethod
    def register_task(cls, func):
        cls.known_analysis[func.__name__] = func
        return func


def volume_average(data, kdict=None, space='kspace', reduce_all=False):
    """Mean over the volume from k-space (Hermitian weights) or x-space data."""
    if space == 'kspace':
        if hasattr(data, 'k') and hasattr(data, 'require_space'):
            k = data.k
            values = data['kspace']
        else:
            k = kdict
            values = data
            if k is None:
                raise ValueError("volume_average: pass k vectors via kdict for plain arrays")
        nd = values.ndim
        if nd == 3:
            if k['x'][..., 0] == 0:
                s = 2 * values[..., 1:].sum() + values[..., 0].sum()
            else:
                s = 2 * values.sum()
        else:
            if k['x'][0, ...] == 0:
                s = 2 * values[1:, ...].sum() + values[0, ...].sum()
            else:
                s = 2 * values.sum()
        return reduce_sum(np.asarray(s), reduce_all=reduce_all)
    if space == 'xspace':
        values = data if isinstance(data, np.ndarray) else data['xspace']
        return reduce_mean(values)
    raise ValueError("volume_average: must be either xspace or kspace")


def _ksum_sq(field):
    """sum over components of <|c|^2> (device reduction)."""
    return sum(c._plan.kreduce(c._kdev(), 0) for i, c in field)


@VolumeAverageSet.register_task
def ekin(data, scratch, space='kspace'):
    if space == 'kspace':
        return 0.5 * _ksum_sq(data['u'])
    return sum(0.5 * c._plan.xreduce(c._xdev(), 2) for i, c in data['u']) / data['u'][0]._plan.npts


@VolumeAverageSet.register_task
def emag(data, scratch, space='kspace'):
    if space == 'kspace':
        return 0.5 * _ksum_sq(data['B'])
    return sum(0.5 * c._plan.xreduce(c._xdev(), 2) for i, c in data['B']) / data['B'][0]._plan.npts


@VolumeAverageSet.register_task
def enstrophy(data, scratch, space='kspace'):
    """0.5 <w_z^2>, 2D (volume_average.py:191-200)."""
    u = data['u']
    return 0.5 * u[0]._plan.kreduce(u[0]._kdev(), 1, u[1]._kdev())


def _comp_sq(comp):
    return comp._plan.kreduce(comp._kdev(), 0)


@VolumeAverageSet.register_task
def ux2(data, scratch, space='kspace'):
    return _comp_sq(data['u']['x'])


@VolumeAverageSet.register_task
def uy2(data, scratch, space='kspace'):
    return _comp_sq(data['u']['y'])


@VolumeAverageSet.register_task
def uz2(data, scratch, space='kspace'):
    return _comp_sq(data['u']['z'])


@VolumeAverageSet.register_task
def bx2(data, scratch, space='kspace'):
    return _comp_sq(data['B']['x'])


@VolumeAverageSet.register_task
def by2(data, scratch, space='kspace'):
    return _comp_sq(data['B']['y'])


@VolumeAverageSet.register_task
def bz2(data, scratch, space='kspace'):
    return _comp_sq(data['B']['z'])


@VolumeAverageSet.register_task
def temp2(data, scratch, space='kspace'):
    return _comp_sq(data['T'][0])


@VolumeAverageSet.register_task
def comp_mean(data, scratch, fname, cindex):
    c = data[fname][cindex]
    return c._plan.xreduce(c._xdev(), 1) / c._plan.npts


def _xmean(comp):
    return comp._plan.xreduce(comp._xdev(), 1) / comp._plan.npts


for _name, _f, _c in (('uxm', 'u', 'x'), ('uym', 'u', 'y'), ('uzm', 'u', 'z'),
                      ('bxm', 'B', 'x'), ('bym', 'B', 'y'), ('bzm', 'B', 'z')):
    def _mk(f=_f, c=_c):
        def task(data, scratch, space='kspace'):
            return _xmean(data[f][c])
        return task
    _t = _mk()
    _t.__name__ = _name
    VolumeAverageSet.register_task(_t)


for _name, _f, _c in (('ux_max', 'u', 'x'), ('uy_max', 'u', 'y'), ('uz_max', 'u', 'z'),
                      ('bx_max', 'B', 'x'), ('by_max', 'B', 'y'), ('bz_max', 'B', 'z')):
    def _mk(f=_f, c=_c):
        def task(data, scratch):
            return reduce_max(data[f][c]['xspace'])
        return task
    _t = _mk()
    _t.__name__ = _name
    VolumeAverageSet.register_task(_t)


@VolumeAverageSet.register_task
def vort_cenk(data, scratch):
    """Centroid wavenumber of McWilliams (1990) (volume_average.py:202-212)."""
    k2 = data['u']['x'].k2(no_zero=True)
    en = np.zeros(k2.shape)
    for i in range(data['u'].ndim):
        a = data['u'][i]['kspace']
        en += 0.5 * (a * a.conj()).real
    en.flat[0] = 0.
    return (k2 ** 1.5 * en).sum() / (k2 * en).sum()


@VolumeAverageSet.register_task
def energy_dissipation(data, scratch):
    """2 nu <|curl u|^2> / 2 (volume_average.py:254-264)."""
    u = data['u']
    d = {}
    for c in 'xyz'[:u.ndim]:
        for dd in 'xyz'[:u.ndim]:
            d[c + dd] = u[c].deriv(dd).copy()
    if u.ndim == 3:
        w2 = (np.abs(d['zy'] - d['yz']) ** 2 + np.abs(d['xz'] - d['zx']) ** 2
              + np.abs(d['yx'] - d['xy']) ** 2)
    else:
        w2 = np.abs(d['yx'] - d['xy']) ** 2
    en_dis = 2 * data.parameters['nu'] * (0.5 * w2).real
    return volume_average(en_dis, kdict=scratch['scalar'].k)


@VolumeAverageSet.register_task
def thermal_energy_dissipation(data, scratch):
    """kappa <|grad T|^2> (volume_average.py:266-275)."""
    T = data['T']
    s = sum(np.abs(T.deriv(d)) ** 2 for d in 'xyz'[:T.ndim])
    return volume_average(data.parameters['kappa'] * s, kdict=T.k)


@VolumeAverageSet.register_task
def divergence(data, scratch):
    """<div u> (volume_average.py:277-284)."""
    u = data['u']
    plan = u[0]._plan
    out = scratch['scalar']
    plan.kcombine(out._kout(), [(u[i]._kdev(), 1.0, i) for i in range(u.ndim)])
    out._mark('kspace')
    return volume_average(out)


@VolumeAverageSet.register_task
def divergence_sum(data, scratch):
    """sum |i k.u| over the k array (volume_average.py:286-294)."""
    u = data['u']
    plan = u[0]._plan
    out = scratch['scalar']
    plan.kcombine(out._kout(), [(u[i]._kdev(), 1.0, i) for i in range(u.ndim)])
    out._mark('kspace')
    return float(np.abs(out['kspace']).sum())
