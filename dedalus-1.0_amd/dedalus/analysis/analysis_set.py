"""Cadenced analysis tasks (dedalus/analysis/analysis_set.py:41-860).

AnalysisSet / AnalysisTask / VolumeAverage / TrackMode / PowerSpectrum keep
the reference's interfaces and output files.  Plots (Snapshot frames,
spectrum figures) are written with matplotlib when it is importable.
"""

import os

import numpy as np

from dedalus.config import decfg
from dedalus.utils.parallelism import com_sys


class AnalysisSet:

    def __init__(self, data, ti):
        self.data = data
        self.ti = ti
        self.tasks = []

    def add(self, task):
        self.tasks.append(task)
        task._an = self
        task._n = len(self.tasks)
        task.setup(self.data, self.ti.iteration)

    def run(self):
        for task in self.tasks:
            if self.ti.iteration % task.cadence == 0:
                task.run(self.data, self.ti.iteration)

    def cleanup(self):
        for task in self.tasks:
            task.cleanup(self.data, self.ti.iteration)


class AnalysisTask:

    def __init__(self, cadence, *args, **kwargs):
        self.cadence = cadence

    def setup(self, data, it):
        pass

    def run(self, data, it):
        pass

    def cleanup(self, data, it):
        pass


def _plane(comp, space, axis, index):
    arr = comp[space]
    nd = arr.ndim
    if nd == 2:
        return arr
    ax = {'z': 0, 'y': 1, 'x': 2}[axis] if space == 'xspace' else {'y': 0, 'z': 1, 'x': 2}[axis]
    idx = arr.shape[ax] // 2 if index == 'middle' else int(index)
    return np.take(arr, idx, axis=ax)


class Snapshot(AnalysisTask):
    """Planes of every field component, saved as frames (analysis_set.py:79-467)."""

    def __init__(self, cadence, space=None, axis=None, index=None, units=None,
                 cmap=None, even_scale=True, dpi=None):
        self.cadence = cadence
        self.space = space or decfg.get('analysis', 'snapshot_space')
        self.axis = axis or decfg.get('analysis', 'snapshot_axis')
        self.index = index if index is not None else decfg.get('analysis', 'snapshot_index')
        self.cmap = cmap or decfg.get('analysis', 'snapshot_cmap')
        self.dpi = dpi or decfg.getint('analysis', 'snapshot_dpi')

    def setup(self, data, it):
        os.makedirs('frames', exist_ok=True)

    def run(self, data, it):
        planes = {}
        for fname, field in data:
            for ci, comp in field:
                p = _plane(comp, self.space, self.axis, self.index)
                planes["%s%s" % (fname, field.ctrans[ci])] = np.abs(p) if np.iscomplexobj(p) else p
        base = 'frames/snap_%s_n%07i' % (self.space[0], it)
        np.savez(base + '.npz', **planes)
        try:
            import matplotlib
            matplotlib.use('AGG')
            import matplotlib.pyplot as plt
        except ImportError:
            return
        n = len(planes)
        fig, axes = plt.subplots(1, n, figsize=(4 * n, 4), squeeze=False)
        for ax, (name, p) in zip(axes[0], planes.items()):
            ax.imshow(p, cmap=self.cmap, origin='lower')
            ax.set_title(name)
        fig.suptitle('t = %6.3f' % data.time)
        fig.savefig(base + '.png', dpi=self.dpi)
        plt.close(fig)


class TrackMode(AnalysisTask):
    """Complex amplitudes of chosen modes to text (analysis_set.py:470-582)."""

    def __init__(self, cadence, fieldlist=None, modelist=(), indexlist=()):
        self.cadence = cadence
        self.fieldlist = fieldlist
        self.modelist = list(modelist)
        self.indexlist = list(indexlist)

    def setup(self, data, it):
        if self.fieldlist is None:
            self.fieldlist = list(data.fields.keys())
        cols = '# time' + ''.join('\t' + str(list(m)) for m in self.modelist)
        cols += ''.join('\t' + str(tuple(i)) for i in self.indexlist if i is not None)
        for fname in self.fieldlist:
            field = data[fname]
            for ci, comp in field:
                with open('%s%s_mode_amplitudes.dat' % (fname, field.ctrans[ci]), 'a') as f:
                    f.write("# Dedalus Mode Amplitudes\n" + cols + '\n')

    def run(self, data, it):
        for fname in self.fieldlist:
            field = data[fname]
            for ci, comp in field:
                kd = comp['kspace']
                amps = []
                for mode in self.modelist:
                    index = comp.find_mode(mode)
                    amps.append(kd[index] if index else 0.)
                for index in self.indexlist:
                    amps.append(kd[tuple(index)] if index is not None else 0.)
                with open('%s%s_mode_amplitudes.dat' % (fname, field.ctrans[ci]), 'a') as f:
                    f.write('%s\t' % data.time + '\t'.join(repr(complex(a)) for a in amps) + '\n')


def compute_spectrum(field, norm=1.):
    """Shell-binned power as PowerSpectrum._compute_spectrum (analysis_set.py:778-815).

    Per k-array entry: power = sum_c |u_c|^2 times 2 pi |k| (2D) or
    4 pi |k|^2 (3D); bins [k_i, k_i + dk) over [0, |kny|), at most 100,
    bin value = mean of the nonzero samples in the bin.  Returns
    (bin centres, values).
    """
    c0 = field[0]
    kmag = np.sqrt(c0.k2())
    power = np.zeros_like(kmag)
    for ci, comp in field:
        power += np.abs(comp['kspace']) ** 2
    if field.ndim == 2:
        p1 = norm * power * 2. * np.pi * kmag
    else:
        p1 = norm * power * 4. * np.pi * kmag ** 2
    kmax = np.sqrt(np.sum(np.asarray(c0.kny) ** 2))
    n = int(min(np.min(c0.global_shape['xspace']) / 2., 100))
    kbottom = np.linspace(0, kmax, n, endpoint=False)
    ktop = kbottom + kbottom[1]
    spectrum = np.zeros_like(kbottom)
    counts = np.zeros_like(kbottom)
    nz = p1 != 0
    for i in range(kbottom.size):
        mask = (kmag >= kbottom[i]) & (kmag < ktop[i]) & nz
        spectrum[i] = np.sum(p1[mask])
        counts[i] = float(np.sum(mask))
    counts[counts == 0] = 1.
    return kbottom + kbottom[1] / 2., spectrum / counts


class PowerSpectrum(AnalysisTask):
    """Power spectra of fields, written to <field>_power_spectra.dat (analysis_set.py:585-815)."""

    def __init__(self, cadence, fieldlist=None, norm=1., write=True, plot=True,
                 loglog=True, nyquistlines=False, dealiasinglines=True, dpi=None):
        self.cadence = cadence
        self.fieldlist = fieldlist
        self.norm = norm
        self.write = write
        self.plot = plot
        self.loglog = loglog
        self.dpi = dpi or decfg.getint('analysis', 'powerspectrum_dpi')

    def setup(self, data, it):
        if self.fieldlist is None:
            self.fieldlist = list(data.fields.keys())
        self.firstrun = [True] * len(self.fieldlist)
        if self.write:
            for fname in self.fieldlist:
                with open('%s_power_spectra.dat' % fname, 'w') as f:
                    f.write("# Dedalus Power Spectrum\n")
        if self.plot:
            os.makedirs('frames', exist_ok=True)

    def _compute_spectrum(self, field, norm):
        return compute_spectrum(field, norm)

    def run(self, data, it):
        results = []
        for col, fname in enumerate(self.fieldlist):
            k, spec = self._compute_spectrum(data[fname], self.norm)
            results.append((fname, k, spec))
            if self.write:
                with open('%s_power_spectra.dat' % fname, 'a') as f:
                    if self.firstrun[col]:
                        f.write('# k\n# ' + '\t'.join(repr(float(v)) for v in k) + '\n')
                    f.write('%s\t' % data.time + '\t'.join(repr(float(v)) for v in spec) + '\n')
            self.firstrun[col] = False
        if self.plot:
            try:
                import matplotlib
                matplotlib.use('AGG')
                import matplotlib.pyplot as plt
            except ImportError:
                return
            fig, axes = plt.subplots(1, len(results), figsize=(4 * len(results), 4), squeeze=False)
            for ax, (fname, k, spec) in zip(axes[0], results):
                good = spec > 0
                if self.loglog:
                    ax.loglog(k[good], spec[good])
                else:
                    ax.semilogy(k[good], spec[good])
                ax.set_title(fname)
            fig.suptitle('t = %6.3f' % data.time)
            fig.savefig('frames/power_spectra_n%07i.png' % it, dpi=self.dpi)
            plt.close(fig)


class VolumeAverage(AnalysisTask):
    """Run a VolumeAverageSet at a cadence (analysis_set.py:818-838)."""

    def __init__(self, cadence, va):
        self.cadence = cadence
        self.volume_average_object = va

    def run(self, data, it):
        self.volume_average_object.run()
