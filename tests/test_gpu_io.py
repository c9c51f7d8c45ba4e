"""Host-boundary I/O of the drop-in package (GPU): snapshots and restart,
TrackMode and PowerSpectrum outputs, the AnalysisSet cadence runner.

Reference behaviour: time_step.py:100-151 (advance snapshots at iteration 0
and every save_cadence), restart.py:33-99, analysis_set.py:41-61 (cadence),
:470-582 (TrackMode), :778-815 (PowerSpectrum._compute_spectrum).
"""

import glob
import json
import os

import numpy as np
import pytest

from helpers import kfields

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _tmpdir(tmp_path, monkeypatch):
    monkeypatch.chdir(tmp_path)


def mods():
    from dedalus.config import decfg
    decfg.set('utils', 'loglevel', 'warning')
    import logging
    logging.getLogger("Dedalus").setLevel(logging.WARNING)
    import dedalus.mods as m
    return m


def case(*a):
    print("case:", *a, flush=True)


def hydro(m, shape=(32, 32), nu=2e-3):
    RHS = m.IncompressibleHydro(shape, m.FourierRepresentation)
    RHS.parameters['nu'] = nu
    data = RHS.create_fields(0.)
    for i, a in enumerate(kfields('hydro', shape, 81)['u']):
        data['u'][i]['kspace'] = a
    return RHS, data


def test_snapshot_restart_round_trip():
    """Snapshots at iterations 0, 3, 6; restarting from the last one gives the
    state and time it saved, and one more step from it equals one more step of
    the original run bit for bit."""
    m = mods()
    case("snapshot / restart 32^2")
    RHS, data = hydro(m)
    ti = m.RK2mid(RHS)
    ti.save_cadence = 3
    dt = 1e-2
    saved = None
    for it in range(7):
        if it == 6:
            saved = [np.array(data['u'][i]['kspace']) for i in range(2)]
            t6 = data.time
        ti.advance(data, dt)
    snaps = sorted(glob.glob('snap_*'))
    assert snaps == ['snap_00000', 'snap_00001', 'snap_00002']
    with open(glob.glob('snap_00002/dedalus_obj_*.json')[0]) as f:
        meta = json.load(f)
    assert meta['physics'] == 'IncompressibleHydro' and meta['shape'] == [32, 32]
    assert meta['parameters']['nu'] == 2e-3 and meta['iteration'] == 6
    with np.load(glob.glob('snap_00002/data.cpu*.npz')[0], allow_pickle=False) as f:
        assert 'fields/u/0' in f.files and str(f['fields/u/0.space']) in ('kspace', 'xspace')
    RHS2, data2 = hydro(m)
    from dedalus.utils.restart import restart_into
    restart_into(data2, 'snap_00002')
    assert data2.time == pytest.approx(t6, abs=0)
    for i in range(2):
        assert np.array_equal(np.asarray(data2['u'][i]['kspace']), saved[i])
    ti2 = m.RK2mid(RHS2)
    ti2.do_advance(data2, dt)
    for i in range(2):
        assert np.array_equal(np.asarray(data2['u'][i]['kspace']), np.asarray(data['u'][i]['kspace']))


def test_track_mode_and_cadence():
    """TrackMode writes the k-space amplitudes of the chosen modes at the
    AnalysisSet cadence."""
    m = mods()
    case("track mode")
    RHS, data = hydro(m)
    ti = m.RK2mid(RHS)
    ti.save_cadence = 1000
    an = m.AnalysisSet(data, ti)
    an.add(m.TrackMode(2, fieldlist=['u'], modelist=[(1.0, 2.0)], indexlist=[(3, 1)]))
    expect = []
    for it in range(5):
        an.run()
        if it % 2 == 0:
            c0 = data['u'][0]
            kd = np.asarray(c0['kspace'])
            expect.append((kd[c0.find_mode((1.0, 2.0))], kd[3, 1]))
        ti.advance(data, 1e-2)
    an.cleanup()
    lines = [l for l in open('ux_mode_amplitudes.dat') if not l.startswith('#')]
    assert len(lines) == 3
    for line, (a, b) in zip(lines, expect):
        vals = line.split('\t')[1:]
        assert complex(vals[0].strip()) == complex(a)
        assert complex(vals[1].strip()) == complex(b)


def test_power_spectrum_bins():
    """compute_spectrum keeps the reference's binning: bins of width
    |k_Nyquist| / n over [0, |k_Nyquist|) with n = min(N/2, 100), power times
    2 pi |k| in 2D, and each bin the mean of its nonzero samples."""
    m = mods()
    from dedalus.analysis.analysis_set import compute_spectrum
    case("power spectrum")
    RHS = m.IncompressibleHydro((32, 32), m.FourierRepresentation)
    data = RHS.create_fields(0.)
    u, v = data['u'][0], data['u'][1]
    ku = np.zeros(u['kspace'].shape, dtype=complex)
    kv = np.zeros_like(ku)
    ku[1, 0] = 0.3          # |k| = 1 (kx = 1)
    kv[0, 1] = 0.4j         # |k| = 1 (ky = 1)
    ku[3, 4] = 0.1          # |k| = 5
    u['kspace'] = ku
    v['kspace'] = kv
    k, spec = compute_spectrum(data['u'])
    kmax = np.hypot(16.0, 16.0)
    width = kmax / 16
    assert len(k) == 16 and np.allclose(k, (np.arange(16) + 0.5) * width)
    two_pi = 2 * np.pi
    expect = np.zeros(16)
    expect[int(1 / width)] = np.mean([two_pi * 0.09, two_pi * 0.16])   # two samples at |k| = 1
    expect[int(5 / width)] = two_pi * 5 * 0.01
    assert np.allclose(spec, expect, rtol=1e-14, atol=0)


def steady_forcing(data, i):
    """A self-contained forcing function: its source is all a restart has."""
    import numpy as np
    f = np.zeros(data['u'][i]['kspace'].shape, dtype=complex)
    f[1, 2] = 0.05j if i == 0 else -0.02
    f[2, 1] = 0.03 if i == 0 else 0.04j
    return f


@pytest.mark.parametrize("forced", [False, True], ids=["fused", "forced"])
def test_restart_continues_bit_exact(forced):
    """restart(snap_dir) rebuilds (RHS, data, ti) from a snapshot (restart.py:33-69):
    class, parameters, forcing function rebound by name, fields, clock and
    stopping controls; running it to the stop iteration equals the
    uninterrupted run bit for bit."""
    m = mods()
    from dedalus.utils.restart import restart
    case("restart", "forced" if forced else "fused")
    RHS, data = hydro(m, nu=1.5e-3)
    RHS.parameters['viscosity_order'] = 2
    if forced:
        RHS.set_velocity_forcing(steady_forcing)
    ti = m.RK2mid(RHS)
    ti.save_cadence = 3
    ti.stop_iteration = 9
    dt = 1e-2
    while ti.ok:
        ti.advance(data, dt)
    final = [np.array(data['u'][i]['kspace']) for i in range(2)]
    assert sorted(glob.glob('snap_*')) == ['snap_00000', 'snap_00001', 'snap_00002']

    RHS2, data2, ti2 = restart('snap_00002')
    assert type(RHS2) is m.IncompressibleHydro and type(ti2) is m.RK2mid
    assert RHS2.parameters['nu'] == 1.5e-3 and RHS2.parameters['viscosity_order'] == 2
    assert ti2.iteration == 6 and ti2.stop_iteration == 9 and ti2.save_cadence == 3
    assert data2.time == pytest.approx(6 * dt, abs=1e-15)
    if forced:
        assert RHS2.forcing_functions['VelocityForcing'].__name__ == 'steady_forcing'
    else:
        assert not RHS2.forcing_functions
    while ti2.ok:
        ti2.advance(data2, dt)
    assert ti2.iteration == 9
    for i in range(2):
        assert np.array_equal(np.asarray(data2['u'][i]['kspace']), final[i])
    # the restarted run continues the snapshot numbering
    assert os.path.isdir('snap_00003')


def test_restart_missing_forcing_raises(tmp_path):
    """A snapshot naming a forcing function its source does not define
    cannot be restarted (restart.py:56-59)."""
    m = mods()
    from dedalus.utils.restart import restart
    case("restart, missing forcing")
    RHS, data = hydro(m)
    RHS.set_velocity_forcing(steady_forcing)
    ti = m.RK2mid(RHS)
    ti.advance(data, 1e-2)
    with open('snap_00000/forcing_functions.py', 'w') as f:
        f.write("def other(data, i):\n    return 0\n")
    with pytest.raises(ValueError):
        restart('snap_00000')
