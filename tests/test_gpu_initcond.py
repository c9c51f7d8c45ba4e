"""Setup helpers of the drop-in package on the GPU: CFL time step,
vorticity-to-velocity initial conditions, turbulent initial conditions
(init_cond.py:129-133, 281-341, 383-438; physics.py:601-610; time_step.py:160-178)."""

import numpy as np
import pytest

from oracle import numpy_path as npth

pytestmark = pytest.mark.gpu

L2 = (2 * np.pi, 2 * np.pi)


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


def xspace(shape, k):
    c = npth.Component(shape, L2)
    c['kspace'] = np.array(k)
    return np.array(c['xspace'])


def vorticity_and_divergence(shape, ku, kv):
    cu, cv = npth.Component(shape, L2), npth.Component(shape, L2)
    cu['kspace'] = np.array(ku)
    cv['kspace'] = np.array(kv)
    return cv.deriv('x') - cu.deriv('y'), cu.deriv('x') + cv.deriv('y')


def test_cfl_dt():
    """dt = CFL * min(dx) / sqrt(max_i max_x u_i^2), limited to 1.05 x the previous dt."""
    m = mods()
    from dedalus.init_cond.seeded import taylor_green_2d
    case("cfl dt")
    shape = (64, 48)
    RHS = m.IncompressibleHydro(shape, m.FourierRepresentation)
    data = RHS.create_fields(0.)
    u = taylor_green_2d(shape)
    u = [1.7 * u[0], 0.4 * u[1]]
    for i in range(2):
        data['u'][i]['kspace'] = u[i]
    ti = m.RK2mid(RHS)
    ti.CFL = 0.4
    umax = np.sqrt(max((xspace(shape, a) ** 2).max() for a in u))
    dx = min(2 * np.pi / 64, 2 * np.pi / 48)
    dt = ti.cfl_dt(data)
    assert dt == pytest.approx(0.4 * dx / umax, rel=1e-13)
    ti.dt_old = dt / 2
    assert ti.cfl_dt(data) == pytest.approx(1.05 * dt / 2, rel=1e-15)


def test_vorticity_wave_and_mit_vortices():
    """The velocity built from a vorticity through psi = w / k^2 is solenoidal
    and has that vorticity (every mode but the mean)."""
    m = mods()
    from dedalus.init_cond.init_cond import MIT_vortices, vorticity_wave
    case("vorticity initial conditions")
    shape = (64, 64)
    RHS = m.IncompressibleHydro(shape, m.FourierRepresentation)
    data = RHS.create_fields(0.)
    vorticity_wave(data, (2.0, 3.0), 0.3 + 0.1j)
    w, div = vorticity_and_divergence(shape, data['u'][0]['kspace'], data['u'][1]['kspace'])
    assert np.abs(div).max() < 1e-15
    i0 = data['u'][0].find_mode((2.0, 3.0))
    assert w[i0] == pytest.approx((0.3 + 0.1j) / 2, rel=1e-14)
    w[i0] = 0
    assert np.abs(w).max() < 1e-15
    data2 = RHS.create_fields(0.)
    MIT_vortices(data2)
    w2, div2 = vorticity_and_divergence(shape, data2['u'][0]['kspace'], data2['u'][1]['kspace'])
    assert np.abs(div2).max() < 1e-13
    y, x = np.meshgrid(np.arange(64) * 2 * np.pi / 64, np.arange(64) * 2 * np.pi / 64, indexing='ij')
    wx = (np.exp(-((x - np.pi) ** 2 + (y - np.pi + np.pi / 4) ** 2) / 0.2)
          + np.exp(-((x - np.pi) ** 2 + (y - np.pi - np.pi / 4) ** 2) / 0.2)
          - 0.5 * np.exp(-((x - np.pi - np.pi / 4) ** 2 + (y - np.pi - np.pi / 4) ** 2) / 0.4))
    cw = npth.Component(shape, L2)
    cw['xspace'] = wx
    ref = np.array(cw['kspace'])
    ref[0, 0] = 0
    w2[0, 0] = 0
    assert np.abs(w2 - ref).max() < 1e-12 * np.abs(ref).max()


def test_turb_new_energy():
    """turb_new gives a solenoidal field with the requested total energy."""
    m = mods()
    from dedalus.init_cond.init_cond import mcwilliams_spec, turb_new
    from dedalus.init_cond.seeded import mean_square
    case("turb_new")
    shape = (64, 64)
    RHS = m.IncompressibleHydro(shape, m.FourierRepresentation)
    data = RHS.create_fields(0.)
    np.random.seed(3)
    turb_new(data, mcwilliams_spec, tot_en=0.7, k0=6.)
    ku, kv = data['u'][0]['kspace'], data['u'][1]['kspace']
    assert 0.5 * mean_square([ku, kv], shape) == pytest.approx(0.7, rel=1e-12)
    _, div = vorticity_and_divergence(shape, ku, kv)
    assert np.abs(div).max() < 1e-12 * np.abs(ku).max()
    ex = 0.5 * np.mean(xspace(shape, ku) ** 2 + xspace(shape, kv) ** 2)
    assert ex == pytest.approx(0.7, rel=1e-12)
