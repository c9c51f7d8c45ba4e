"""The drop-in API (``from dedalus.mods import *``) on the GPU vs the oracle.

These read like the reference's sample scripts: build a physics object,
create fields, set k-space initial conditions, advance with RK2mid /
RK2trap, and compare with the oracle on the same inputs.
"""

import os

import numpy as np
import pytest

from helpers import kfields, random_k, rel_l2
from oracle import numpy_path as npth

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


def set_fields(data, kf):
    for name, comps in kf.items():
        for i, a in enumerate(comps):
            data[name][i]['kspace'] = a


def compare(data, ref, tol=1e-12):
    for name, comps in ref.items():
        for i, r in enumerate(comps):
            got = data[name][i]['kspace']
            assert rel_l2(got, r) < tol, (name, i, rel_l2(got, r))


def test_taylor_green_64_closed_form():
    """The 2D Taylor-Green workload: 64^2, nu = 1e-3, RK2mid dt = 1e-3, 200 steps."""
    m = mods()
    from dedalus.init_cond.seeded import workload
    case("taylor-green 64^2 200 steps")
    w = workload('taylor_green_64')
    RHS = m.IncompressibleHydro(w['shape'], m.FourierRepresentation)
    RHS.parameters['nu'] = 1e-3
    data = RHS.create_fields(0.)
    set_fields(data, w['fields'])
    ti = m.RK2mid(RHS)
    ti.stop_iteration = 200
    while ti.ok:
        ti.advance(data, 1e-3)
    assert ti.iteration == 200 and abs(data.time - 0.2) < 1e-12
    fac = np.exp(-2 * 1e-3 * 0.2)    # |k|^2 = 2
    for i in range(2):
        assert rel_l2(data['u'][i]['kspace'], w['fields']['u'][i] * fac) < 1e-13
    ref = npth.run('hydro', w['shape'], w['fields'], 1e-3, 200, nu=1e-3)
    compare(data, ref)
    assert os.path.isdir('snap_00000')


@pytest.mark.parametrize("kind,shape,p,integ", [
    ('hydro', (32, 32), {'nu': 2e-3}, 'RK2mid'),
    ('hydro', (16, 16, 16), {'nu': 5e-3, 'viscosity_order': 2}, 'RK2trap'),
    ('boussinesq', (32, 32), {'nu': 1e-3, 'kappa': 1e-3, 'boussinesq_direction': 'y'}, 'RK2mid'),
    ('boussinesq', (8, 16, 16), {'nu': 1e-3, 'kappa': 2e-3, 'boussinesq_direction': 'z'}, 'RK2trap'),
    ('mhd', (32, 32), {'nu': 1e-3, 'eta': 1e-3, 'rho0': 1.5}, 'RK2mid'),
    ('mhd', (16, 16, 16), {'nu': 1e-3, 'eta': 2e-3}, 'RK2mid'),
], ids=str)
def test_advance_matches_oracle(kind, shape, p, integ):
    m = mods()
    case("advance", kind, shape, p, integ)
    cls = {'hydro': m.IncompressibleHydro, 'boussinesq': m.BoussinesqHydro,
           'mhd': m.IncompressibleMHD}[kind]
    RHS = cls(shape, m.FourierRepresentation)
    RHS.parameters.update(p)
    data = RHS.create_fields(0.)
    kf = kfields(kind, shape, 50)
    set_fields(data, kf)
    ti = (m.RK2mid if integ == 'RK2mid' else m.RK2trap)(RHS)
    for _ in range(6):
        ti.advance(data, 5e-3)
    ph = npth.make_physics(kind, shape, **dict(p))
    odata = npth.load_state(ph, kf)
    oti = (npth.RK2mid if integ == 'RK2mid' else npth.RK2trap)(ph)
    for _ in range(6):
        oti.do_advance(odata, 5e-3)
    compare(data, npth.read_state(odata))
    # the integrator's own storage after the last step: deriv1 = F(y_n),
    # deriv2 = F(data2) (RK2mid) or F of the ETD1 state (RK2trap), data2 the
    # midpoint (time_step.py:286, 297, 370, 381)
    compare(ti.deriv1, npth.read_state(oti.deriv1))
    compare(ti.deriv2, npth.read_state(oti.deriv2))
    if integ == 'RK2mid':
        compare(ti.data2, npth.read_state(oti.data2))
    IF = ti.deriv1['u'][0].integrating_factor
    assert IF is not None and np.allclose(np.asarray(IF), p['nu'] * data['u'][0].k2() ** p.get('viscosity_order', 1))


def test_forcing_path_matches_oracle():
    """A Python velocity forcing runs the per-stage path (RHS + device ETD)."""
    m = mods()
    shape = (32, 32)
    case("forcing", shape)
    kf = kfields('hydro', shape, 60)
    fk = [random_k(shape, 61 + i) * 0.1 for i in range(2)]

    def forcing(data, i):
        return fk[i]

    RHS = m.IncompressibleHydro(shape, m.FourierRepresentation)
    RHS.parameters['nu'] = 1e-3
    RHS.set_velocity_forcing(forcing)
    data = RHS.create_fields(0.)
    set_fields(data, kf)
    ti = m.RK2mid(RHS)
    for _ in range(4):
        ti.advance(data, 1e-2)
    ph = npth.make_physics('hydro', shape, nu=1e-3)
    ph.forcing_functions['VelocityForcing'] = forcing
    od = npth.load_state(ph, kf)
    oti = npth.RK2mid(ph)
    for _ in range(4):
        oti.do_advance(od, 1e-2)
    compare(data, npth.read_state(od))


def test_space_state_machine():
    """Lazy transforms, buffer-preserving assignment, transform counts."""
    m = mods()
    shape = (16, 24)
    case("state machine", shape)
    L = (1.0, 3.0)
    RHS = m.IncompressibleHydro(shape, m.FourierRepresentation, length=L)
    data = RHS.create_fields(0.)
    rep = data['u']['x']
    oc = npth.Component(shape, L)
    x = np.random.default_rng(7).standard_normal(shape)
    rep['xspace'] = x
    oc['xspace'] = x
    assert rep._curr_space == 'xspace'
    k = rep['kspace']
    assert rep.fwd_count == 1 and rep._curr_space == 'kspace'
    assert rel_l2(k, oc['kspace']) < 1e-14
    # host write into the k array, then back to x space
    k[1, 2] += 0.25 - 0.5j
    oc.kdata[1, 2] += 0.25 - 0.5j
    xx = rep['xspace']
    assert rep.rev_count == 1
    assert rel_l2(xx, oc['xspace']) < 1e-14
    # k space again: dealiased data of the round trip
    assert rel_l2(rep['kspace'], oc['kspace']) < 1e-14
    # scalar assignment and the wavenumber helpers
    rep['kspace'] = 0.
    assert np.abs(rep['kspace']).max() == 0
    assert np.array_equal(rep.k2(), oc.k2())
    for nm in ('x', 'y'):
        assert np.array_equal(rep.k[nm], oc.k[nm])
    assert np.array_equal(rep.kny, oc.kny)
    assert rep.find_mode((2 * np.pi / L[1] * 2, 2 * np.pi / L[0] * 3)) == (2, 3)


def test_operators_match_oracle():
    """gradX, divX, XgradY, XcrossY, curlX, laplace_solve, div_free (device kernels)."""
    m = mods()
    shape = (16, 12, 8)
    case("operators", shape)
    RHS = m.IncompressibleMHD(shape, m.FourierRepresentation)
    data = RHS.create_fields(0.)
    kf = kfields('mhd', shape, 70)
    set_fields(data, kf)
    RHS._finalize()
    aux = RHS.aux_fields
    ph = npth.make_physics('mhd', shape)
    od = npth.load_state(ph, kf)
    ph.aux = npth.State(shape, ph.length, ph.aux_list)
    oaux = ph.aux
    # divX + laplace_solve
    RHS.divX(data['u'], aux['mathscalar'])
    ph.divX(od['u'], oaux['mathscalar'])
    assert rel_l2(aux['mathscalar']['kspace'], oaux['mathscalar']['kspace']) < 1e-14
    RHS.laplace_solve(aux['mathscalar'], aux['mathscalar'])
    ph.laplace_solve(oaux['mathscalar'], oaux['mathscalar'])
    assert rel_l2(aux['mathscalar']['kspace'], oaux['mathscalar']['kspace']) < 1e-14
    # curlX of B
    RHS.curlX(data['B'], aux['mathvector'])
    ph.curlX(od['B'], oaux['mathvector'])
    for i in range(3):
        assert rel_l2(aux['mathvector'][i]['kspace'], oaux['mathvector'][i]['kspace']) < 1e-14
    # XcrossY (moves B to x space in both)
    RHS.XcrossY(aux['mathvector'], data['B'], aux['mathvector2'])
    ph.XcrossY(oaux['mathvector'], od['B'], oaux['mathvector2'])
    for i in range(3):
        assert rel_l2(aux['mathvector2'][i]['xspace'], oaux['mathvector2'][i]['xspace']) < 1e-13
    assert data['B']['x']._curr_space == 'xspace'
    # XgradY
    deriv = RHS.create_fields(0.)
    oderiv = ph.create_fields()
    RHS.XgradY(data['u'], data['u'], aux['mathscalar'], aux['mathvector'], deriv['u'])
    ph.XgradY(od['u'], od['u'], oaux['mathscalar'], oaux['mathvector'], oderiv['u'])
    for i in range(3):
        assert rel_l2(deriv['u'][i]['xspace'], oderiv['u'][i]['xspace']) < 1e-13
    # div_free = pressure projection
    data['u'].div_free()
    ph.pressure_projection(None, od)
    for i in range(3):
        assert rel_l2(data['u'][i]['kspace'], od['u'][i]['kspace']) < 1e-14


def test_volume_averages():
    m = mods()
    shape = (32, 32)
    case("volume averages", shape)
    RHS = m.IncompressibleHydro(shape, m.FourierRepresentation)
    data = RHS.create_fields(0.)
    kf = kfields('hydro', shape, 80)
    set_fields(data, kf)
    vs = m.VolumeAverageSet(data)
    e = vs.known_analysis['ekin'](data, vs.scratch)
    ex = vs.known_analysis['ekin'](data, vs.scratch, space='xspace')
    z = vs.known_analysis['enstrophy'](data, vs.scratch)
    oc = npth.Component(shape, (2 * np.pi,) * 2)
    xs = []
    for a in kf['u']:
        oc['kspace'] = a
        xs.append(oc['xspace'].copy())
    e_ref = 0.5 * sum((x ** 2).mean() for x in xs)
    assert abs(e - e_ref) < 1e-13 * e_ref and abs(ex - e_ref) < 1e-13 * e_ref
    wk = 1j * oc.k['x'] * kf['u'][1] - 1j * oc.k['y'] * kf['u'][0]
    oc['kspace'] = wk
    z_ref = 0.5 * (oc['xspace'] ** 2).mean()
    assert abs(z - z_ref) < 1e-12 * z_ref
    # divergence: Hermitian-weighted sum of the k-space divergence (volume_average.py:71-106)
    d = 1j * oc.k['x'] * kf['u'][0] + 1j * oc.k['y'] * kf['u'][1]
    d_ref = 2 * d[1:, :].sum() + d[0, :].sum()
    assert abs(vs.known_analysis['divergence'](data, vs.scratch) - d_ref) < 1e-12 * abs(d_ref)


@pytest.mark.parametrize("kind,shape,p", [
    ('hydro', (32, 32), {'nu': 2e-3}),
    ('boussinesq', (32, 32), {'nu': 1e-3, 'kappa': 1e-3, 'boussinesq_direction': 'y'}),
    ('mhd', (16, 16, 16), {'nu': 1e-3, 'eta': 1e-3}),
], ids=str)
def test_subclass_runs_reference_operator_sequence(kind, shape, p):
    """A user subclass that inherits RHS gets the reference's class-dependent
    behaviour (physics.py:575, 697, 805: no pressure projection): RHS and
    three RK2mid steps against the oracle's subclass of the same class."""
    m = mods()
    case("subclass", kind, shape)
    base = {'hydro': m.IncompressibleHydro, 'boussinesq': m.BoussinesqHydro,
            'mhd': m.IncompressibleMHD}[kind]
    obase = {'hydro': npth.IncompressibleHydro, 'boussinesq': npth.BoussinesqHydro,
             'mhd': npth.IncompressibleMHD}[kind]
    Sub = type('Sub', (base,), {})
    OSub = type('OSub', (obase,), {})
    RHS = Sub(shape, m.FourierRepresentation)
    RHS.parameters.update(p)
    kf = kfields(kind, shape, 70)
    data = RHS.create_fields(0.)
    set_fields(data, kf)
    deriv = RHS.create_fields(0.)
    RHS.RHS(data, deriv)
    q = dict(p)
    oph = OSub(shape) if kind != 'boussinesq' else OSub(shape, None, q.pop('boussinesq_direction'))
    oph.parameters.update(q)
    odata = npth.load_state(oph, kf)
    oderiv = oph.create_fields()
    oph.RHS(odata, oderiv)
    compare(deriv, npth.read_state(oderiv))
    # fresh objects for the run: the first RHS call sets the integrating
    # factors on the deriv it is given (physics.py:538-540), as in the reference
    RHS = Sub(shape, m.FourierRepresentation)
    RHS.parameters.update(p)
    data = RHS.create_fields(0.)
    set_fields(data, kf)
    oph = OSub(shape) if kind != 'boussinesq' else OSub(shape, None, p['boussinesq_direction'])
    oph.parameters.update(q)
    odata = npth.load_state(oph, kf)
    ti = m.RK2mid(RHS)
    oti = npth.RK2mid(oph)
    for _ in range(3):
        ti.advance(data, 5e-3)
        oti.do_advance(odata, 5e-3)
    compare(data, npth.read_state(odata))


def _aniso_factors(self, deriv):
    """Integrating factors that are not coeff |k|^(2 order): an anisotropic
    hyperviscosity nu (kx^2 + 3 ky^2)^2 plus nu/2 |k|^2, one array per
    velocity component (what a subclass may set; physics.py:504-513 sets
    nu |k|^(2 vo))."""
    nu = self.parameters['nu']
    for _, comp in deriv['u']:
        kx, ky = comp.k['x'], comp.k['y']
        comp.integrating_factor = nu * (kx ** 2 + 3. * ky ** 2) ** 2 + 0.5 * nu * comp.k2()


@pytest.mark.parametrize("shape,integ", [((32, 24), 'RK2mid'), ((32, 24), 'RK2trap'),
                                         ((16, 12, 20), 'RK2mid'), ((16, 12, 20), 'RK2trap')],
                         ids=str)
def test_array_integrating_factors(shape, integ):
    """A subclass whose integrating factors are arbitrary per-mode arrays
    (time_step.py:289-304, 373-388 pass them to the sweeps as -IF) advances
    through ddx_update_if: four steps against the oracle's subclass with the
    same arrays, and the integrator's derivatives after the last one."""
    m = mods()
    case("array integrating factors", shape, integ)
    Sub = type('Sub', (m.IncompressibleHydro,), {'_setup_integrating_factors': _aniso_factors})
    OSub = type('OSub', (npth.IncompressibleHydro,), {'_setup_integrating_factors': _aniso_factors})
    p = {'nu': 4e-3}
    RHS = Sub(shape, m.FourierRepresentation)
    RHS.parameters.update(p)
    data = RHS.create_fields(0.)
    kf = kfields('hydro', shape, 80)
    set_fields(data, kf)
    oph = OSub(shape)
    oph.parameters.update(p)
    odata = npth.load_state(oph, kf)
    ti = (m.RK2mid if integ == 'RK2mid' else m.RK2trap)(RHS)
    oti = (npth.RK2mid if integ == 'RK2mid' else npth.RK2trap)(oph)
    assert not ti._fusable()
    for _ in range(4):
        ti.advance(data, 4e-3)
        oti.do_advance(odata, 4e-3)
    IF = ti.deriv1['u'][0].integrating_factor
    assert isinstance(IF, np.ndarray)
    compare(data, npth.read_state(odata))
    compare(ti.deriv1, npth.read_state(oti.deriv1))
    compare(ti.deriv2, npth.read_state(oti.deriv2))


@pytest.mark.parametrize("f64", [True, False], ids=["f64", "f32"])
@pytest.mark.parametrize("shape", [(24, 20), (12, 16, 10)], ids=str)
def test_update_if_sweeps(shape, f64):
    """ddx_update_if for each sweep kind against the oracle's sweeps
    (forward_step_cy_{2,3}d.pyx restated) on random states and a random
    positive factor, including |Z| < 0.5 (Taylor branch) and IF = 0 modes."""
    m = mods()
    from dedalus import _hip
    from dedalus.data_objects import kgrid
    case("update_if", shape, "f64" if f64 else "f32")
    plan = _hip.Plan(shape, (2 * np.pi,) * len(shape), _hip.F64 if f64 else _hip.F32)
    rng = np.random.default_rng(11)
    ksh = kgrid.kshape(shape)
    mask = kgrid.dealias_mask(shape, (2 * np.pi,) * len(shape))
    IF = rng.uniform(0., 40., ksh)
    IF[rng.random(ksh) < 0.1] = 0.
    vals = [random_k(shape, 30 + i) for i in range(3)]
    bufs = []
    for v in vals:
        b = plan.alloc_k()
        plan.upload_k(b, v)
        bufs.append(b)
    ifb = plan.upload_real_k(IF)
    out = plan.alloc_k()
    dt = 0.03
    nd = len(shape)
    for kind, name in ((_hip.EULER, 'euler'), (_hip.ETD1, 'etd1'), (_hip.ETD2RK1, 'etd2rk1'),
                       (_hip.ETD2RK2, 'etd2rk2')):
        plan.update_if(kind, bufs[0], out, bufs[1], bufs[2], ifb, dt)
        got = plan.download_k(out)
        ref = np.zeros(ksh, dtype=np.complex128)
        if name == 'euler':
            npth.euler(vals[0], ref, vals[1], dt)
        elif name == 'etd1':
            npth.etd1(vals[0], ref, vals[1], -IF, dt, nd)
        elif name == 'etd2rk1':
            npth.etd2rk1(vals[0], ref, vals[1], vals[2], -IF, dt, nd)
        else:
            npth.etd2rk2(vals[0], ref, vals[1], vals[2], -IF, dt, nd)
        ref[mask] = 0.
        assert rel_l2(got, ref) < (1e-14 if f64 else 2e-7), (name, rel_l2(got, ref))


RULE_CASES = [
    ('hydro', (32, 32), {'nu': 2e-3}, 'RK2mid'),
    ('hydro', (64, 48), {'nu': 2e-3}, 'RK2trap'),
    ('hydro', (16, 16, 16), {'nu': 5e-3}, 'RK2mid'),
    ('boussinesq', (32, 32), {'nu': 1e-3, 'kappa': 1e-3, 'boussinesq_direction': 'y'}, 'RK2mid'),
    ('mhd', (16, 16, 16), {'nu': 1e-3, 'eta': 2e-3}, 'RK2trap'),
    ('mhd', (32, 32, 32), {'nu': 1e-3, 'eta': 1e-3}, 'RK2mid'),
]


class _Dealiasing:
    """decfg's FFT.dealiasing and the oracle's for the duration of a test."""

    def __init__(self, rule):
        self.rule = rule

    def __enter__(self):
        from dedalus.config import decfg
        self.old = decfg.get('FFT', 'dealiasing')
        decfg.set('FFT', 'dealiasing', self.rule)
        self.oold = npth.DEALIASING
        npth.DEALIASING = self.rule

    def __exit__(self, *exc):
        from dedalus.config import decfg
        decfg.set('FFT', 'dealiasing', self.old)
        npth.DEALIASING = self.oold


@pytest.mark.parametrize("rule", ['2/3 spherical', 'None'])
@pytest.mark.parametrize("kind,shape,p,integ", RULE_CASES, ids=str)
def test_dealias_rules_advance(rule, kind, shape, p, integ):
    """The other dealiasing methods of set_dealiasing (representations.py:
    359-382) through the fused step: RHS on its own, then five steps, and the
    integrator's derivatives, against the oracle under the same rule."""
    m = mods()
    case("dealias rule", rule, kind, shape, integ)
    with _Dealiasing(rule):
        cls = {'hydro': m.IncompressibleHydro, 'boussinesq': m.BoussinesqHydro,
               'mhd': m.IncompressibleMHD}[kind]
        kf = kfields(kind, shape, 120)
        RHS = cls(shape, m.FourierRepresentation)
        RHS.parameters.update(p)
        data = RHS.create_fields(0.)
        set_fields(data, kf)
        assert data['u'][0]._plan.dealias_rule == {'2/3 spherical': 1, 'None': 2}[rule]
        deriv = RHS.create_fields(0.)
        RHS.RHS(data, deriv)
        ph = npth.make_physics(kind, shape, **dict(p))
        odata = npth.load_state(ph, kf)
        oderiv = ph.create_fields()
        ph.RHS(odata, oderiv)
        compare(deriv, npth.read_state(oderiv))
        RHS = cls(shape, m.FourierRepresentation)
        RHS.parameters.update(p)
        data = RHS.create_fields(0.)
        set_fields(data, kf)
        ti = (m.RK2mid if integ == 'RK2mid' else m.RK2trap)(RHS)
        ph = npth.make_physics(kind, shape, **dict(p))
        odata = npth.load_state(ph, kf)
        oti = (npth.RK2mid if integ == 'RK2mid' else npth.RK2trap)(ph)
        for _ in range(5):
            ti.advance(data, 5e-3)
            oti.do_advance(odata, 5e-3)
        compare(data, npth.read_state(odata))
        compare(ti.deriv1, npth.read_state(oti.deriv1))
        compare(ti.deriv2, npth.read_state(oti.deriv2))


@pytest.mark.parametrize("f64", [True, False], ids=["f64", "f32"])
@pytest.mark.parametrize("rule", ['2/3 spherical', 'None'])
@pytest.mark.parametrize("shape", [(64, 64), (128, 96), (32, 32, 32), (16, 24, 20)], ids=str)
def test_dealias_rules_transforms(shape, rule, f64):
    """Plans of each rule: forward transform of a random x field against
    the oracle's forward() (which dealiases its output), backward of k data
    that also fills the modes the rule removes (the spherical rule's shell
    inside the compact cube) against backward(), which dealiases its input
    in place first -- the device buffer is zeroed there too."""
    from dedalus import _hip
    from helpers import random_x
    case("dealias transforms", shape, rule, "f64" if f64 else "f32")
    code = {'2/3 spherical': _hip.DEALIAS_SPHERICAL, 'None': _hip.DEALIAS_NONE}[rule]
    plan = _hip.Plan(shape, (2 * np.pi,) * len(shape), _hip.F64 if f64 else _hip.F32, dealias=code)
    tol = 1e-13 if f64 else 2e-6
    with _Dealiasing(rule):
        c = npth.Component(shape, (2 * np.pi,) * len(shape))
    x = random_x(shape, 5)
    c['xspace'] = x
    kref = c['kspace'].copy()
    xd, kd = plan.alloc_x(), plan.alloc_k()
    plan.upload_x(xd, x)
    plan.fwd(xd, kd)
    assert rel_l2(plan.download_k(kd), kref) < tol
    # k input with every mode of the 2/3 cube filled (Hermitian from x space)
    with _Dealiasing('2/3'):
        c23 = npth.Component(shape, (2 * np.pi,) * len(shape))
    c23['xspace'] = random_x(shape, 6)
    kin = c23['kspace'].copy()
    if rule == 'None':
        c['xspace'] = random_x(shape, 6)
        kin = c['kspace'].copy()
    plan.upload_k(kd, kin)
    plan.bwd(kd, xd)
    c['kspace'] = kin
    xref = c['xspace'].copy()
    assert rel_l2(plan.download_x(xd), xref) < tol
    masked = kin.copy()
    masked[c.dealias_mask()] = 0.
    assert rel_l2(plan.download_k(kd), masked) < tol


def test_fused_rhs_refuses_undealiased_plans():
    """ddx_rhs / ddx_step on a plan without dealiasing fail with
    DDX_ERR_UNSUPPORTED before launching anything (their rotational form is
    the reference's only for dealiased products); the drop-in then runs the
    operator sequence instead (test_dealias_rules_advance)."""
    from dedalus import _hip
    case("fused rhs refuses dealias None")
    plan = _hip.Plan((32, 32), (2 * np.pi, 2 * np.pi), _hip.F64, dealias=_hip.DEALIAS_NONE)
    prm = _hip.make_params(_hip.HYDRO, nu=1e-3)
    st = [plan.alloc_k() for _ in range(2)]
    d1 = [plan.alloc_k() for _ in range(2)]
    with pytest.raises(_hip.DDXError) as e:
        plan.rhs(prm, st, d1)
    assert e.value.code == -3 and "dealiasing" in str(e.value)
    with pytest.raises(_hip.DDXError) as e:
        plan.step(_hip.RK2TRAP, prm, st, None, d1, None, 1, 1e-3)
    assert e.value.code == -3


def test_set_dealiasing_moves_the_data():
    """set_dealiasing on an existing component (representations.py:359-382)
    rebinds it to a plan of the new rule and keeps its data: k space as set,
    and a backward transform then dealiases by the new rule, as the
    reference's backward() does."""
    m = mods()
    from dedalus.data_objects.representations import FourierRepresentation
    case("set_dealiasing migration")
    shape = (32, 32)
    with _Dealiasing('2/3'):
        c = FourierRepresentation(None, shape, (2 * np.pi, 2 * np.pi))
        oc = npth.Component(shape, (2 * np.pi, 2 * np.pi))
    kin = random_k(shape, 77)
    c['kspace'] = kin
    plan23 = c._plan
    c.set_dealiasing('2/3 spherical')
    assert c._plan is not plan23 and c._plan.dealias_rule == 1 and c.nmodes is None
    assert rel_l2(c['kspace'], kin) < 1e-15
    x = c['xspace'].copy()
    oc.set_dealiasing('2/3 spherical')
    oc['kspace'] = kin
    assert rel_l2(x, oc['xspace']) < 1e-13
    assert not np.abs(c['kspace'][c.dealias_mask()]).any()
    c.set_dealiasing('2/3')
    assert c._plan.dealias_rule == 0
