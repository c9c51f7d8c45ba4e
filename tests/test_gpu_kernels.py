"""Parity of every libddx kernel with the numpy oracle (GPU).

Each test prints the case it runs before touching the device.
"""

import os

import numpy as np
import pytest

from helpers import kfields, random_k, random_x, rel_l2
from oracle import numpy_path as npth

pytestmark = pytest.mark.gpu

SHAPES = [(8, 8), (6, 8), (16, 16), (64, 64), (48, 64), (30, 20), (128, 128), (512, 512),
          (8, 8, 8), (4, 6, 8), (16, 16, 16), (12, 10, 16), (48, 2, 48), (32, 32, 32),
          # size-specialised x-row transforms (two rows per team; odd row counts
          # leave the last team one row)
          (5, 64), (3, 5, 64), (4, 6, 128), (12, 256), (2, 3, 512)]
LENGTHS = {2: (1.0, 2.5), 3: (2 * np.pi, 1.5, 3.0)}
FIELDS = {'hydro': ['u'], 'boussinesq': ['u', 'T'], 'mhd': ['u', 'B']}
BDIR = {'x': 0, 'y': 1, 'z': 2}


def hip():
    from dedalus import _hip
    _hip.init()
    return _hip


def case(*args):
    print("case:", *args, flush=True)


def params_struct(hp, kind, nd, p):
    phys = {'hydro': hp.HYDRO, 'boussinesq': hp.BOUSSINESQ, 'mhd': hp.MHD}[kind]
    return hp.make_params(phys, nu=p.get('nu', 0.), viscosity_order=p.get('viscosity_order', 1),
                          kappa=p.get('kappa', 0.), eta=p.get('eta', 0.), g=p.get('g', 1.),
                          alpha_t=p.get('alpha_t', 1.), beta=p.get('beta', 1.),
                          bdir=BDIR[p.get('boussinesq_direction', 'z' if nd == 3 else 'y')],
                          rho0=p.get('rho0', 1.))


def upload_fields(plan, kind, kf):
    bufs = []
    for name in FIELDS[kind]:
        for arr in kf[name]:
            b = plan.alloc_k()
            plan.upload_k(b, arr)
            bufs.append(b)
    return bufs


def download_fields(plan, kind, nd, bufs):
    out, i = {}, 0
    for name in FIELDS[kind]:
        nc = nd if name in ('u', 'B') else 1
        out[name] = [plan.download_k(bufs[i + c]) for c in range(nc)]
        i += nc
    return out


@pytest.mark.parametrize("shape", SHAPES, ids=str)
def test_pack_unpack(shape):
    hp = hip()
    case("pack/unpack", shape)
    L = (2 * np.pi,) * len(shape)
    plan = hp.get_plan(shape, L)
    comp = npth.Component(shape, L)
    kref = np.random.default_rng(1).standard_normal(comp.kshape) * (1 + 1j)
    b = plan.alloc_k()
    plan.upload_k(b, kref)
    back = plan.download_k(b)
    expect = np.where(comp.dealias_mask(), 0, kref)
    assert np.array_equal(back, expect)


@pytest.mark.parametrize("shape", SHAPES, ids=str)
def test_forward(shape):
    hp = hip()
    L = LENGTHS[len(shape)]
    case("forward", shape, L)
    plan = hp.get_plan(shape, L)
    comp = npth.Component(shape, L)
    x = random_x(shape, 3)
    comp['xspace'] = x
    ref = comp['kspace']
    xd = plan.alloc_x()
    kc = plan.alloc_k()
    plan.upload_x(xd, x)
    plan.fwd(xd, kc)
    got = plan.download_k(kc)
    assert rel_l2(got, ref) < 1e-14


@pytest.mark.parametrize("shape", SHAPES, ids=str)
@pytest.mark.parametrize("hermitian", [True, False])
def test_backward(shape, hermitian):
    hp = hip()
    L = LENGTHS[len(shape)]
    case("backward", shape, L, hermitian)
    plan = hp.get_plan(shape, L)
    comp = npth.Component(shape, L)
    if hermitian:
        k = random_k(shape, 4, L)
    else:
        rng = np.random.default_rng(5)
        k = rng.standard_normal(comp.kshape) + 1j * rng.standard_normal(comp.kshape)
        k[comp.dealias_mask()] = 0
    comp['kspace'] = k
    ref = comp['xspace']
    kc = plan.alloc_k()
    xd = plan.alloc_x()
    plan.upload_k(kc, k)
    plan.bwd(kc, xd)
    got = plan.download_x(xd)
    assert rel_l2(got, ref) < 1e-14


@pytest.mark.parametrize("shape,nf", [((64, 64), 3), ((48, 64), 2), ((16, 16, 16), 3),
                                      ((4, 6, 128), 5), ((5, 64), 2), ((64, 450), 16)], ids=str)
def test_transforms_of_several_fields(shape, nf):
    """ddx_fwd_many / ddx_bwd_many equal nf single-field calls: bit for bit
    where the size-specialised row kernels run (each transforms two rows of
    one field); to rounding where the generic row kernel pairs two fields in
    one complex transform."""
    hp = hip()
    L = LENGTHS[len(shape)]
    case("fwd/bwd many", shape, nf)
    plan = hp.Plan(shape, L)
    ks = [plan.alloc_k() for _ in range(nf)]
    xs = [plan.alloc_x() for _ in range(nf)]
    for f in range(nf):
        plan.upload_k(ks[f], random_k(shape, 200 + f, L))
    plan.bwd_many(ks, xs)
    xm = [plan.download_x(b) for b in xs]
    x1 = []
    for f in range(nf):
        plan.bwd(ks[f], xs[f])
        x1.append(plan.download_x(xs[f]))
    fast = shape[-1] in (64, 128, 256, 512)

    def same(a, b):
        return np.array_equal(a, b) if fast else rel_l2(a, b) < 1e-15

    for f in range(nf):
        assert same(xm[f], x1[f]), f
    for f in range(nf):
        plan.upload_x(xs[f], random_x(shape, 300 + f))
    plan.fwd_many(xs, ks)
    km = [plan.download_k(b) for b in ks]
    for f in range(nf):
        plan.fwd(xs[f], ks[f])
        assert same(km[f], plan.download_k(ks[f])), f


@pytest.mark.parametrize("shape", [(16, 16), (8, 6, 8)], ids=str)
def test_deriv_and_project(shape):
    hp = hip()
    nd = len(shape)
    L = LENGTHS[nd]
    case("deriv/project", shape)
    plan = hp.get_plan(shape, L)
    comps = [npth.Component(shape, L) for _ in range(nd)]
    ks = [random_k(shape, 10 + i, L) for i in range(nd)]
    bufs = []
    for c, k in zip(comps, ks):
        c['kspace'] = k
        b = plan.alloc_k()
        plan.upload_k(b, k)
        bufs.append(b)
    out = plan.alloc_k()
    for d, name in enumerate('xyz'[:nd]):
        plan.deriv(bufs[0], out, d)
        assert rel_l2(plan.download_k(out), comps[0].deriv(name)) < 1e-15
    # projection against the oracle's divX / laplace_solve / deriv sequence
    ph = npth.IncompressibleHydro(shape, L)
    ph.aux = npth.State(shape, L, ph.aux_list)
    st = ph.create_fields()
    for i in range(nd):
        st['u'][i]['kspace'] = ks[i]
    ph.pressure_projection(None, st)
    plan.project(bufs)
    for i in range(nd):
        assert rel_l2(plan.download_k(bufs[i]), st['u'][i]['kspace']) < 1e-14


@pytest.mark.parametrize("ndim", [2, 3])
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
@pytest.mark.parametrize("order", [1, 2])
def test_update(ndim, kind, order):
    hp = hip()
    shape = (32, 32) if ndim == 2 else (16, 12, 16)
    L = LENGTHS[ndim]
    case("update", shape, kind, order)
    plan = hp.get_plan(shape, L)
    comp = npth.Component(shape, L)
    y, f1, f2 = (random_k(shape, s, L) for s in (20, 21, 22))
    coeff, dt = 0.37, 0.013
    IF = coeff * comp.k2() ** order
    ref = np.empty_like(y)
    if kind == 0:
        npth.euler(y, ref, f1, dt)
    elif kind == 1:
        npth.etd1(y, ref, f1, -IF, dt, ndim)
    elif kind == 2:
        npth.etd2rk1(y, ref, f1, f2, -IF, dt, ndim)
    else:
        npth.etd2rk2(y, ref, f1, f2, -IF, dt, ndim)
    b = [plan.alloc_k() for _ in range(4)]
    for buf, arr in zip(b, (y, f1, f2)):
        plan.upload_k(buf, arr)
    plan.update(kind, b[0], b[3], b[1], b[2], coeff, order, dt)
    assert rel_l2(plan.download_k(b[3]), ref) < 1e-14
    # in place
    plan.update(kind, b[0], b[0], b[1], b[2], coeff, order, dt)
    assert rel_l2(plan.download_k(b[0]), ref) < 1e-14


RHS_CASES = [
    ('hydro', (16, 16), {'nu': 1e-2}),
    ('hydro', (64, 48), {'nu': 1e-3}),
    ('hydro', (16, 12, 8), {'nu': 1e-2}),
    ('hydro', (32, 32, 32), {}),
    ('boussinesq', (32, 32), {'boussinesq_direction': 'y', 'g': 1.3, 'alpha_t': 0.7, 'beta': 1.1}),
    ('boussinesq', (16, 8, 16), {'boussinesq_direction': 'z', 'kappa': 1e-3}),
    ('mhd', (32, 32), {'rho0': 1.3}),
    ('mhd', (16, 16, 16), {'rho0': 0.8}),
    ('mhd', (8, 12, 16), {}),
    # size-specialised 3D kernels (every extent in 64..512): the x-row kernels
    # of 6, 10 and 12 fields, the spectral z pass and the z / y column passes
    ('hydro', (64, 64, 64), {'nu': 1e-3}),
    ('boussinesq', (64, 64, 64), {'boussinesq_direction': 'z', 'g': 1.1, 'alpha_t': 0.9, 'beta': 0.8}),
    ('boussinesq', (64, 128, 64), {'boussinesq_direction': 'x'}),
    ('mhd', (64, 64, 64), {'rho0': 1.2}),
    ('mhd', (128, 64, 64), {}),
    ('mhd', (64, 64, 128), {'rho0': 0.9}),
    # the n = 256 kernels of the 256^3 MHD workload, one axis at a time
    ('mhd', (64, 64, 256), {}),
    ('mhd', (64, 256, 64), {'rho0': 1.1}),
    ('mhd', (256, 64, 64), {}),
]


@pytest.mark.parametrize("kind,shape,p", RHS_CASES, ids=lambda v: str(v))
def test_rhs(kind, shape, p):
    hp = hip()
    nd = len(shape)
    L = LENGTHS[nd]
    case("rhs", kind, shape, p)
    kf = kfields(kind, shape, 30, L)
    ref = npth.rhs(kind, shape, kf, L, **dict(p))
    plan = hp.get_plan(shape, L)
    st = upload_fields(plan, kind, kf)
    dv = [plan.alloc_k() for _ in st]
    plan.rhs(params_struct(hp, kind, nd, p), st, dv)
    got = download_fields(plan, kind, nd, dv)
    for name in FIELDS[kind]:
        for c in range(len(ref[name])):
            assert rel_l2(got[name][c], ref[name][c]) < 1e-12, (name, c)


STEP_CASES = [
    ('hydro', (32, 32), {'nu': 5e-3}, 'RK2mid', 1e-2, 10),
    ('hydro', (32, 32), {'nu': 5e-3, 'viscosity_order': 2}, 'RK2trap', 1e-2, 10),
    ('hydro', (16, 16, 16), {'nu': 1e-2}, 'RK2mid', 1e-2, 5),
    ('boussinesq', (32, 32), {'boussinesq_direction': 'y', 'nu': 1e-3, 'kappa': 2e-3}, 'RK2mid', 1e-2, 10),
    ('mhd', (32, 32), {'nu': 1e-3, 'eta': 2e-3}, 'RK2mid', 1e-2, 10),
    ('mhd', (16, 16, 16), {'nu': 1e-3, 'eta': 1e-3}, 'RK2mid', 5e-3, 5),
    ('mhd', (16, 16, 16), {'nu': 1e-3, 'eta': 1e-3}, 'RK2trap', 5e-3, 5),
    ('hydro', (64, 64, 64), {'nu': 2e-3}, 'RK2trap', 5e-3, 3),
    ('boussinesq', (64, 64, 64), {'boussinesq_direction': 'y', 'nu': 1e-3, 'kappa': 2e-3}, 'RK2trap', 5e-3, 3),
    ('mhd', (64, 64, 64), {'nu': 1e-3, 'eta': 2e-3}, 'RK2mid', 5e-3, 3),
    ('mhd', (64, 64, 64), {'nu': 1e-3, 'eta': 1e-3, 'viscosity_order': 2}, 'RK2trap', 2e-3, 3),
]


@pytest.mark.parametrize("kind,shape,p,integ,dt,nsteps", STEP_CASES, ids=lambda v: str(v))
def test_steps(kind, shape, p, integ, dt, nsteps):
    hp = hip()
    nd = len(shape)
    L = (2 * np.pi,) * nd
    case("steps", kind, shape, p, integ, nsteps)
    kf = kfields(kind, shape, 40, L)
    ref = npth.run(kind, shape, kf, dt, nsteps, integ, L, **dict(p))
    plan = hp.get_plan(shape, L)
    st = upload_fields(plan, kind, kf)
    stage = [plan.alloc_k() for _ in st]
    d1 = [plan.alloc_k() for _ in st]
    d2 = [plan.alloc_k() for _ in st]
    integrator = hp.RK2MID if integ == 'RK2mid' else hp.RK2TRAP
    plan.step(integrator, params_struct(hp, kind, nd, p), st, stage, d1, d2, nsteps, dt)
    got = download_fields(plan, kind, nd, st)
    for name in FIELDS[kind]:
        for c in range(len(ref[name])):
            assert rel_l2(got[name][c], ref[name][c]) < 1e-12, (name, c)


@pytest.mark.parametrize("kind,shape", [('hydro', (64, 64)), ('mhd', (16, 16, 16))], ids=str)
def test_graph_replays_match_eager(kind, shape):
    """Each graph replay, with host work and syncs between calls, equals eager launches bit for bit."""
    hp = hip()
    nd = len(shape)
    L = (2 * np.pi,) * nd
    case("graph replay", kind, shape)
    kf = kfields(kind, shape, 95, L)
    prm = params_struct(hp, kind, nd, {'nu': 1e-3, 'eta': 1e-3})
    runs = []
    for graph in (False, True):
        plan = hp.Plan(shape, L)
        plan.set_graph(graph)
        st = upload_fields(plan, kind, kf)
        bufs = [[plan.alloc_k() for _ in st] for _ in range(3)]
        snaps = []
        for call in range(5):
            plan.step(hp.RK2MID, prm, st, bufs[0], bufs[1], bufs[2], 2, 1e-3)
            snaps.append([plan.download_k(b) for b in st])
            _ = np.linalg.norm(snaps[-1][0])      # host work between replays
        runs.append(snaps)
    for a, b in zip(*runs):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("kind,shape", [('hydro', (64, 64)), ('mhd', (16, 16, 16))], ids=str)
def test_graph_after_workspace_growth(kind, shape):
    """A captured step, then a multi-field transform that reallocates the
    plan's workspace, then more steps: equal to eager launches bit for bit
    (the graph must not replay onto the freed workspace)."""
    hp = hip()
    nd = len(shape)
    L = (2 * np.pi,) * nd
    case("graph after workspace growth", kind, shape)
    kf = kfields(kind, shape, 97, L)
    prm = params_struct(hp, kind, nd, {'nu': 1e-3, 'eta': 1e-3})
    runs = []
    for graph in (False, True):
        plan = hp.Plan(shape, L)
        plan.set_graph(graph)
        st = upload_fields(plan, kind, kf)
        bufs = [[plan.alloc_k() for _ in st] for _ in range(3)]
        snaps = []
        plan.step(hp.RK2MID, prm, st, bufs[0], bufs[1], bufs[2], 2, 1e-3)
        snaps.append([plan.download_k(b) for b in st])
        # 14 fields: more than any right-hand side's workspace
        xs = [plan.alloc_x() for _ in range(14)]
        ks = [plan.alloc_k() for _ in range(14)]
        for i in range(14):
            plan.copy(ks[i], st[i % len(st)])
        plan.bwd_many(ks, xs)
        plan.fwd_many(xs, ks)
        for _ in range(2):
            plan.step(hp.RK2MID, prm, st, bufs[0], bufs[1], bufs[2], 2, 1e-3)
            snaps.append([plan.download_k(b) for b in st])
        snaps.append([plan.download_k(b) for b in ks])
        runs.append(snaps)
    for a, b in zip(*runs):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_single_steps_with_changing_dt():
    """Single steps whose dt changes every call (a CFL condition) run eagerly;
    once a step repeats it is captured and replayed.  Every state equals the
    plan without graphs bit for bit, and the last one the oracle."""
    hp = hip()
    shape = (32, 32)
    L = (2 * np.pi,) * 2
    case("single steps, changing dt", shape)
    kf = kfields('hydro', shape, 98, L)
    prm = params_struct(hp, 'hydro', 2, {'nu': 2e-3})
    dts = [1e-2, 9e-3, 8.5e-3, 8e-3, 8e-3, 8e-3, 8e-3, 7e-3, 7e-3]
    runs = []
    for graph in (False, True):
        plan = hp.Plan(shape, L)
        plan.set_graph(graph)
        st = upload_fields(plan, 'hydro', kf)
        bufs = [[plan.alloc_k() for _ in st] for _ in range(3)]
        snaps = []
        for dt in dts:
            plan.step(hp.RK2MID, prm, st, bufs[0], bufs[1], bufs[2], 1, dt)
            snaps.append([plan.download_k(b) for b in st])
        runs.append(snaps)
    for a, b in zip(*runs):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    ph = npth.make_physics('hydro', shape, L, nu=2e-3)
    data = npth.load_state(ph, kf)
    ti = npth.RK2mid(ph)
    for dt in dts:
        ti.do_advance(data, dt)
    ref = npth.read_state(data)
    for c in range(2):
        assert rel_l2(runs[1][-1][c], ref['u'][c]) < 1e-12


@pytest.mark.parametrize("kind,shape,integ", [('hydro', (64, 48), 'RK2mid'), ('mhd', (16, 16, 16), 'RK2mid'),
                                                ('boussinesq', (32, 64), 'RK2trap')], ids=str)
def test_step_deriv2_optional(kind, shape, integ):
    """ddx_step without deriv2 buffers gives the same state bit for bit; with
    them (RK2mid) deriv2 equals ddx_rhs of the midpoint state left in stage,
    and deriv1 ddx_rhs of the last step's start."""
    hp = hip()
    nd = len(shape)
    L = (2 * np.pi,) * nd
    case("deriv2 optional", kind, shape, integ)
    kf = kfields(kind, shape, 99, L)
    p = {'nu': 1e-3, 'eta': 2e-3, 'kappa': 1e-3}
    prm = params_struct(hp, kind, nd, p)
    integrator = hp.RK2MID if integ == 'RK2mid' else hp.RK2TRAP
    out = []
    for keep in (True, False):
        plan = hp.Plan(shape, L)
        st = upload_fields(plan, kind, kf)
        bufs = [[plan.alloc_k() for _ in st] for _ in range(3)]
        plan.step(integrator, prm, st, bufs[0], bufs[1], bufs[2] if keep else None, 3, 2e-3)
        y3 = [plan.download_k(b) for b in st]
        out.append(y3)
        if keep and integ == 'RK2mid':
            d2 = [plan.download_k(b) for b in bufs[2]]
            chk = [plan.alloc_k() for _ in st]
            plan.rhs(prm, bufs[0], chk)
            for a, b in zip(d2, [plan.download_k(b) for b in chk]):
                assert np.array_equal(a, b)
            # deriv1 = F(y_2): replay two steps from the start
            st2 = upload_fields(plan, kind, kf)
            plan.step(integrator, prm, st2, bufs[0], bufs[1], bufs[2], 2, 2e-3)
            plan.rhs(prm, st2, chk)
            want = [plan.download_k(b) for b in chk]
            plan.step(integrator, prm, st2, bufs[0], bufs[1], bufs[2], 1, 2e-3)
            for a, b in zip([plan.download_k(b) for b in bufs[1]], want):
                assert np.array_equal(a, b)
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("shape,integ", [((32, 32), 'RK2mid'), ((16, 16, 16), 'RK2trap'),
                                         ((64, 48), 'RK2mid')], ids=str)
def test_steps_with_constant_forcing(shape, integ):
    """ddx_step with a forcing array equals the oracle with a state-independent
    VelocityForcing function; a second call with other forcing buffers re-keys the graph."""
    hp = hip()
    nd = len(shape)
    L = (2 * np.pi,) * nd
    case("forced steps", shape, integ)
    kf = kfields('hydro', shape, 50, L)
    fk = [0.3 * random_k(shape, 60 + i, L) for i in range(nd)]
    p = {'nu': 2e-3}
    plan = hp.get_plan(shape, L)
    prm = params_struct(hp, 'hydro', nd, p)
    integrator = hp.RK2MID if integ == 'RK2mid' else hp.RK2TRAP
    for scale in (1.0, -0.5):
        fs = [scale * f for f in fk]
        ref = npth.run('hydro', shape, kf, 1e-2, 6, integ, L, forcing=fs, **p)
        st = upload_fields(plan, 'hydro', kf)
        fb = []
        for f in fs:
            b = plan.alloc_k()
            plan.upload_k(b, f)
            fb.append(b)
        bufs = [[plan.alloc_k() for _ in st] for _ in range(3)]
        plan.step(integrator, prm, st, bufs[0], bufs[1], bufs[2], 3, 1e-2, forcing=fb)
        plan.step(integrator, prm, st, bufs[0], bufs[1], bufs[2], 3, 1e-2, forcing=fb)
        got = download_fields(plan, 'hydro', nd, st)
        for c in range(nd):
            assert rel_l2(got['u'][c], ref['u'][c]) < 1e-12, (scale, c)


@pytest.mark.parametrize("kind,shape,integ", [('hydro', (64, 64), 'RK2mid'),
                                                ('boussinesq', (64, 128), 'RK2trap'),
                                                ('mhd', (64, 48), 'RK2mid'),
                                                ('hydro', (64, 64), 'RK2trap')], ids=str)
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_fused_y_epilogue_matches_separate(kind, shape, integ, precision):
    """2D steps on 64-point y axes run the forward y pass inside the epilogue;
    the states equal those of separate launches (DDX_NO_FUSED_YEPI=1) bit for
    bit, and stay within the oracle's tolerance."""
    import os
    hp = hip()
    nd = len(shape)
    L = (2 * np.pi,) * nd
    case("fused y epilogue", kind, shape, integ, precision)
    prec = hp.F64 if precision == "f64" else hp.F32
    p = {'nu': 2e-3, 'kappa': 1e-3, 'eta': 3e-3, 'boussinesq_direction': 'y'}
    kf = kfields(kind, shape, 75, L)
    integrator = hp.RK2MID if integ == 'RK2mid' else hp.RK2TRAP
    out = []
    for sep in (False, True):
        if sep:
            os.environ["DDX_NO_FUSED_YEPI"] = "1"
        try:
            plan = hp.Plan(shape, L, prec)
        finally:
            os.environ.pop("DDX_NO_FUSED_YEPI", None)
        st = upload_fields(plan, kind, kf)
        bufs = [[plan.alloc_k() for _ in st] for _ in range(3)]
        plan.step(integrator, params_struct(hp, kind, nd, p), st, bufs[0], bufs[1], bufs[2], 4, 5e-3)
        out.append(download_fields(plan, kind, nd, st))
    for name in FIELDS[kind]:
        for c in range(len(out[0][name])):
            assert np.array_equal(out[0][name][c], out[1][name][c]), (name, c)
    if precision == "f64":
        ref = npth.run(kind, shape, kf, 5e-3, 4, integ, L, **dict(p))
        for name in FIELDS[kind]:
            for c in range(len(ref[name])):
                assert rel_l2(out[0][name][c], ref[name][c]) < 1e-12, (name, c)


ZSPLIT_CASES = [('mhd', (64, 64, 64), {}), ('hydro', (32, 256, 64), {}),
                ('boussinesq', (256, 32, 32), {'boussinesq_direction': 'z'}),
                ('boussinesq', (16, 256, 16), {'boussinesq_direction': 'x'}),
                ('mhd', (12, 20, 24), {}), ('hydro', (10, 8, 6), {})]


@pytest.mark.parametrize("kind,shape,extra", ZSPLIT_CASES, ids=str)
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_z_split_matches_unsplit(kind, shape, extra, precision):
    """FP64 3D right-hand sides form the fields with a d/dz term (wx, wy, Jx,
    Jy, dz T) and the plain copies before the z transforms, and the others
    (wz, Jz, dx T, dy T) in the y pass from plain z-pass outputs.  The states
    agree with the unsplit dataflow (DDX_NO_ZSPLIT=1, every derived field
    formed before the z pass) to rounding, and with the oracle within 1e-12.
    FP32 plans do not split: the two runs are then identical."""
    import os
    hp = hip()
    L = LENGTHS[3]
    case("z split", kind, shape, precision)
    prec = hp.F64 if precision == "f64" else hp.F32
    p = {'nu': 2e-3, 'kappa': 1e-3, 'eta': 3e-3}
    p.update(extra)
    kf = kfields(kind, shape, 91, L)
    out = []
    for unsplit in (False, True):
        if unsplit:
            os.environ["DDX_NO_ZSPLIT"] = "1"
        try:
            plan = hp.Plan(shape, L, prec)
        finally:
            os.environ.pop("DDX_NO_ZSPLIT", None)
        st = upload_fields(plan, kind, kf)
        bufs = [[plan.alloc_k() for _ in st] for _ in range(3)]
        plan.step(hp.RK2MID, params_struct(hp, kind, 3, p), st, bufs[0], bufs[1], bufs[2], 3, 2e-3)
        out.append(download_fields(plan, kind, 3, st))
    for name in FIELDS[kind]:
        for c in range(len(out[0][name])):
            if precision == "f64":
                assert rel_l2(out[0][name][c], out[1][name][c]) < 1e-14, (name, c)
            else:
                assert np.array_equal(out[0][name][c], out[1][name][c]), (name, c)
    if precision == "f64":
        ref = npth.run(kind, shape, kf, 2e-3, 3, 'RK2mid', L, **dict(p))
        for name in FIELDS[kind]:
            for c in range(len(ref[name])):
                assert rel_l2(out[0][name][c], ref[name][c]) < 1e-12, (name, c)


FUSE_CASES = [('hydro', (32, 32), 'RK2mid'), ('hydro', (16, 16, 16), 'RK2trap'),
              ('boussinesq', (32, 48), 'RK2mid'), ('boussinesq', (16, 8, 16), 'RK2trap'),
              ('mhd', (16, 16, 16), 'RK2mid'), ('mhd', (32, 32), 'RK2trap'),
              ('boussinesq', (64, 64, 64), 'RK2trap'), ('mhd', (64, 64, 64), 'RK2mid'),
              ('mhd', (16, 32, 16), 'RK2trap')]


@pytest.mark.parametrize("kind,shape,integ", FUSE_CASES, ids=str)
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_fused_update_matches_separate_sweeps(kind, shape, integ, precision):
    """ETD sweeps fused into the RHS epilogues give bit-identical states to
    separate update launches (DDX_NO_FUSED_UPDATE=1)."""
    import os
    hp = hip()
    nd = len(shape)
    L = (2 * np.pi,) * nd
    case("fused update", kind, shape, integ, precision)
    prec = hp.F64 if precision == "f64" else hp.F32
    p = {'nu': 2e-3, 'kappa': 1e-3, 'eta': 3e-3, 'viscosity_order': 1,
         'boussinesq_direction': 'y' if nd == 2 else 'z'}
    kf = kfields(kind, shape, 70, L)
    integrator = hp.RK2MID if integ == 'RK2mid' else hp.RK2TRAP
    out = []
    for fused in (True, False):
        if not fused:
            os.environ["DDX_NO_FUSED_UPDATE"] = "1"
        try:
            plan = hp.Plan(shape, L, prec)
        finally:
            os.environ.pop("DDX_NO_FUSED_UPDATE", None)
        st = upload_fields(plan, kind, kf)
        bufs = [[plan.alloc_k() for _ in st] for _ in range(3)]
        plan.step(integrator, params_struct(hp, kind, nd, p), st, bufs[0], bufs[1], bufs[2], 4, 5e-3)
        out.append([plan.download_k(b) for b in st])
    for a, b in zip(*out):
        assert np.array_equal(a, b)


RADIX_SHAPES = [(64, 64), (128, 128), (256, 256), (512, 512), (16, 64, 128), (8, 256, 64),
                (512, 16, 64), (4, 8, 512)]


@pytest.mark.parametrize("shape", RADIX_SHAPES, ids=str)
def test_radix_variant_transforms(shape):
    """The radix variant of the size-specialised rows and columns
    (DDX_TUNE_RADIX = 1: rows 8.8 / 8.4.4 / 8.8.4 / 8.4.4.4, columns 4.4.4 /
    8.4.4 / 8.8.4 / 16.8.4) transforms like the oracle."""
    hp = hip()
    L = LENGTHS[len(shape)]
    case("radix variant transforms", shape)
    plan = hp.Plan(shape, L)
    plan.set_tuning(hp.TUNE_RADIX, 1)
    comp = npth.Component(shape, L)
    x = random_x(shape, 31)
    comp['xspace'] = x
    ref = comp['kspace']
    xd, kc = plan.alloc_x(), plan.alloc_k()
    plan.upload_x(xd, x)
    plan.fwd(xd, kc)
    assert rel_l2(plan.download_k(kc), ref) < 1e-14
    k = random_k(shape, 32, L)
    comp['kspace'] = k
    xr = comp['xspace']
    plan.upload_k(kc, k)
    plan.bwd(kc, xd)
    assert rel_l2(plan.download_x(xd), xr) < 1e-14


@pytest.mark.parametrize("kind,shape", [('hydro', (64, 64)), ('hydro', (128, 128)), ('boussinesq', (256, 64)),
                                        ('mhd', (16, 64, 128)), ('mhd', (8, 64, 512)), ('hydro', (512, 16, 64))],
                         ids=str)
def test_radix_variant_steps(kind, shape):
    """Three RK2mid steps with the radix variant: the oracle to 1e-12, the
    built-in plans to rounding."""
    hp = hip()
    nd = len(shape)
    L = (2 * np.pi,) * nd
    case("radix variant steps", kind, shape)
    p = {'nu': 1e-3, 'eta': 1e-3, 'kappa': 1e-3}
    if kind == 'boussinesq':
        p['boussinesq_direction'] = 'y' if nd == 2 else 'z'
    kf = kfields(kind, shape, 33, L)
    ref = npth.run(kind, shape, kf, 2e-3, 3, 'RK2mid', L, **dict(p))
    prm = params_struct(hp, kind, nd, p)
    got = []
    for radix in (1, 0):
        plan = hp.Plan(shape, L)
        plan.set_tuning(hp.TUNE_RADIX, radix)
        st = upload_fields(plan, kind, kf)
        bufs = [[plan.alloc_k() for _ in st] for _ in range(3)]
        plan.step(hp.RK2MID, prm, st, bufs[0], bufs[1], None, 3, 2e-3)
        got.append(download_fields(plan, kind, nd, st))
    for name in FIELDS[kind]:
        for c in range(len(ref[name])):
            assert rel_l2(got[0][name][c], ref[name][c]) < 1e-12, (name, c)
            assert rel_l2(got[0][name][c], got[1][name][c]) < 1e-13, (name, c)


def test_radix_variant_fp32():
    """FP32 with the radix variant against the FP64 oracle (128^2 hydro)."""
    hp = hip()
    shape = (128, 128)
    L = (2 * np.pi,) * 2
    case("radix variant fp32", shape)
    kf = kfields('hydro', shape, 34, L)
    ref = npth.run('hydro', shape, kf, 2e-3, 3, 'RK2mid', L, nu=1e-3)
    plan = hp.Plan(shape, L, hp.F32)
    plan.set_tuning(hp.TUNE_RADIX, 1)
    st = upload_fields(plan, 'hydro', kf)
    bufs = [[plan.alloc_k() for _ in st] for _ in range(3)]
    plan.step(hp.RK2MID, params_struct(hp, 'hydro', 2, {'nu': 1e-3}), st, bufs[0], bufs[1], None, 3, 2e-3)
    for c in range(2):
        assert rel_l2(plan.download_k(st[c]), ref['u'][c]) < 2e-6


@pytest.mark.parametrize("f64", [True, False], ids=["f64", "f32"])
@pytest.mark.parametrize("shape", [(64, 64), (48, 64), (30, 20), (16, 16, 16), (12, 10, 16),
                                   (4, 6, 128)], ids=str)
def test_bwd_combine(shape, f64):
    """ddx_bwd_combine against the oracle's deriv() then backward(): the
    field itself, one derivative with a coefficient, and curl-type
    differences of two sources (the vorticity of curlX, physics.py:374-405)."""
    hp = hip()
    nd = len(shape)
    L = LENGTHS[nd]
    case("bwd_combine", shape, "f64" if f64 else "f32")
    plan = hp.Plan(shape, L, hp.F64 if f64 else hp.F32)
    srcs = [random_k(shape, 400 + i, L) for i in range(nd)]
    ks = []
    for a in srcs:
        b = plan.alloc_k()
        plan.upload_k(b, a)
        ks.append(b)
    names = ['x', 'y', 'z'][:nd]
    recipes = [[(0, -1, 1.0)], [(1, 0, -2.5)],
               [(1, 0, 1.0), (0, 1, -1.0)]]
    if nd == 3:
        recipes += [[(2, 1, 1.0), (1, 2, -1.0)], [(0, 2, 0.5), (2, -1, 3.0)]]
    xs = [plan.alloc_x() for _ in recipes]
    plan.bwd_combine(ks, recipes, xs)
    comps = []
    for a in srcs:
        c = npth.Component(shape, L)
        c['kspace'] = a
        comps.append(c)
    for q, r in enumerate(recipes):
        acc = np.zeros_like(srcs[0])
        for s_, d_, c_ in r:
            term = comps[s_]['kspace'].copy() if d_ < 0 else comps[s_].deriv(names[d_]).copy()
            acc = acc + c_ * term
        out = npth.Component(shape, L)
        out['kspace'] = acc
        want = out['xspace']
        got = plan.download_x(xs[q])
        assert rel_l2(got, want) < (1e-13 if f64 else 2e-6), (q, rel_l2(got, want))


@pytest.mark.parametrize("kind,shape,p,integ", [
    ('hydro', (64, 64), {'nu': 2e-3}, 'RK2mid'),
    ('hydro', (30, 20), {'nu': 2e-3}, 'RK2trap'),
    ('hydro', (16, 16, 16), {'nu': 5e-3}, 'RK2mid'),
    ('boussinesq', (32, 64), {'nu': 1e-3, 'kappa': 1e-3, 'boussinesq_direction': 'y'}, 'RK2mid'),
    ('mhd', (8, 16, 16), {'nu': 1e-3, 'eta': 1e-3}, 'RK2trap'),
], ids=str)
def test_steps_with_x_space_forcing(kind, shape, p, integ):
    """ddx_step_xforced with an x-space forcing f equals ddx_step with the
    k-space forcing forward(f) (1/N, dealiased) to rounding: the forcing
    joins the products before their forward transform instead."""
    hp = hip()
    nd = len(shape)
    L = LENGTHS[nd]
    case("x-space forcing", kind, shape, integ)
    plan = hp.Plan(shape, L)
    kf = kfields(kind, shape, 90, L)
    prm = params_struct(hp, kind, nd, p)
    nforce = nd + (1 if kind == 'boussinesq' else 0)
    xf, kfz = [], []
    for i in range(nforce):
        xb = plan.alloc_x()
        plan.upload_x(xb, 0.3 * random_x(shape, 500 + i))
        kb = plan.alloc_k()
        plan.fwd(xb, kb)
        xf.append(xb)
        kfz.append(kb)
    nst = sum(nd if n in ('u', 'B') else 1 for n in FIELDS[kind])
    xforcing = xf + [None] * (nst - nforce)
    kforcing = kfz + [None] * (nst - nforce)
    integ_c = hp.RK2MID if integ == 'RK2mid' else hp.RK2TRAP
    results = []
    for mode in ('k', 'x'):
        st = upload_fields(plan, kind, kf)
        stage = [plan.alloc_k() for _ in st]
        d1 = [plan.alloc_k() for _ in st]
        d2 = [plan.alloc_k() for _ in st]
        if mode == 'k':
            plan.step(integ_c, prm, st, stage, d1, d2, 3, 4e-3, forcing=kforcing)
        else:
            plan.step(integ_c, prm, st, stage, d1, d2, 3, 4e-3, xforcing=xforcing)
        results.append(download_fields(plan, kind, nd, st))
    for name in results[0]:
        for a, b in zip(results[0][name], results[1][name]):
            assert rel_l2(b, a) < 1e-13, (name, rel_l2(b, a))
    if kind == 'mhd':
        # the induction equation has no forcing hook
        st = upload_fields(plan, kind, kf)
        bad = xf + [xf[0]] + [None] * (nst - nforce - 1)
        with pytest.raises(hp.DDXError):
            plan.step(integ_c, prm, st, [plan.alloc_k() for _ in st], [plan.alloc_k() for _ in st],
                      None, 1, 4e-3, xforcing=bad)


XIN_RECIPES = {
    'hydro': [[(0, -1, 1.0)], [(1, -1, 1.0)], [(1, 0, 1.0), (0, 1, -1.0)]],
    'boussinesq': [[(0, -1, 1.0)], [(1, -1, 1.0)], [(1, 0, 1.0), (0, 1, -1.0)],
                   [(2, -1, 1.0)], [(2, 0, 1.0)], [(2, 1, 1.0)]],
    'mhd': [[(0, -1, 1.0)], [(1, -1, 1.0)], [(1, 0, 1.0), (0, 1, -1.0)],
            [(2, -1, 1.0)], [(3, -1, 1.0)], [(3, 0, 1.0), (2, 1, -1.0)]],
}


@pytest.mark.parametrize("kind,shape,p,integ,prec,batch", [
    ('hydro', (64, 64), {'nu': 1e-3}, 'RK2mid', 'f64', None),
    ('hydro', (32, 128), {'nu': 2e-3}, 'RK2trap', 'f64', None),
    ('hydro', (64, 64), {'nu': 1e-3}, 'RK2mid', 'f64', 3),
    ('hydro', (48, 256), {'nu': 1e-3}, 'RK2mid', 'f32', None),
    ('boussinesq', (64, 64), {'nu': 1e-3, 'kappa': 1e-3, 'boussinesq_direction': 'y'}, 'RK2mid', 'f64', None),
    ('mhd', (16, 128), {'nu': 1e-3, 'eta': 1e-3}, 'RK2trap', 'f64', None),
    ('hydro', (30, 20), {'nu': 2e-3}, 'RK2mid', 'f64', None),
], ids=str)
def test_step_with_x_space_inputs(kind, shape, p, integ, prec, batch):
    """ddx_step_xin, given the backward transforms of the first right-hand
    side's inputs (ddx_bwd_combine of the state with the physics' recipes),
    equals ddx_step_xforced to rounding (the transform pairs rows where the
    step pairs fields); and it reads them: inputs of another state change the
    result.  An x extent without the size-specialised x-row kernel (20)
    ignores them, bit for bit."""
    hp = hip()
    nd = len(shape)
    L = LENGTHS[nd]
    case("x-space step inputs", kind, shape, integ, prec, batch)
    precision = hp.F64 if prec == 'f64' else hp.F32
    plan = hp.Plan(shape, L, precision, batch=batch) if batch else hp.Plan(shape, L, precision)
    prm = params_struct(hp, kind, nd, p)
    integ_c = hp.RK2MID if integ == 'RK2mid' else hp.RK2TRAP

    def upload(seed):
        kf = kfields(kind, shape, seed, L)
        bufs = []
        for name in FIELDS[kind]:
            for arr in kf[name]:
                b = plan.alloc_k()
                plan.upload_k(b, np.stack([arr] * batch) if batch else arr)
                bufs.append(b)
        return bufs

    nforce = nd + (1 if kind == 'boussinesq' else 0)
    xf = []
    for i in range(nforce):
        xb = plan.alloc_x()
        xs = (batch,) + tuple(shape) if batch else shape
        plan.upload_x(xb, 0.3 * random_x(xs, 700 + i))
        xf.append(xb)
    ns = sum(nd if n in ('u', 'B') else 1 for n in FIELDS[kind])
    xforcing = xf + [None] * (ns - nforce)
    recipes = XIN_RECIPES[kind]
    results = []
    for mode in ('plain', 'xin', 'other'):
        st = upload(41)
        xin = [plan.alloc_x() for _ in recipes]
        plan.bwd_combine(upload(43) if mode == 'other' else st, recipes, xin)
        stage = [plan.alloc_k() for _ in st]
        d1 = [plan.alloc_k() for _ in st]
        if mode == 'plain':
            plan.step(integ_c, prm, st, stage, d1, None, 1, 4e-3, xforcing=xforcing)
        else:
            plan.step(integ_c, prm, st, stage, d1, None, 1, 4e-3, xforcing=xforcing, xin=xin)
        results.append([plan.download_k(b) for b in st])
    tol = 1e-13 if prec == 'f64' else 1e-5
    applies = shape[-1] in (64, 128, 256, 512)
    for c, (a, b, o) in enumerate(zip(*results)):
        e = rel_l2(b, a)
        print("  component %d: xin vs plain %.2e, other state's xin vs plain %.2e" % (c, e, rel_l2(o, a)))
        if applies:
            assert e < tol, (c, e)
        else:
            assert np.array_equal(a, b) and np.array_equal(a, o)
    if applies:
        assert max(rel_l2(o, a) for a, o in zip(results[0], results[2])) > 1e-6
    # one step only; 2D only
    st = upload(41)
    xin = [plan.alloc_x() for _ in recipes]
    with pytest.raises(hp.DDXError):
        plan.step(integ_c, prm, st, [plan.alloc_k() for _ in st], [plan.alloc_k() for _ in st], None,
                  2, 4e-3, xforcing=xforcing, xin=xin)


def test_step_x_space_inputs_3d_refused():
    hp = hip()
    case("x-space step inputs refused in 3D")
    plan = hp.Plan((16, 16, 16), LENGTHS[3])
    prm = params_struct(hp, 'hydro', 3, {'nu': 1e-3})
    st = [plan.alloc_k() for _ in range(3)]
    for b in st:
        plan.zero(b)
    xin = [plan.alloc_x() for _ in range(6)]
    with pytest.raises(hp.DDXError):
        plan.step(hp.RK2MID, prm, st, [plan.alloc_k() for _ in st], [plan.alloc_k() for _ in st], None,
                  1, 1e-3, xin=xin)
