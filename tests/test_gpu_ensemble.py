"""Batched 2D plans (ensembles of independent members) on the GPU.

Every member of a batched plan must equal the same field run through a
plain 2D plan bit for bit (the kernels are the same; the member index takes
the place of the z slab), and the plain plan is held to the oracle by the
kernel tests; a few members are also checked against the oracle directly.
"""

import numpy as np
import pytest

from helpers import kfields, random_k, random_x, rel_l2
from oracle import numpy_path as npth

pytestmark = pytest.mark.gpu

L2 = (2 * np.pi, 2 * np.pi)
SHAPES = [(32, 32), (64, 64), (48, 32), (128, 128), (256, 256), (512, 512)]


def hip():
    from dedalus import _hip
    _hip.init()
    return _hip


def case(*args):
    print("case:", *args, flush=True)


def nb_for(shape):
    return 3 if shape[0] * shape[1] >= 256 * 256 else 7


@pytest.mark.parametrize("shape", SHAPES, ids=str)
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_transforms_match_single_plans(shape, precision):
    hp = hip()
    nb = nb_for(shape)
    prec = hp.F64 if precision == "f64" else hp.F32
    case("batched transforms", shape, nb, precision)
    bp = hp.Plan(shape, L2, prec, batch=nb)
    sp = hp.Plan(shape, L2, prec)
    xs = np.stack([random_x(shape, 100 + m) for m in range(nb)])
    ks = np.stack([random_k(shape, 200 + m, L2) for m in range(nb)])
    xb, kb = bp.alloc_x(), bp.alloc_k()
    bp.upload_x(xb, xs)
    bp.fwd(xb, kb)
    kgot = bp.download_k(kb)
    bp.upload_k(kb, ks)
    assert np.array_equal(bp.download_k(kb), np.stack([sp.download_k(_up(sp, k)) for k in ks]))
    bp.bwd(kb, xb)
    xgot = bp.download_x(xb)
    xs1, ks1 = sp.alloc_x(), sp.alloc_k()
    for m in range(nb):
        sp.upload_x(xs1, xs[m])
        sp.fwd(xs1, ks1)
        assert np.array_equal(kgot[m], sp.download_k(ks1)), m
        sp.upload_k(ks1, ks[m])
        sp.bwd(ks1, xs1)
        assert np.array_equal(xgot[m], sp.download_x(xs1)), m


def _up(plan, k):
    b = plan.alloc_k()
    plan.upload_k(b, k)
    return b


STEP_CASES = [('hydro', (64, 64), {'nu': 1e-3}), ('hydro', (128, 128), {'nu': 5e-4}),
              ('boussinesq', (32, 48), {'boussinesq_direction': 'y', 'nu': 1e-3, 'kappa': 1e-3}),
              ('mhd', (64, 64), {'nu': 1e-3, 'eta': 1e-3}), ('hydro', (256, 256), {'nu': 1e-4})]
FIELDS = {'hydro': ['u'], 'boussinesq': ['u', 'T'], 'mhd': ['u', 'B']}


def params(hp, kind, p):
    phys = {'hydro': hp.HYDRO, 'boussinesq': hp.BOUSSINESQ, 'mhd': hp.MHD}[kind]
    return hp.make_params(phys, nu=p.get('nu', 0.), kappa=p.get('kappa', 0.), eta=p.get('eta', 0.),
                          bdir={'x': 0, 'y': 1}[p.get('boussinesq_direction', 'y')])


@pytest.mark.parametrize("kind,shape,p", STEP_CASES, ids=str)
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_steps_match_single_plans(kind, shape, p, precision):
    hp = hip()
    nb = nb_for(shape)
    prec = hp.F64 if precision == "f64" else hp.F32
    case("batched steps", kind, shape, nb, precision)
    prm = params(hp, kind, p)
    members = [kfields(kind, shape, 300 + m, L2) for m in range(nb)]
    comps = [arr for name in FIELDS[kind] for arr in members[0][name]]
    ncomp = len(comps)
    stacked = [np.stack([[a for name in FIELDS[kind] for a in mem[name]][c] for mem in members])
               for c in range(ncomp)]
    fk = np.stack([0.1 * random_k(shape, 400 + m, L2) for m in range(nb)])
    bp = hp.Plan(shape, L2, prec, batch=nb)
    st = [_up(bp, a) for a in stacked]
    fb = [_up(bp, fk)] + [None] * (ncomp - 1)
    bufs = [[bp.alloc_k() for _ in st] for _ in range(3)]
    bp.step(hp.RK2MID, prm, st, bufs[0], bufs[1], bufs[2], 4, 1e-3, forcing=fb)
    got = [bp.download_k(b) for b in st]
    sp = hp.Plan(shape, L2, prec)
    for m in range(nb):
        s1 = [_up(sp, stacked[c][m]) for c in range(ncomp)]
        f1 = [_up(sp, fk[m])] + [None] * (ncomp - 1)
        b1 = [[sp.alloc_k() for _ in s1] for _ in range(3)]
        sp.step(hp.RK2MID, prm, s1, b1[0], b1[1], b1[2], 4, 1e-3, forcing=f1)
        for c in range(ncomp):
            assert np.array_equal(got[c][m], sp.download_k(s1[c])), (m, c)


def test_members_match_oracle():
    hp = hip()
    shape, nb = (32, 32), 4
    case("batched members vs oracle", shape, nb)
    p = {'nu': 2e-3}
    members = [kfields('hydro', shape, 500 + m, L2) for m in range(nb)]
    bp = hp.Plan(shape, L2, hp.F64, batch=nb)
    st = [_up(bp, np.stack([mem['u'][c] for mem in members])) for c in range(2)]
    bufs = [[bp.alloc_k() for _ in st] for _ in range(3)]
    bp.step(hp.RK2MID, params(hp, 'hydro', p), st, bufs[0], bufs[1], bufs[2], 5, 1e-2)
    got = [bp.download_k(b) for b in st]
    for m in range(nb):
        ref = npth.run('hydro', shape, members[m], 1e-2, 5, 'RK2mid', L2, **p)
        for c in range(2):
            assert rel_l2(got[c][m], ref['u'][c]) < 1e-12, (m, c)


def test_ensemble_graph_matches_eager_and_oracle():
    """Two groups (32^2 x 3, 64^2 x 2) in FP32: replays of one-step and
    multi-step graphs equal eager group steps bit for bit; members stay within
    FP32 error of the FP64 oracle."""
    hip()
    from dedalus.ensemble import Ensemble, vorticity_member
    case("ensemble graph vs eager vs oracle")
    groups = (((32, 32), 3), ((64, 64), 2))
    runs = []
    # eager; one-step graphs; three-step graphs (4 steps = one three-step
    # replay and one one-step replay); a four-step graph (one replay per call)
    for graph in (None, 1, 3, 4):
        ens = Ensemble(groups, nu=5e-4, dt=1e-3, precision="f32")
        ens.seed_members(1000)
        if graph:
            ens.capture(steps_per_graph=graph)
            ens.seed_members(1000)
        snaps = []
        for call in range(3):
            ens.advance(4)
            ens.sync()
            snaps.append([g.read() for g in ens.groups])
        runs.append(snaps)
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            for ga, gb in zip(a, b):
                for (ua, va), (ub, vb) in zip(ga, gb):
                    assert np.array_equal(ua, ub) and np.array_equal(va, vb)
    # FP32 members vs the FP64 oracle after 12 steps
    m = 0
    for (shape, n), gsnap in zip(groups, runs[1][-1]):
        for j in range(n):
            u0 = vorticity_member(shape, L2, 1000 + m)
            ref = npth.run('hydro', shape, {'u': u0}, 1e-3, 12, 'RK2mid', L2, nu=5e-4)
            for c in range(2):
                err = rel_l2(gsnap[j][c], ref['u'][c])
                assert err < 2e-6, (shape, j, c, err)
            m += 1


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_member_reductions(precision):
    """Per-member energy and enstrophy of an ensemble from device reductions
    equal the host Parseval sums of each member's downloaded fields."""
    hip()
    from dedalus.ensemble import Ensemble
    from dedalus.init_cond.seeded import mean_square
    from dedalus.data_objects.kgrid import wavenumbers
    case("member reductions", precision)
    ens = Ensemble((((32, 32), 5), ((64, 48), 3)), nu=1e-3, dt=1e-3, precision=precision)
    ens.seed_members(2000)
    ens.advance(3)
    ens.sync()
    e = ens.energies()
    z = ens.enstrophies()
    assert e.shape == (8,) and z.shape == (8,)
    m = 0
    tol = 1e-13 if precision == "f64" else 1e-6
    for g in ens.groups:
        k, _, ksh = wavenumbers(g.shape, L2)
        for u, v in g.read():
            assert abs(e[m] - 0.5 * mean_square([u, v], g.shape)) <= tol * e[m]
            w = 1j * k['x'] * v - 1j * k['y'] * u
            assert abs(z[m] - 0.5 * mean_square([w], g.shape)) <= tol * z[m]
            m += 1


def test_autotune_keeps_results():
    """Ensemble.autotune picks kernel variants per group (radix plan, tile
    width, XCD order, fused forward y + epilogue); it restores the members'
    state bit for bit, and the tuned graph's steps equal the untuned ones
    (bit for bit with the built-in radix plan, to FP32 rounding with the
    other)."""
    hip()
    from dedalus.ensemble import Ensemble
    case("ensemble autotune")
    groups = (((64, 64), 3), ((256, 256), 2))
    runs = []
    for tune in (False, True):
        ens = Ensemble(groups, nu=5e-4, dt=1e-3, precision="f32")
        ens.seed_members(2000)
        if tune:
            before = [g.read() for g in ens.groups]
            chosen = ens.autotune(steps=2, reps=1)
            print("chosen", chosen, flush=True)
            assert set(chosen) == {(64, 64), (256, 256)}
            for gb, g in zip(before, ens.groups):
                for (ua, va), (ub, vb) in zip(gb, g.read()):
                    assert np.array_equal(ua, ub) and np.array_equal(va, vb)
        ens.capture()
        ens.seed_members(2000)
        ens.advance(4)
        ens.sync()
        runs.append([g.read() for g in ens.groups])
    for ga, gb, shape in zip(runs[0], runs[1], [g[0] for g in groups]):
        exact = chosen[shape]["radix"] == 0
        for (ua, va), (ub, vb) in zip(ga, gb):
            if exact:
                assert np.array_equal(ua, ub) and np.array_equal(va, vb)
            else:
                assert rel_l2(ua, ub) < 1e-5 and rel_l2(va, vb) < 1e-5


def test_tuning_knobs_keep_results():
    """ddx_plan_set_tuning choices: the XCD order and the fused forward y +
    epilogue give bit-identical steps (batched 64^2 and 256^2 plans, FP64 and
    FP32); the tile width does in FP32, while in FP64 its 16-column tiles
    exchange through a buffer of reals, compiled to other instruction
    sequences: results agree to rounding."""
    hp = hip()
    case("tuning knobs")
    cases = (((64, 64), hp.TUNE_FUSED_YEPI, (0, 1)), ((64, 64), hp.TUNE_XCD_ORDER, (0, 1, 2)),
             ((256, 256), hp.TUNE_XCD_ORDER, (0, 1, 2)), ((256, 256), hp.TUNE_WIDE_TILES, (0, 1)))
    nb = 3
    for prec in (hp.F64, hp.F32):
        for shape, knob, values in cases:
            members = [kfields('hydro', shape, 700 + m, L2) for m in range(nb)]
            outs = []
            for value in (-1,) + values:
                bp = hp.Plan(shape, L2, prec, batch=nb)
                bp.set_tuning(knob, value)
                st = [bp.alloc_k() for _ in range(2)]
                for c in range(2):
                    bp.upload_k(st[c], np.stack([members[m]['u'][c] for m in range(nb)]))
                bufs = [[bp.alloc_k() for _ in range(2)] for _ in range(3)]
                bp.step(hp.RK2MID, hp.make_params(hp.HYDRO, nu=1e-3), st, bufs[0], bufs[1], bufs[2], 3, 1e-3)
                outs.append([bp.download_k(b) for b in st])
            exact = not (knob == hp.TUNE_WIDE_TILES and prec == hp.F64)
            for o in outs[1:]:
                for a, b in zip(outs[0], o):
                    if exact:
                        assert np.array_equal(a, b), (shape, knob, prec)
                    else:
                        assert rel_l2(b, a) < 1e-14, (shape, knob, prec)
    with pytest.raises(hp.DDXError):
        hp.Plan((64, 64), L2).set_tuning(hp.TUNE_XCD_ORDER, 3)
