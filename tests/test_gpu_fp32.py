"""FP32 transforms and nonlinear products against the FP64 oracle (GPU).

FP64 stays in the twiddle generation, the projection / epilogue and the ETD
phi functions; the transforms and x-space products run in FP32.  Each test
prints the per-field relative L2 error it measured.
"""

import numpy as np
import pytest

from helpers import kfields, random_k, random_x, rel_l2
from oracle import numpy_path as npth

pytestmark = pytest.mark.gpu

FIELDS = {'hydro': ['u'], 'boussinesq': ['u', 'T'], 'mhd': ['u', 'B']}


def hip():
    from dedalus import _hip
    _hip.init()
    return _hip


def case(*a):
    print("case:", *a, flush=True)


def up(plan, kind, kf):
    bufs = []
    for name in FIELDS[kind]:
        for a in kf[name]:
            b = plan.alloc_k()
            plan.upload_k(b, a)
            bufs.append(b)
    return bufs


def down(plan, kind, nd, bufs):
    out, i = {}, 0
    for name in FIELDS[kind]:
        nc = nd if name in ('u', 'B') else 1
        out[name] = [plan.download_k(bufs[i + c]) for c in range(nc)]
        i += nc
    return out


@pytest.mark.parametrize("shape", [(64, 64), (48, 64), (512, 512), (32, 32, 32), (16, 12, 8)], ids=str)
def test_fp32_transforms(shape):
    hp = hip()
    L = (2 * np.pi,) * len(shape)
    case("fp32 transforms", shape)
    plan = hp.get_plan(shape, L, hp.F32)
    c = npth.Component(shape, L)
    x = random_x(shape, 5)
    c['xspace'] = x
    kref = c['kspace'].copy()
    xd, kc = plan.alloc_x(), plan.alloc_k()
    plan.upload_x(xd, x)
    plan.fwd(xd, kc)
    ef = rel_l2(plan.download_k(kc), kref)
    plan.bwd(kc, xd)
    c['kspace'] = kref
    eb = rel_l2(plan.download_x(xd), c['xspace'])
    print("fp32 %s: forward rel L2 %.2e, round trip rel L2 %.2e" % (shape, ef, eb))
    assert ef < 2e-6 and eb < 4e-6


@pytest.mark.parametrize("kind,shape,p,dt,n", [
    ('hydro', (32, 32, 32), {'nu': 2e-3}, 5e-3, 10),
    ('mhd', (16, 16, 16), {'nu': 1e-3, 'eta': 1e-3}, 5e-3, 10),
    ('boussinesq', (64, 64), {'nu': 1e-3, 'kappa': 1e-3, 'boussinesq_direction': 'y'}, 5e-3, 10),
], ids=str)
def test_fp32_steps_error_per_field(kind, shape, p, dt, n):
    hp = hip()
    nd = len(shape)
    L = (2 * np.pi,) * nd
    case("fp32 steps", kind, shape)
    kf = kfields(kind, shape, 90, L)
    ref = npth.run(kind, shape, kf, dt, n, 'RK2mid', L, **dict(p))
    plan = hp.get_plan(shape, L, hp.F32)
    st = up(plan, kind, kf)
    stage, d1, d2 = ([plan.alloc_k() for _ in st] for _ in range(3))
    phys = {'hydro': hp.HYDRO, 'boussinesq': hp.BOUSSINESQ, 'mhd': hp.MHD}[kind]
    prm = hp.make_params(phys, nu=p.get('nu', 0.), kappa=p.get('kappa', 0.), eta=p.get('eta', 0.),
                         bdir={'x': 0, 'y': 1, 'z': 2}[p.get('boussinesq_direction', 'z')])
    plan.step(hp.RK2MID, prm, st, stage, d1, d2, n, dt)
    got = down(plan, kind, nd, st)
    for name in FIELDS[kind]:
        for c in range(len(ref[name])):
            e = rel_l2(got[name][c], ref[name][c])
            print("fp32 %s %s%d after %d steps: rel L2 vs FP64 oracle = %.2e" % (kind, name, c, n, e))
            assert e < 2e-5, (name, c, e)


def test_fp32_hydro3d_128_vs_fp64():
    """The 128^3 hydro workload, 100 steps: FP32 path against the FP64 GPU path."""
    hp = hip()
    from dedalus.init_cond.seeded import workload
    case("fp32 vs fp64 hydro3d_128 100 steps")
    w = workload('hydro3d_128')
    res = {}
    for prec in (hp.F64, hp.F32):
        plan = hp.Plan(w['shape'], w['length'], prec)
        st = up(plan, 'hydro', w['fields'])
        stage, d1, d2 = ([plan.alloc_k() for _ in st] for _ in range(3))
        prm = hp.make_params(hp.HYDRO, nu=w['params']['nu'])
        plan.step(hp.RK2MID, prm, st, stage, d1, d2, w['nsteps'], w['dt'])
        res[prec] = down(plan, 'hydro', 3, st)
    for c in range(3):
        e = rel_l2(res[hp.F32]['u'][c], res[hp.F64]['u'][c])
        print("hydro3d_128 u%d after 100 steps: FP32 vs FP64 rel L2 = %.2e" % (c, e))
        assert e < 1e-5
