"""ddx_step_noalloc (the entry point that allocates nothing) against ddx_step (GPU).

Two calls in a row on one GPU share the static device scratch, the second on
a smaller grid than the first, so it runs on the first call's leftovers:
each result must equal ddx_step's on the same inputs bit for bit (the same
kernels on another workspace).  A grid whose workspace exceeds the scratch
is refused on the host before any launch.
"""

import os

import numpy as np
import pytest

from helpers import kfields

pytestmark = pytest.mark.gpu

FIELDS = {'hydro': ['u'], 'boussinesq': ['u', 'T'], 'mhd': ['u', 'B']}


def hip():
    from dedalus import _hip
    _hip.init()
    return _hip


def params(hp, kind, nd):
    phys = {'hydro': hp.HYDRO, 'boussinesq': hp.BOUSSINESQ, 'mhd': hp.MHD}[kind]
    return hp.make_params(phys, nu=2e-3, kappa=1e-3, eta=1.5e-3, bdir=nd - 1)


def run(hp, kind, shape, nsteps, dt, noalloc, keep_d2, env=None):
    nd = len(shape)
    L = (2 * np.pi,) * nd
    if env:
        os.environ[env] = "1"
    try:
        plan = hp.Plan(shape, L)
    finally:
        if env:
            os.environ.pop(env, None)
    kf = kfields(kind, shape, 31, L)
    st = []
    for name in FIELDS[kind]:
        for a in kf[name]:
            b = plan.alloc_k()
            plan.upload_k(b, a)
            st.append(b)
    bufs = [[plan.alloc_k() for _ in st] for _ in range(3)]
    d2 = bufs[2] if keep_d2 else None
    fn = plan.step_noalloc if noalloc else plan.step
    fn(hp.RK2MID, params(hp, kind, nd), st, bufs[0], bufs[1], d2, nsteps, dt)
    out = [plan.download_k(b) for b in st] + [plan.download_k(b) for b in bufs[1]]
    if keep_d2:
        out += [plan.download_k(b) for b in bufs[2]]
    return out


@pytest.mark.parametrize("env", [None, "DDX_NO_FUSED_UPDATE"], ids=["fused", "unfused"])
def test_noalloc_matches_step_two_calls(env):
    hp = hip()
    cases = [('mhd', (32, 32, 32), 3, 2e-3, True),      # larger workspace first
             ('hydro', (32, 32), 2, 1e-2, False)]       # then a smaller grid
    for kind, shape, n, dt, keep in cases:
        print("case: noalloc", env or "fused", kind, shape, n, "steps", flush=True)
        want = run(hp, kind, shape, n, dt, False, keep, env)
        got = run(hp, kind, shape, n, dt, True, keep, env)
        assert len(got) == len(want)
        for a, b in zip(got, want):
            assert np.all(np.isfinite(a))
            assert np.array_equal(a, b)


def test_noalloc_refuses_oversized_grid():
    hp = hip()
    print("case: noalloc oversized 128^3 MHD", flush=True)
    shape = (128, 128, 128)
    plan = hp.Plan(shape, (2 * np.pi,) * 3)
    st = [plan.alloc_k() for _ in range(6)]
    bufs = [[plan.alloc_k() for _ in range(6)] for _ in range(2)]
    with pytest.raises(hp.DDXError) as e:
        plan.step_noalloc(hp.RK2MID, params(hp, 'mhd', 3), st, bufs[0], bufs[1], None, 1, 1e-3)
    assert e.value.code == -1 and "static scratch" in str(e.value)
