"""The benchmark workloads for their own step counts against the oracle's
trajectories (GPU).

The fixtures under tests/golden/ hold the numpy oracle's state after a
workload's full run (tests/golden/make_fixtures.py: every kept mode with
|k| <= 16 in 2D or 12 in 3D, the L2 norm of each whole component and the
shell spectrum).  Here the HIP step runs the same count from the same seeded
initial condition:

- FP64 (the plain method): relative L2 <= 1e-12 on the stored modes of each
  component, on its whole-array norm, and on the spectrum;
- FP32 (128^3 hydro, and the 768-member ensemble as it is benchmarked): the
  error of each field against the same FP64 trajectory is printed and held
  to the bound written in the test.
"""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

FP64_CASES = ("boussinesq2d_512", "hydro3d_128", "mhd3d_256",
              "ensemble_m0", "ensemble_m1", "ensemble_m256", "ensemble_m512")


def hip():
    from dedalus import _hip
    _hip.init()
    return _hip


def fixture(name):
    path = os.path.join(GOLDEN, name + ".npz")
    assert os.path.exists(path), "missing fixture %s (tests/golden/make_fixtures.py)" % path
    return np.load(path)


def workload(name):
    import make_fixtures
    return make_fixtures.workloads()[name]()


def _import_maker():
    import sys
    if GOLDEN not in sys.path:
        sys.path.insert(0, GOLDEN)


def params_of(hp, w):
    prm = dict(w["params"])
    kind = w["kind"]
    nd = len(w["shape"])
    if kind == "hydro":
        return hp.make_params(hp.HYDRO, nu=prm["nu"])
    if kind == "mhd":
        return hp.make_params(hp.MHD, nu=prm["nu"], eta=prm["eta"], rho0=prm.get("rho0", 1.0))
    bdir = {"x": 0, "y": 1, "z": 2}[prm.get("boussinesq_direction", "z" if nd == 3 else "y")]
    return hp.make_params(hp.BOUSSINESQ, nu=prm["nu"], kappa=prm["kappa"], g=prm["g"],
                          alpha_t=prm["alpha_t"], beta=prm["beta"], bdir=bdir)


def initial(w):
    names = {"hydro": ["u"], "boussinesq": ["u", "T"], "mhd": ["u", "B"]}[w["kind"]]
    return [a for nm in names for a in w["fields"][nm]]


def run_hip(hp, w, precision):
    plan = hp.Plan(tuple(w["shape"]), tuple(w["length"]), precision)
    st = []
    for a in initial(w):
        b = plan.alloc_k()
        plan.upload_k(b, a)
        st.append(b)
    bufs = [[plan.alloc_k() for _ in st] for _ in range(3)]
    plan.step(hp.RK2MID, params_of(hp, w), st, bufs[0], bufs[1], bufs[2], int(w["nsteps"]), w["dt"])
    return [plan.download_k(b) for b in st]


def spectrum(comps, w):
    _import_maker()
    import make_fixtures
    nd = len(w["shape"])
    return make_fixtures.shell_spectrum(comps[:nd], tuple(w["shape"]), tuple(w["length"]))


def compare(fx, got, label):
    """Worst relative errors over the components: stored modes, whole norms."""
    worst_modes = worst_norm = 0.0
    for c, g in enumerate(got):
        idx, val = fx["idx_%d" % c], fx["val_%d" % c]
        e = float(np.linalg.norm(g.ravel()[idx] - val) / np.linalg.norm(val))
        n = abs(float(np.linalg.norm(g)) - float(fx["norm_%d" % c])) / float(fx["norm_%d" % c])
        print("  %s component %d: modes rel L2 %.2e, norm rel %.2e" % (label, c, e, n), flush=True)
        worst_modes, worst_norm = max(worst_modes, e), max(worst_norm, n)
    return worst_modes, worst_norm


@pytest.mark.parametrize("name", FP64_CASES)
def test_golden_fp64_full_run(name):
    _import_maker()
    hp = hip()
    fx = fixture(name)
    w = workload(name)
    print("case: golden FP64", name, tuple(w["shape"]), w["nsteps"], "steps", flush=True)
    ic = initial(w)
    assert int(fx["ncomp"]) == len(ic)
    for c, a in enumerate(ic):
        # the seeded generator still makes the fixture's initial condition
        assert abs(np.linalg.norm(a) - float(fx["ic_%d" % c])) <= 1e-15 * float(fx["ic_%d" % c])
    got = run_hip(hp, w, hp.F64)
    wm, wn = compare(fx, got, name)
    assert wm <= 1e-12 and wn <= 1e-12, (name, wm, wn)
    s, sref = spectrum(got, w), fx["spectrum"]
    serr = float(np.linalg.norm(s[:sref.size] - sref) / np.linalg.norm(sref))
    print("  spectrum rel L2 %.2e" % serr, flush=True)
    assert serr <= 1e-12


def test_golden_fp32_hydro3d_128():
    """The 128^3 hydro workload in FP32 for its 100 steps against the FP64
    oracle trajectory: per-field error of the FP32 transforms and products
    (FP64 projection and ETD arithmetic)."""
    _import_maker()
    hp = hip()
    fx = fixture("hydro3d_128")
    w = workload("hydro3d_128")
    print("case: golden FP32 hydro3d_128, 100 steps", flush=True)
    got = run_hip(hp, w, hp.F32)
    wm, wn = compare(fx, got, "hydro3d_128 FP32")
    # measured 2-3e-7 per field (FP32 rounding of the transforms)
    assert wm <= 2e-6 and wn <= 2e-6, (wm, wn)


def test_golden_fp32_ensemble768():
    """The 768-member FP32 ensemble as benchmarked (three batched groups
    captured together, many steps per graph launch: here 100, so the run is
    five replays) for its 500 steps; members 0, 1, 256 and 512 against their
    FP64 oracle trajectories."""
    _import_maker()
    from dedalus.ensemble import ENSEMBLE_768, ensemble_768
    import make_fixtures
    hip()
    ens = ensemble_768()
    print("case: golden FP32 ensemble_768, 500 steps", flush=True)
    ens.capture(steps_per_graph=100)    # takes one eager step: reseed after it
    ens.seed_members(ENSEMBLE_768["seed0"])
    ens.advance(ENSEMBLE_768["nsteps"])
    ens.sync()
    members = []
    for g in ens.groups:
        members.extend(g.read())
    worst = 0.0
    for m in make_fixtures.ENSEMBLE_MEMBERS:
        fx = fixture("ensemble_m%d" % m)
        wm, wn = compare(fx, list(members[m]), "member %d FP32" % m)
        worst = max(worst, wm, wn)
    # FP32 transforms over 500 steps of 2D turbulence
    assert worst <= 1e-4, worst
