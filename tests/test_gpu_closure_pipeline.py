"""Closure pipeline on the GPU: coarse-graining and the sub-grid tendency
against the numpy oracle, graph-captured coarse runs against eager ones."""

import numpy as np
import pytest
import torch

from helpers import rel_l2
from oracle import numpy_path as npth

pytestmark = pytest.mark.gpu

L2 = (2 * np.pi, 2 * np.pi)


def case(*args):
    print("case:", *args, flush=True)


def box_ref(x, f):
    ny, nx = x.shape
    return x.reshape(ny // f, f, nx // f, f).mean(axis=(1, 3))


def filter_ref(kfine, fine_shape, f):
    cf = npth.Component(fine_shape, L2)
    cf['kspace'] = kfine
    xc = box_ref(cf['xspace'].copy(), f)
    cc = npth.Component(xc.shape, L2)
    cc['xspace'] = xc
    return cc['kspace'].copy()


@pytest.mark.parametrize("fine,f", [((128, 128), 4), ((512, 512), 8), ((96, 64), 2)], ids=str)
def test_pair_matches_oracle(fine, f):
    from dedalus.closure.sgs import CoarseGrainer, initial_velocity
    from dedalus import _hip
    _hip.init()
    case("sgs pair", fine, f)
    s = torch.cuda.Stream()
    cg = CoarseGrainer(fine, f, L2, stream=s.cuda_stream)
    u = initial_velocity(fine, L2, 5)
    F, C = cg.fine, cg.coarse
    st = []
    for a in u:
        b = F.alloc_k()
        F.upload_k(b, a)
        st.append(b)
    ubar = [C.alloc_k() for _ in range(2)]
    tau = [C.alloc_k() for _ in range(2)]
    cg.pair(st, ubar, tau)
    xb = [C.alloc_x() for _ in range(3)]
    cg.features_x(ubar, xb)
    got_u = [C.download_k(b) for b in ubar]
    got_t = [C.download_k(b) for b in tau]
    got_x = [C.download_x(b) for b in xb]
    coarse = cg.coarse_shape
    ub_ref = [filter_ref(a, fine, f) for a in u]
    nl_f = npth.rhs('hydro', fine, {'u': u}, L2)['u']
    nl_c = npth.rhs('hydro', coarse, {'u': ub_ref}, L2)['u']
    tau_ref = [filter_ref(a, fine, f) - b for a, b in zip(nl_f, nl_c)]
    for c in range(2):
        assert np.isfinite(got_u[c]).all() and np.isfinite(got_t[c]).all()
        assert rel_l2(got_u[c], ub_ref[c]) < 1e-13, c
        assert rel_l2(got_t[c], tau_ref[c]) < 1e-12, c
    # features: u, v, omega in x space
    cu = [npth.Component(coarse, L2) for _ in range(2)]
    for c in range(2):
        cu[c]['kspace'] = ub_ref[c]
    w = npth.Component(coarse, L2)
    w['kspace'] = cu[1].deriv('x') - cu[0].deriv('y')
    for got, ref in zip(got_x, [cu[0]['xspace'], cu[1]['xspace'], w['xspace']]):
        assert rel_l2(got, ref) < 1e-13


def _pipe_small():
    from dedalus.closure import pipeline as P
    return P, P.Pipeline(P.small_config())


def test_coarse_graph_replays_match_eager():
    """The captured coarse closure steps (features, network, forcing
    transform, RK2mid step), one step per graph and three steps per graph,
    replay bit-identically to eager steps, with host reads and other device
    work between replays."""
    import torch
    from dedalus.closure import ClosureNet
    from dedalus.closure.sgs import initial_velocity
    case("coarse closure graph replay")
    P, pipe = _pipe_small()
    C = pipe.cg.coarse
    net = ClosureNet(cin=3, cout=2, width=16, depth=3, seed=4).cuda()
    scales = (np.array([0.7, 0.7, 3.0]), np.array([0.1, 0.1]))
    u0 = initial_velocity(pipe.coarse_shape, L2, 9)
    ub = []
    for a in u0:
        b = C.alloc_k()
        C.upload_k(b, a)
        ub.append(b)
    results = []
    for graph in (0, 1, 3):
        for with_net in (False, True):
            r = pipe.coarse_run(net if with_net else None, scales if with_net else None)
            r.set_state(ub)
            if graph:
                r.capture(steps_per_graph=graph)
                r.set_state(ub)
            else:
                C.set_graph(False)
                r.one_step()          # the same warm-up step capture() takes
                r.set_state(ub)
            snaps = []
            for call in range(4):
                r.advance(3)
                snaps.append([C.download_k(b) for b in r.state])
                _ = torch.ones(1000, device="cuda").sum().item()   # other device work
            results.append(snaps)
    for a, b in ((0, 2), (1, 3), (0, 4), (1, 5)):
        for sa, sb in zip(results[a], results[b]):
            for x, y in zip(sa, sb):
                assert np.isfinite(x).all()
                assert np.array_equal(x, y)
    # the closure changes the run
    assert not np.array_equal(results[0][-1][0], results[1][-1][0])


def test_closure_reduces_spectrum_error_small():
    """End to end on 128^2 -> 32^2: pairs, training, and coarse runs with and
    without the closure against the coarse-grained fine run."""
    from dedalus.closure import ClosureNet
    case("closure end to end 128 -> 32")
    P, pipe = _pipe_small()
    X, Y, info = pipe.generate_pairs()
    assert X.shape == (32, 32, 32, 3) and Y.shape == (32, 32, 32, 2)
    assert bool(torch.isfinite(X).all()) and bool(torch.isfinite(Y).all())
    assert all(i["cfl"] < 0.5 for i in info)
    net = ClosureNet(cin=3, cout=2, width=16, depth=3, seed=0).cuda()
    scales, hist = P.train(net, X, Y, iters=1500, batch=8)
    print("training history (iteration, batch loss, validation loss):", hist)
    # the held-out loss falls (single-batch losses are noisy under the
    # random symmetry maps)
    assert hist[-1][2] < 0.7 * hist[0][2]
    res = P.evaluate(pipe, net, scales, T=4.0, seeds=(21, 22))
    print("spectrum error", res["err"])
    assert np.isfinite(res["err"]["closure"])
    assert res["err"]["closure"] < 0.5 * res["err"]["none"]
    # the batched evaluation is the mean of the members' own evaluations
    single = [P.evaluate(pipe, net, scales, T=4.0, seeds=(s,)) for s in (21, 22)]
    for name in ("truth", "none", "closure"):
        both = np.array(res["spectra"][name])
        avg = 0.5 * (np.array(single[0]["spectra"][name]) + np.array(single[1]["spectra"][name]))
        assert np.allclose(both, avg, rtol=1e-10, atol=1e-14 * avg.max()), name


def test_batched_closure_run_matches_single_runs():
    """A batched coarse closure run (network on B images, batched plan) gives
    every member bit-identically the single-member run's result."""
    from dedalus.closure import ClosureNet
    from dedalus.closure.sgs import initial_velocity
    case("batched closure run")
    P, pipe = _pipe_small()
    C = pipe.cg.coarse
    net = ClosureNet(cin=3, cout=2, width=16, depth=3, seed=6).cuda()
    scales = (np.array([0.7, 0.7, 3.0]), np.array([0.1, 0.1]))
    nb = 3
    u = [initial_velocity(pipe.coarse_shape, L2, 30 + m) for m in range(nb)]
    rb = pipe.coarse_run(net, scales, batch=nb)
    for c in range(2):
        rb.C.upload_k(rb.state[c], np.stack([u[m][c] for m in range(nb)]))
    rb.capture()
    for c in range(2):
        rb.C.upload_k(rb.state[c], np.stack([u[m][c] for m in range(nb)]))
    rb.advance(5)
    got = [rb.C.download_k(b) for b in rb.state]
    for m in range(nb):
        r1 = pipe.coarse_run(net, scales)
        for c in range(2):
            C.upload_k(r1.state[c], u[m][c])
        r1.capture()
        for c in range(2):
            C.upload_k(r1.state[c], u[m][c])
        r1.advance(5)
        for c in range(2):
            assert np.array_equal(got[c][m], C.download_k(r1.state[c])), (m, c)


def test_batched_pairs_match_single():
    """Sub-grid pairs of a batched coarse-grainer equal per-member single
    coarse-grainers bit for bit."""
    import torch
    from dedalus.closure.sgs import CoarseGrainer, initial_velocity
    from dedalus import _hip
    _hip.init()
    case("batched sgs pairs")
    fine, f, nb = (128, 128), 4, 3
    s = torch.cuda.Stream()
    cb = CoarseGrainer(fine, f, L2, stream=s.cuda_stream, batch=nb)
    u = [initial_velocity(fine, L2, 40 + m) for m in range(nb)]
    st = []
    for c in range(2):
        b = cb.fine.alloc_k()
        cb.fine.upload_k(b, np.stack([u[m][c] for m in range(nb)]))
        st.append(b)
    ubar = [cb.coarse.alloc_k() for _ in range(2)]
    tau = [cb.coarse.alloc_k() for _ in range(2)]
    cb.pair(st, ubar, tau)
    gu = [cb.coarse.download_k(b) for b in ubar]
    gt = [cb.coarse.download_k(b) for b in tau]
    c1 = CoarseGrainer(fine, f, L2, stream=s.cuda_stream)
    for m in range(nb):
        st1 = []
        for c in range(2):
            b = c1.fine.alloc_k()
            c1.fine.upload_k(b, u[m][c])
            st1.append(b)
        ub1 = [c1.coarse.alloc_k() for _ in range(2)]
        ta1 = [c1.coarse.alloc_k() for _ in range(2)]
        c1.pair(st1, ub1, ta1)
        for c in range(2):
            assert np.array_equal(gu[c][m], c1.coarse.download_k(ub1[c])), (m, c)
            assert np.array_equal(gt[c][m], c1.coarse.download_k(ta1[c])), (m, c)
