#!/usr/bin/env python3
"""Throughput of the MI355X pseudospectral step (one JSON line on stdout).

A step is one RK2mid step (two right-hand sides, two ETD sweeps) of the
workload, with the state already resident in HBM.  Default workload: 3D
incompressible MHD, 256^3, FP64, seeded solenoidal u and B with
E(k) ~ k^4 exp(-2 (k/4)^2), rms 1 each, nu = eta = 1e-3, dt = 2.5e-4.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Workloads: mhd3d_256 (default), hydro3d_128, boussinesq2d_512,
taylor_green_64, ensemble_768 (768 FP32 2D members in one graph per step),
closure_conv (the closure network's convolutions alone, K passes).

Fields of the JSON line besides the standard ones:
  value         whole-job throughput of the timed region (K steps, state
                resident in HBM, one HIP graph per step replayed K times)
  roofline      dominant kernel class (most GPU time): algorithmic bytes
                per launch / mean launch duration, both measured with HIP
                events on the plan's stream over a second, instrumented
                timed region of the same K steps (eager launches, an event
                pair per launch); traffic = HBM bytes per launch from the
                rocprofv3 FETCH_SIZE / WRITE_SIZE summary in profiles/ when
                one exists for this kernel and workload, else null
  step_algorithmic  the whole step's algorithmic bytes (every kernel's
                input and output arrays once, summed over the instrumented
                pass, per step) x K / wall of the timed region, against
                8 TB/s: the step's roofline fraction on the bytes this
                dataflow must move
  model_speed   B_model per point-step (the byte model of the reference's
                minimal dataflow: 27 transforms on full layouts) x points x
                K / wall, against 8 TB/s: a speed ratio against that model,
                not a fraction of measured bytes (this dataflow moves fewer)
  hbm_measured  the metric's %HBM: HBM bytes per step measured by
                rocprofv3 FETCH_SIZE + WRITE_SIZE (the step summary in
                profiles/ for this workload and precision) x K / wall of
                the timed region, against 8 TB/s; null without a summary
  cpu_baseline  the numpy oracle (a restatement of the reference's numpy
                path) timed on this host: one warm-up step and the median
                of three timed steps of the same workload, one thread
  closure       (default workload) the metric's second part: the learned
                sub-grid closure on 512^2 -> 64^2 forced 2D turbulence --
                spectrum and enstrophy errors of coarse runs with and
                without it against the coarse-grained fine run (trained on
                the configuration's 2000 pairs, and on 4000 beside it),
                the speedup over the plain 512^2 FP64 run for single runs
                and for both batched at 64 members, and the MFMA rate of
                its convolutions
  source_hash   SHA-256 prefix of the kernel sources; rocprof summaries in
                profiles/ carry the hash of the sources they measured, and
                their figures are used only when it matches
"""

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "dedalus-1.0_amd"))

METRIC = "Mpt-steps/s at 256³ FP64 + %HBM peak; closure speedup at equal spectrum err"
HBM_PEAK_GBS = 8000.0

# bytes per point-step of the byte model of the reference's minimal dataflow
# (27 transforms on full layouts for 3D MHD; FP64), per workload
B_MODEL_PER_POINT_STEP = {
    "mhd3d_256": 1206.0,
    "hydro3d_128": 654.4,
    "boussinesq2d_512": 521.3,
    "taylor_green_64": 375.5,
    "ensemble_768": 370.4,   # 2D hydro, 8 transforms per RHS; halved for its FP32 state
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


TUNE_KNOBS = {"wide": "TUNE_WIDE_TILES", "xcd": "TUNE_XCD_ORDER", "yepi": "TUNE_FUSED_YEPI",
              "radix": "TUNE_RADIX"}


def parse_tune(text):
    """'radix=1,wide=0' -> {'radix': 1, 'wide': 0} (kernel choices, ddx_plan_set_tuning)."""
    out = {}
    for item in filter(None, (text or "").split(",")):
        k, _, v = item.partition("=")
        if k not in TUNE_KNOBS:
            raise SystemExit("--tune: unknown knob %r (one of %s)" % (k, ", ".join(TUNE_KNOBS)))
        out[k] = int(v)
    return out


def gpu_run(w, steps, warmup, precision="f64", tune=None):
    from dedalus import _hip

    nd = len(w["shape"])
    plan = _hip.Plan(w["shape"], w["length"], _hip.F64 if precision == "f64" else _hip.F32)
    for k, v in (tune or {}).items():
        plan.set_tuning(getattr(_hip, TUNE_KNOBS[k]), v)
    names = {"hydro": ["u"], "boussinesq": ["u", "T"], "mhd": ["u", "B"]}[w["kind"]]
    state = []
    for nm in names:
        for arr in w["fields"][nm]:
            b = plan.alloc_k()
            plan.upload_k(b, arr)
            state.append(b)
    stage = [plan.alloc_k() for _ in state]
    d1 = [plan.alloc_k() for _ in state]
    # ddx_step does not store the second right-hand side without deriv2
    # buffers (the reference's deriv2 is scratch to the step itself)
    d2 = None
    p = w["params"]
    phys = {"hydro": _hip.HYDRO, "boussinesq": _hip.BOUSSINESQ, "mhd": _hip.MHD}[w["kind"]]
    bdir = {"x": 0, "y": 1, "z": 2}[p.get("boussinesq_direction", "z" if nd == 3 else "y")]
    prm = _hip.make_params(phys, nu=p.get("nu", 0.), viscosity_order=p.get("viscosity_order", 1),
                           kappa=p.get("kappa", 0.), eta=p.get("eta", 0.), g=p.get("g", 1.),
                           alpha_t=p.get("alpha_t", 1.), beta=p.get("beta", 1.), bdir=bdir,
                           rho0=p.get("rho0", 1.))
    integ = _hip.RK2MID if w["integrator"] == "RK2mid" else _hip.RK2TRAP
    dt = w["dt"]
    log("warmup %d steps" % warmup)
    plan.step(integ, prm, state, stage, d1, d2, warmup, dt)
    plan.sync()
    log("timed %d steps" % steps)
    plan.sync()
    t0 = time.perf_counter()
    plan.step(integ, prm, state, stage, d1, d2, steps, dt)
    plan.sync()
    t1 = time.perf_counter()
    # the same K steps again with a HIP event pair around every launch
    # (per-kernel durations for the roofline; eager launches, no graph)
    log("instrumented %d steps" % steps)
    plan.prof_enable(True)
    plan.sync()
    t2 = time.perf_counter()
    plan.step(integ, prm, state, stage, d1, d2, steps, dt)
    plan.sync()
    t3 = time.perf_counter()
    prof = plan.prof_read()
    plan.prof_enable(False)
    prof["_wall_instrumented_s"] = t3 - t2
    # sanity: finite state after the run
    import numpy as np
    u0 = plan.download_k(state[0])
    if not np.all(np.isfinite(u0)):
        raise RuntimeError("non-finite state after the timed steps")
    return t1 - t0, prof


def ensemble_run(steps, warmup):
    """The 768-member FP32 ensemble: all three groups in one graph, the timed
    steps replayed as one launch of a graph of that many steps."""
    from dedalus.ensemble import ensemble_768, ENSEMBLE_768
    ens = ensemble_768()
    log("ensemble: %d members, %d points" % (ens.nmembers, ens.points))
    ens.capture(steps_per_graph=steps)
    ens.seed_members(ENSEMBLE_768["seed0"])
    ens.advance(warmup)
    ens.sync()
    log("timed %d steps" % steps)
    t0 = time.perf_counter()
    ens.advance(steps)
    ens.sync()
    t1 = time.perf_counter()
    log("instrumented %d steps" % steps)
    for g in ens.groups:
        g.plan.prof_enable(True)
    t2 = time.perf_counter()
    ens.step_eager(steps)
    ens.sync()
    t3 = time.perf_counter()
    prof = {}
    for g in ens.groups:
        for k, (n, ms, by) in g.plan.prof_read().items():
            a = prof.get(k, (0, 0.0, 0.0))
            prof[k] = (a[0] + n, a[1] + ms, a[2] + by)
        g.plan.prof_enable(False)
    import numpy as np
    e = ens.energies()
    if not np.all(np.isfinite(e)):
        raise RuntimeError("non-finite ensemble state after the timed steps")
    prof["_wall_instrumented_s"] = t3 - t2
    # each group alone (its own graph, the same K steps): the concurrent
    # ensemble step against the sum and the maximum of its parts
    alone = {}
    for g in ens.groups:
        g.plan.set_graph(True)
        g.step(ens.params, 2, ens.dt)
        g.plan.sync()
        ta = time.perf_counter()
        g.step(ens.params, steps, ens.dt)
        g.plan.sync()
        alone["%dx%d^2" % (g.n, g.shape[0])] = round((time.perf_counter() - ta) / steps * 1e3, 4)
    # the autotuner's choice per group, then the same timed region again
    ens.seed_members(ENSEMBLE_768["seed0"])
    chosen = ens.autotune(steps=4, reps=2)
    ens.capture(steps_per_graph=steps)
    ens.seed_members(ENSEMBLE_768["seed0"])
    ens.advance(warmup)
    ens.sync()
    t4 = time.perf_counter()
    ens.advance(steps)
    ens.sync()
    tuned = (time.perf_counter() - t4) / steps * 1e3
    prof["_ensemble"] = {"groups_alone_ms_per_step": alone,
                         "sum_of_groups_ms": round(sum(alone.values()), 4),
                         "max_group_ms": round(max(alone.values()), 4),
                         "builtin_ms_per_step": round((t1 - t0) / steps * 1e3, 4),
                         "tuned_ms_per_step": round(tuned, 4),
                         "tuned_choice": {"%dx%d" % k: v for k, v in chosen.items()}}
    return t1 - t0, prof, ens


def ensemble_cpu_baseline():
    """One RK2mid step of one member per grid size on the numpy oracle, scaled
    to the ensemble's point count (FP64 oracle; the GPU state is FP32)."""
    sys.path.insert(0, ROOT)
    import numpy as np
    from oracle import numpy_path as npth
    from dedalus.ensemble import ENSEMBLE_768, vorticity_member
    L = (2 * np.pi, 2 * np.pi)
    tot_t, tot_pts = 0.0, 0
    for shape, n in ENSEMBLE_768["groups"]:
        u0 = vorticity_member(shape, L, ENSEMBLE_768["seed0"])
        ph = npth.make_physics("hydro", shape, L, nu=ENSEMBLE_768["nu"])
        data = npth.load_state(ph, {"u": u0})
        ti = npth.RK2mid(ph)
        ti.do_advance(data, ENSEMBLE_768["dt"])      # warm-up step
        ts = []
        for _ in range(5):
            t0 = time.perf_counter()
            ti.do_advance(data, ENSEMBLE_768["dt"])
            ts.append(time.perf_counter() - t0)
        tot_t += float(np.median(ts)) * n
        tot_pts += shape[0] * shape[1] * n
    return {
        "value": tot_pts / tot_t / 1e6,
        "unit": "Mpt-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": "one member per grid size (64^2, 128^2, 256^2) on the numpy oracle (FP64, one "
                  "thread): 1 warm-up + 5 timed RK2mid steps, median, weighted by 256 members "
                  "each; %s" % host_description(),
    }


def closure_summary(batches=(1, 64), t_eval=20.0):
    """The learned closure on 512^2 -> 64^2 forced 2D turbulence: pairs from
    the fine solver, training on the configuration's 2000 pairs (and, beside
    it, on all 4000 generated), spectrum and enstrophy errors of coarse runs
    against the coarse-grained fine run over t in [5, 10] and [10, 20],
    throughput against the plain 512^2 run (single runs, and both batched at
    the same member count), and the MFMA rate of the network's forward
    convolutions."""
    import numpy as np
    import torch
    from dedalus.closure import ClosureNet, pipeline as P
    cfg = dict(P.CLOSURE_512_64)
    t0 = time.perf_counter()
    pipe = P.Pipeline(cfg)
    X, Y, info = pipe.generate_pairs()
    t_pairs = time.perf_counter() - t0
    log("closure: %d pairs in %.1f s" % (X.shape[0], t_pairs))
    nseeds = len(cfg["train_seeds"])
    sets = {"closure_%d" % cfg["train_pairs"]: P.train_subset(X, Y, nseeds, cfg["train_pairs"]),
            "closure_%d" % X.shape[0]: (X, Y)}
    main = "closure_%d" % cfg["train_pairs"]
    iters = cfg["train_iters"]
    nets, train = {}, {}
    for label, (XX, YY) in sets.items():
        net = ClosureNet(cin=3, cout=2, width=cfg["width"], depth=cfg["depth"], seed=0).cuda()
        t1 = time.perf_counter()
        scales, hist = P.train(net, XX, YY, iters=iters)
        train[label] = {"pairs": int(XX.shape[0]), "train_s": round(time.perf_counter() - t1, 1),
                        "final_loss": round(hist[-1][1], 4), "val_loss": round(hist[-1][2], 4)}
        log("closure: %s trained %d iterations in %.1f s" % (label, iters, train[label]["train_s"]))
        nets[label] = (net, scales)
    windows = [(t_eval / 4, t_eval / 2), (t_eval / 2, t_eval)]
    res = P.evaluate(pipe, nets=nets, T=t_eval, seeds=cfg["eval_seeds"], windows=windows)
    log("closure: spectrum error %s, enstrophy error %s" % (res["errs"], res["enstrophy_errs"]))
    ratio = cfg["dt_ratio"]
    thr = {}
    for b in batches:
        t_fine = P.time_fine_batched(pipe, b, 200 if b == 1 else 50)
        r = pipe.coarse_run(*nets[main], batch=None if b == 1 else b)
        r.capture()
        t = P.time_steps(pipe, r, 200 if b == 1 else 50)
        plain = pipe.coarse_run(None, batch=None if b == 1 else b)
        plain.capture()
        t_plain = P.time_steps(pipe, plain, 200 if b == 1 else 50)
        thr[str(b)] = {"coarse_closure_ms_per_step": round(t, 4),
                       "coarse_plain_ms_per_step": round(t_plain, 4),
                       "fine_ms_per_step": round(t_fine, 4),
                       "speedup": round(ratio * t_fine / t, 2)}
    r = lambda d: {k: round(v, 4) for k, v in d.items()}
    sp1, sp64 = thr["1"]["speedup"], thr[str(batches[-1])]["speedup"]
    # MFMA rate of the forward convolutions on 256 coarse images
    from dedalus.closure.net import Inference
    net = nets[main][0]
    B, (H, W) = 256, pipe.coarse_shape
    with torch.cuda.stream(pipe.stream):
        inf = Inference(net, B, H, W)
        inf.x.normal_()
        inf.run()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 20
        ev0.record()
        for _ in range(reps):
            inf.run()
        ev1.record()
        ev1.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    ch = net.chans
    alg = sum(2 * 9 * ch[l] * ch[l + 1] for l in range(len(ch) - 1)) * B * H * W
    padded = net.flops_per_pixel() * B * H * W
    return {
        "config": "512^2 -> 64^2 (8x8 box), nu 1e-4, dt 2.5e-4 / 2e-3, forcing |k| = 4 rms 0.1; "
                  "network 3-32-32-32-32-2, %d parameters" % sum(p.numel() for p in net.parameters()),
        "pairs_generated": int(X.shape[0]), "pairs_s": round(t_pairs, 1),
        "train_pairs": cfg["train_pairs"], "train_iters": iters, "train": train,
        "spectrum_err": r(res["errs"][0]),
        "enstrophy_err": r(res["enstrophy_errs"][0]),
        "spectrum_err_t10_20": r(res["errs"][1]),
        "enstrophy_err_t10_20": r(res["enstrophy_errs"][1]),
        "spectrum_err_per_seed": {k: [round(x, 4) for x in v] for k, v in res["member_errs"][0].items()},
        "spectrum_err_def": "mean over 1 <= k <= %d of |E_model/E_truth - 1|, Hermitian shell sums "
                            "averaged over %d held-out seeds and t in [%g, %g] (also [%g, %g]); "
                            "enstrophy <w^2>/2 (volume_average.py:191-200), |Z/Z_truth - 1| of the "
                            "same averages; training set for the headline: the configuration's "
                            "%d pairs" % (res["kc"], len(res["seeds"]), windows[0][0], windows[0][1],
                                         windows[1][0], windows[1][1], cfg["train_pairs"]),
        "throughput": thr,
        "speedup_def": "simulated time per wall second of the coarse closure run over the plain "
                       "512^2 FP64 run, dt ratio %d x t_fine / t_coarse; '1': one run each (the "
                       "metric's comparison), '%d': %d members of each batched on one plan "
                       "(like for like)" % (ratio, batches[-1], batches[-1]),
        "meets_20x": {"single_run": bool(sp1 >= 20.0), "batched_like_for_like": bool(sp64 >= 20.0)},
        "conv_roofline": {"bound": "mfma", "kernel": "conv3x3 forward (5 layers, 256 images 64^2)",
                          "achieved": round(alg / (ms * 1e-3) / 1e12, 2),
                          "peak": 2500.0, "unit": "TFLOP/s",
                          "frac": round(alg / (ms * 1e-3) / 1e12 / 2500.0, 5),
                          "mfma_padded_TFLOPs": round(padded / (ms * 1e-3) / 1e12, 2),
                          "ms_per_forward": round(ms, 4)},
    }


def copy_run(mib=1024, reps=20):
    """The HBM rate a plain device copy reaches (ddx_copy of `mib` MiB, one
    16-byte chunk per thread), timed with HIP events on the plan's stream:
    the practical ceiling the step's kernels are read against."""
    import torch
    from dedalus import _hip
    _hip.init()
    n = mib << 20
    src, dst = _hip.DevBuf(n), _hip.DevBuf(n)
    plan = _hip.Plan((8, 8), (1.0, 1.0))
    s = torch.cuda.Stream()
    plan.set_stream(s.cuda_stream)
    plan.zero(src)
    plan.copy(dst, src, n)
    plan.sync()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(reps):
        plan.copy(dst, src, n)
    e1.record(s)
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    gbs = 2.0 * n / (ms * 1e-3) / 1e9
    return {"metric": "device copy GB/s (read + write)", "value": round(gbs, 1), "unit": "GB/s",
            "n_gpus": 1, "steps": reps, "higher_is_better": True, "dtype": "u8",
            "data": "synthetic: zeroed buffer", "config": {"workload": "copy", "MiB": mib},
            "ms_per_copy": round(ms, 4),
            "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None}}


def closure_conv_run(batch=256, size=64, reps=20):
    """The closure network's convolutions alone (kernel profiles): the
    preallocated inference pass and one training forward + backward of the
    5-layer width-32 network on `batch` images of size^2."""
    import torch
    from dedalus import _hip
    from dedalus.closure.net import ClosureNet, Inference
    _hip.init()
    B, H = batch, size
    net = ClosureNet(cin=3, cout=2, width=32, depth=5, seed=0).cuda()
    inf = Inference(net, B, H, H)
    inf.x.normal_()
    x = torch.randn(B, H, H, 3, device="cuda")
    inf.run()
    net(x).square().mean().backward()
    torch.cuda.synchronize()
    e = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    e[0].record()
    for _ in range(reps):
        inf.run()
    e[1].record()
    for _ in range(reps):
        net.zero_grad(set_to_none=True)
        net(x).square().mean().backward()
    e[2].record()
    e[2].synchronize()
    fwd_ms = e[0].elapsed_time(e[1]) / reps
    train_ms = e[1].elapsed_time(e[2]) / reps
    ch = net.chans
    alg = sum(2 * 9 * ch[l] * ch[l + 1] for l in range(len(ch) - 1)) * B * H * H
    return {"metric": "closure convolution TFLOP/s (bf16 MFMA, fp32 accumulate)",
            "value": round(alg / fwd_ms / 1e9, 2), "unit": "TFLOP/s", "n_gpus": 1,
            "steps": reps, "higher_is_better": True, "dtype": "bf16",
            "data": "synthetic: random images and seeded weights",
            "config": {"workload": "closure_conv", "batch": B, "size": H,
                       "channels": list(ch)},
            "forward_ms": round(fwd_ms, 4), "train_step_ms": round(train_ms, 4),
            "train_TFLOPs": round(3 * alg / train_ms / 1e9, 2),
            "roofline": {"bound": "mfma", "achieved": round(alg / fwd_ms / 1e9, 2), "peak": 2500.0,
                         "unit": "TFLOP/s", "frac": round(alg / fwd_ms / 1e9 / 2500.0, 5),
                         "traffic": None}}


def step_traffic_from_profiles(workload, precision):
    """HBM bytes per step from the committed rocprofv3 step summary, if any,
    and only if it was measured on the kernels built from the current
    sources (its source_hash line)."""
    path = os.path.join(ROOT, "profiles", "step_%s_hbm_traffic.txt" % workload)
    if not os.path.exists(path):
        return None
    if profile_hash(path) != source_hash():
        log("warning: %s was measured on other kernel sources; hbm_measured is null" % path)
        return None
    key = "hbm_bytes_per_step_%s" % precision
    with open(path) as f:
        for line in f:
            if line.startswith(key):
                return float(line.split("=", 1)[1])
    return None


def traffic_from_profiles(kernel, workload):
    """HBM bytes per launch from a committed rocprofv3 summary, if present."""
    path = os.path.join(ROOT, "profiles", "%s_%s_pmc.txt" % (kernel, workload))
    if not os.path.exists(path):
        return None
    if profile_hash(path) != source_hash():
        log("warning: %s was measured on other kernel sources; traffic is null" % path)
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_launch"):
                return float(line.split("=", 1)[1])
    return None


def host_description():
    import numpy as np
    try:
        import scipy
        sv = scipy.__version__
    except ImportError:
        sv = "absent"
    return "host %s, %d logical CPUs, numpy %s, scipy %s" % (cpu_model(), os.cpu_count() or 0,
                                                             np.__version__, sv)


def cpu_baseline(w, timed=3):
    """The numpy oracle (the reference's numpy path restated) on this host,
    one process: one warm-up RK2mid step of the same workload, then `timed`
    steps; the median step gives Mpt-steps/s.  numpy.fft (pocketfft) and the
    ufuncs run on one thread."""
    sys.path.insert(0, ROOT)
    import numpy as np
    from oracle import numpy_path as npth

    ph = npth.make_physics(w["kind"], w["shape"], w["length"], **dict(w["params"]))
    data = npth.load_state(ph, w["fields"])
    ti = (npth.RK2mid if w["integrator"] == "RK2mid" else npth.RK2trap)(ph)
    log("cpu baseline: 1 warm-up + %d timed oracle steps" % timed)
    ti.do_advance(data, w["dt"])
    ts = []
    for i in range(timed):
        t0 = time.perf_counter()
        ti.do_advance(data, w["dt"])
        ts.append(time.perf_counter() - t0)
        log("cpu baseline: step %d %.2f s" % (i + 1, ts[-1]))
    med = float(np.median(ts))
    npts = float(np.prod(w["shape"]))
    return {
        "value": npts / med / 1e6,
        "unit": "Mpt-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": "RK2mid steps of the same %s workload (%s) on the numpy oracle, one process, "
                  "one thread (pocketfft): 1 warm-up step, then %d timed steps %s s, median "
                  "%.2f s/step; %s"
                  % (w["name"], "x".join(map(str, w["shape"])), timed,
                     "/".join("%.2f" % t for t in ts), med, host_description()),
    }


def source_hash():
    """First 16 hex digits of SHA-256 over the kernel sources and headers
    (dedalus-1.0_amd/csrc/*, include/*.h), in name order."""
    import hashlib
    h = hashlib.sha256()
    for d, pat in (("dedalus-1.0_amd/csrc", None), ("include", ".h")):
        base = os.path.join(ROOT, d)
        for name in sorted(os.listdir(base)):
            if pat and not name.endswith(pat):
                continue
            with open(os.path.join(base, name), "rb") as f:
                h.update(name.encode())
                h.update(f.read())
    return h.hexdigest()[:16]


def profile_hash(path):
    with open(path) as f:
        for line in f:
            if line.startswith("source_hash"):
                return line.split("=", 1)[1].strip()
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="mhd3d_256")
    ap.add_argument("--precision", choices=["f64", "f32"], default="f64",
                    help="f32: FP32 transforms and products (FP64 twiddles, projection, ETD)")
    ap.add_argument("--tune", default="",
                    help="kernel choices for the plan, e.g. radix=1,wide=1 (knobs: %s; "
                         "default: the built-in rule)" % ", ".join(TUNE_KNOBS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-closure", action="store_true",
                    help="skip the learned-closure measurement (the metric's second part)")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="after the timed region, copy this many MiB once with the library's "
                         "copy kernel (a known-size reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the pseudospectral step here is built and measured "
              "for a single MI355X (multi-GPU slab decomposition is design-only); "
              "got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.workload == "closure_conv":
        print(json.dumps(closure_conv_run(reps=args.steps)), flush=True)
        return
    if args.workload == "copy":
        print(json.dumps(copy_run(reps=args.steps)), flush=True)
        return
    if args.workload == "ensemble_768":
        from dedalus.ensemble import ENSEMBLE_768
        wall, prof, ens = ensemble_run(args.steps, args.warmup)
        npts = ens.points
        args.precision = "f32"
        w = {"name": "ensemble_768", "kind": "hydro",
             "shape": [list(g[0]) + [g[1]] for g in ENSEMBLE_768["groups"]],
             "integrator": "RK2mid", "dt": ENSEMBLE_768["dt"],
             "params": {"nu": ENSEMBLE_768["nu"]}}
    else:
        from dedalus.init_cond.seeded import workload
        w = workload(args.workload)
        npts = 1
        for n in w["shape"]:
            npts *= n
        wall, prof = gpu_run(w, args.steps, args.warmup, args.precision, parse_tune(args.tune))
    wall_instr = prof.pop("_wall_instrumented_s")
    ens_info = prof.pop("_ensemble", None)
    if args.calibrate_copy_mib:
        from dedalus import _hip
        n = args.calibrate_copy_mib << 20
        src, dst = _hip.DevBuf(n), _hip.DevBuf(n)
        cp = _hip.Plan((8, 8), (1.0, 1.0))
        cp.zero(src)
        cp.copy(dst, src, n)
        cp.sync()
        log("calibration copy: %d bytes read, %d bytes written" % (n, n))
    ms_step = wall / args.steps * 1e3
    value = npts * args.steps / wall / 1e6

    dom = max(prof, key=lambda k: prof[k][1])
    launches, ms_tot, by_tot = prof[dom]
    per_launch_bytes = by_tot / launches
    per_launch_s = ms_tot / launches / 1e3
    achieved = per_launch_bytes / per_launch_s / 1e9
    roof = {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(dom, w["name"]),
            "alg_bytes_per_launch": per_launch_bytes,
            "mean_launch_us": round(per_launch_s * 1e6, 2),
            "launches": launches,
            "gpu_time_share": round(ms_tot / sum(v[1] for v in prof.values()), 4)}
    kernels = {k: {"launches": v[0], "ms": round(v[1], 3),
                   "GBs": round(v[2] / (v[1] / 1e3) / 1e9, 1) if v[1] > 0 else None}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1][1])}
    balg = B_MODEL_PER_POINT_STEP.get(w["name"])
    if balg and args.precision == "f32":
        balg = balg / 2.0
    model = None
    if balg:
        gbs = balg * npts * args.steps / wall / 1e9
        model = {"bytes_per_point_step": balg, "achieved": round(gbs, 1),
                 "peak": HBM_PEAK_GBS, "unit": "GB/s", "ratio": round(gbs / HBM_PEAK_GBS, 4)}
    alg_step = sum(v[2] for v in prof.values()) / args.steps
    gbs = alg_step * args.steps / wall / 1e9
    step_alg = {"bytes_per_step": alg_step, "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}

    hbm_meas = None
    step_bytes = step_traffic_from_profiles(w["name"], args.precision)
    if step_bytes:
        gbs = step_bytes * args.steps / wall / 1e9
        hbm_meas = {"bytes_per_step": step_bytes, "achieved": round(gbs, 1),
                    "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpt-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "ms_per_step_instrumented": round(wall_instr / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": ("synthetic: seeded members (dedalus.ensemble.vorticity_member, seed 1000 + m)"
                 if w["name"] == "ensemble_768" else
                 "synthetic: seeded initial conditions (dedalus.init_cond.seeded.%s)" % w["name"]),
        "config": {"workload": w["name"], "physics": w["kind"],
                   "grid": list(w["shape"]), "integrator": w["integrator"], "dt": w["dt"],
                   "params": w["params"]},
        "roofline": roof,
        "step_algorithmic": step_alg,
        "model_speed": model,
        "hbm_measured": hbm_meas,
        "kernels": kernels,
        "source_hash": source_hash(),
    }
    if args.tune:
        out["config"]["tuning"] = parse_tune(args.tune)
    if ens_info:
        out["ensemble"] = ens_info
    if not args.no_closure and w["name"] == "mhd3d_256":
        try:
            out["closure"] = closure_summary()
        except Exception as e:  # reported, never silently dropped
            out["closure"] = {"error": "%s: %s" % (type(e).__name__, e)}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = (ensemble_cpu_baseline() if w["name"] == "ensemble_768"
                               else cpu_baseline(w))
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
