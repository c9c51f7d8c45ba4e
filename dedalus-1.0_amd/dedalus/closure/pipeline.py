"""Forced 2D turbulence -> (coarse state, sub-grid tendency) pairs -> trained
closure -> coarse runs compared with the coarse-grained fine run.

Configuration (CLOSURE_512_64):
  fine    512^2 IncompressibleHydro, FP64, nu = 1e-4, RK2mid, dt = 2.5e-4,
          state-independent forcing on 3.5 <= |k| < 4.5 with x-space rms 0.1
          and phases fixed by a seed (the VelocityForcing hook,
          dedalus/physics/physics.py:524-525, 569-572)
  coarse  64^2 (8 x 8 box filter), same nu, dt = 2e-3, forcing = B(f)
  pairs   8 seeds (one batched fine run), every 80 fine steps (0.02 time
          units) after a spin-up of 5 time units, 500 per seed: 4000 pairs;
          the configuration's training set is the 2000 of them taken every
          160 fine steps (train_subset), the same time span at half the
          sampling rate
  training 8000 Adam iterations, batch 32
  network (u, v, omega) -> tau (2 components), 3x3 periodic convolutions,
          widths 32, 5 layers, GELU, BF16 MFMA with FP32 accumulation

Time steps: RK2 amplifies an advected mode with y = |u| k dt by
|G|^2 = 1 + y^4 / 4 per step, which the integrating factor damps by
exp(-2 nu k^2 dt); growth at the 2/3 cutoff k_c stays below damping for
dt < (8 nu / (|u|^4 k_c^2))^(1/3): 5.7e-4 at 512^2 and 2.3e-3 at 64^2 for
|u| = 3.5.  dt = 1e-3 at 512^2 does blow up after about ten time units.

Every launch of a pipeline goes to one torch stream that the plans are bound
to, so the coarse closure step (features, network, forcing transform, RK2mid
step) can be captured as one graph.
"""

import ctypes
import time

import numpy as np
import torch

from .. import _hip
from .net import ClosureNet, Inference, pack, unpack
from .sgs import CoarseGrainer, band_forcing, features_x, initial_velocity, shell_spectrum

CLOSURE_512_64 = dict(
    fine_shape=(512, 512), factor=8, length=(2 * np.pi, 2 * np.pi), nu=1e-4,
    dt_fine=2.5e-4, dt_ratio=8, forcing=dict(kmin=3.5, kmax=4.5, rms=0.1), forcing_seed=7,
    spinup=5.0, sample_every=80, pairs_per_seed=500,
    train_seeds=(101, 102, 103, 104, 105, 106, 107, 108), train_iters=8000,
    eval_seed=201, eval_seeds=(201, 202, 203, 204), width=32, depth=5, train_pairs=2000,
)


def train_subset(X, Y, nseeds, npairs):
    """The first-in-time-order subsample of time-major pairs (pair j * nseeds +
    s is sample j of seed s) with npairs pairs: every (N / npairs)-th sample
    of every seed, so the subset spans the same times."""
    n = X.shape[0]
    if npairs >= n:
        return X, Y
    per = n // nseeds
    keep = per * nseeds // npairs
    if keep * npairs != n or per % keep:
        raise ValueError("%d pairs do not subsample evenly to %d" % (n, npairs))
    sl = lambda A: A.view(per, nseeds, *A.shape[1:])[::keep].reshape(npairs, *A.shape[1:])
    return sl(X), sl(Y)


def small_config(**over):
    """The same recipe on a smaller grid (tests)."""
    cfg = dict(CLOSURE_512_64)
    cfg.update(fine_shape=(128, 128), factor=4, dt_fine=1e-3, dt_ratio=4, spinup=0.5,
               sample_every=10, pairs_per_seed=16, train_seeds=(11, 12), eval_seed=21,
               width=16, depth=3)
    cfg.update(over)
    return cfg


class Pipeline:
    """Plans, buffers and the stream shared by pair generation and coarse runs."""

    def __init__(self, cfg, stream=None):
        _hip.init()
        self.cfg = cfg
        self.stream = stream or torch.cuda.Stream()
        self.cg = CoarseGrainer(cfg["fine_shape"], cfg["factor"], cfg["length"],
                                stream=self.stream.cuda_stream)
        F, C = self.cg.fine, self.cg.coarse
        self.coarse_shape = self.cg.coarse_shape
        self.prm = _hip.make_params(_hip.HYDRO, nu=cfg["nu"])
        L = cfg["length"]
        fc = cfg["forcing"]
        fk = band_forcing(cfg["fine_shape"], L, cfg["forcing_seed"], fc["kmin"], fc["kmax"], fc["rms"])
        self._fk_host = fk
        with torch.cuda.stream(self.stream):
            self.f_fine = [F.alloc_k() for _ in range(2)]
            for b, a in zip(self.f_fine, fk):
                F.upload_k(b, a)
            self.f_coarse = [C.alloc_k() for _ in range(2)]
            for c in range(2):
                self.cg.filter_k(self.f_fine[c], self.f_coarse[c])
            F.sync()

    # ---- fine runs -------------------------------------------------------
    def fine_state(self, seed):
        F = self.cg.fine
        u0 = initial_velocity(self.cfg["fine_shape"], self.cfg["length"], seed)
        st = [F.alloc_k() for _ in range(2)]
        with torch.cuda.stream(self.stream):
            for b, a in zip(st, u0):
                F.upload_k(b, a)
        bufs = [[F.alloc_k() for _ in range(2)] for _ in range(2)]
        return st, bufs

    def fine_advance(self, st, bufs, nsteps):
        self.cg.fine.step(_hip.RK2MID, self.prm, st, bufs[0], bufs[1], None, int(nsteps),
                          self.cfg["dt_fine"], forcing=self.f_fine)

    def max_speed(self, plan, st):
        """max |u| bound sqrt(max u^2 + max v^2) over x space (host sync)."""
        x = plan.alloc_x()
        m = 0.0
        for b in st:
            plan.bwd(b, x)
            m += plan.xreduce(x, 0)
        return float(np.sqrt(m))

    # ---- pairs -----------------------------------------------------------
    def generate_pairs(self, seeds=None, log=None, batched=True):
        """(X [N, m, m, 3], Y [N, m, m, 2]) float32 on the GPU, N = seeds x pairs_per_seed.

        batched: one fine run per seed, all seeds as members of one batched
        plan; pair i = j * len(seeds) + s is sample j of seed s (time-major, so
        the last pairs are the latest samples of every seed).  Otherwise the
        seeds run one after another and pair i = s * pairs_per_seed + j."""
        cfg = self.cfg
        seeds = cfg["train_seeds"] if seeds is None else seeds
        if batched and len(seeds) > 1:
            return self._generate_pairs_batched(list(seeds), log)
        C = self.cg.coarse
        m = self.coarse_shape
        npix = m[0] * m[1]
        n = len(seeds) * cfg["pairs_per_seed"]
        dev = torch.device("cuda", torch.cuda.current_device())
        X = torch.empty((n, m[0], m[1], 3), dtype=torch.float32, device=dev)
        Y = torch.empty((n, m[0], m[1], 2), dtype=torch.float32, device=dev)
        ubar = [C.alloc_k() for _ in range(2)]
        tau = [C.alloc_k() for _ in range(2)]
        xb = [C.alloc_x() for _ in range(3)]
        nspin = int(round(cfg["spinup"] / cfg["dt_fine"]))
        i = 0
        info = []
        with torch.cuda.stream(self.stream):
            for seed in seeds:
                st, bufs = self.fine_state(seed)
                self.fine_advance(st, bufs, nspin)
                umax = []
                for j in range(cfg["pairs_per_seed"]):
                    if j:
                        self.fine_advance(st, bufs, cfg["sample_every"])
                    self.cg.pair(st, ubar, tau)
                    self.cg.features_x(ubar, xb)
                    pack([(int(b), True) for b in xb], npix, 3, bf16=False,
                         f32=X[i].view(npix, 3))
                    C.bwd(tau[0], xb[0])
                    C.bwd(tau[1], xb[1])
                    pack([(int(xb[0]), True), (int(xb[1]), True)], npix, 2, bf16=False,
                         f32=Y[i].view(npix, 2))
                    i += 1
                    if j % max(1, cfg["pairs_per_seed"] // 4) == 0:
                        umax.append(self.max_speed(self.cg.fine, st))
                cfl = max(umax) * cfg["dt_fine"] / (cfg["length"][0] / cfg["fine_shape"][1])
                info.append(dict(seed=seed, max_speed=max(umax), cfl=cfl))
                if log:
                    log("pairs: seed %d done, max |u| %.3f, CFL %.3f" % (seed, max(umax), cfl))
            torch.cuda.current_stream().synchronize()
        return X, Y, info

    def _generate_pairs_batched(self, seeds, log=None):
        cfg = self.cfg
        nb = len(seeds)
        cg = CoarseGrainer(cfg["fine_shape"], cfg["factor"], cfg["length"],
                           stream=self.stream.cuda_stream, batch=nb)
        F, C = cg.fine, cg.coarse
        m = self.coarse_shape
        npix = nb * m[0] * m[1]
        P = cfg["pairs_per_seed"]
        dev = torch.device("cuda", torch.cuda.current_device())
        X = torch.empty((P * nb, m[0], m[1], 3), dtype=torch.float32, device=dev)
        Y = torch.empty((P * nb, m[0], m[1], 2), dtype=torch.float32, device=dev)
        nspin = int(round(cfg["spinup"] / cfg["dt_fine"]))
        with torch.cuda.stream(self.stream):
            f = [F.alloc_k() for _ in range(2)]
            st = [F.alloc_k() for _ in range(2)]
            u0 = [initial_velocity(cfg["fine_shape"], cfg["length"], sd) for sd in seeds]
            for c in range(2):
                F.upload_k(f[c], np.stack([self._fk_host[c]] * nb))
                F.upload_k(st[c], np.stack([u[c] for u in u0]))
            bufs = [[F.alloc_k() for _ in range(2)] for _ in range(2)]
            ubar = [C.alloc_k() for _ in range(2)]
            tau = [C.alloc_k() for _ in range(2)]
            xb = [C.alloc_x() for _ in range(3)]

            def advance(n):
                F.step(_hip.RK2MID, self.prm, st, bufs[0], bufs[1], None, int(n), cfg["dt_fine"],
                       forcing=f)

            advance(nspin)
            umax = []
            for j in range(P):
                if j:
                    advance(cfg["sample_every"])
                cg.pair(st, ubar, tau)
                cg.features_x(ubar, xb)
                pack([(int(b), True) for b in xb], npix, 3, bf16=False,
                     f32=X[j * nb:(j + 1) * nb].reshape(npix, 3))
                C.bwd(tau[0], xb[0])
                C.bwd(tau[1], xb[1])
                pack([(int(xb[0]), True), (int(xb[1]), True)], npix, 2, bf16=False,
                     f32=Y[j * nb:(j + 1) * nb].reshape(npix, 2))
                if j % max(1, P // 4) == 0:
                    umax.append(self.max_speed(F, st))
            torch.cuda.current_stream().synchronize()
        cfl = max(umax) * cfg["dt_fine"] / (cfg["length"][0] / cfg["fine_shape"][1])
        info = [dict(seeds=list(seeds), max_speed=max(umax), cfl=cfl)]
        if log:
            log("pairs: %d seeds batched, max |u| %.3f, CFL %.3f" % (nb, max(umax), cfl))
        return X, Y, info

    # ---- coarse runs -----------------------------------------------------
    def coarse_run(self, net=None, scales=None, batch=None, fused=None):
        return CoarseRun(self, net, scales, batch, fused)


def square_symmetry(A, g, pseudo=()):
    """Element g (0..7) of the square's symmetry group applied to a batch of
    x-space images (B, H, W, C) on an H = W periodic grid whose channels 0, 1
    are the x and y components of a vector, and channels in ``pseudo`` change
    sign under reflections (the vorticity).  Bit 0: transpose (x <-> y), bit
    1: reflect x (column j -> W - 1 - j), bit 2: reflect y.  Reflections about
    the grid's centre map the 8 x 8 blocks of the box filter onto each other,
    so the sub-grid target is exactly equivariant under them."""
    out = A
    sign = torch.ones(A.shape[-1], device=A.device, dtype=A.dtype)
    if g & 1:
        out = out.transpose(1, 2)
        perm = list(range(A.shape[-1]))
        perm[0], perm[1] = 1, 0
        out = out[..., perm]
        for c in pseudo:
            sign[c] = -sign[c]
    if g & 2:
        out = out.flip(2)
        sign[0] = -sign[0]
        for c in pseudo:
            sign[c] = -sign[c]
    if g & 4:
        out = out.flip(1)
        sign[1] = -sign[1]
        for c in pseudo:
            sign[c] = -sign[c]
    return (out * sign).contiguous()


def train(net, X, Y, iters=3000, batch=32, lr=2e-3, seed=0, val_frac=0.1, log=None,
          augment=True):
    """Adam on the normalised MSE; returns (scales, history).  The last
    val_frac of the pairs (the latest samples) is held out.  X channels are
    (u, v, omega), Y channels the sub-grid tendency (tau_u, tau_v).  With
    ``augment`` each batch is mapped by a random symmetry of the square
    (``square_symmetry``): the sub-grid tendency of 2D turbulence on a
    square periodic box, box-filtered on aligned blocks, is equivariant
    under them, so the 2000 pairs act as up to 8 x 2000.  The velocity
    components then share one normalisation scale."""
    n = X.shape[0]
    nval = max(1, int(round(n * val_frac)))
    ntr = n - nval
    xs = X[:ntr].std(dim=(0, 1, 2)).clamp_min(1e-12)
    ys = Y[:ntr].std(dim=(0, 1, 2)).clamp_min(1e-12)
    if augment:
        xs[:2] = xs[:2].square().mean().sqrt()
        ys[:2] = ys[:2].square().mean().sqrt()
    Xn, Yn = X / xs, Y / ys
    opt = torch.optim.Adam(net.parameters(), lr=lr)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=iters, eta_min=lr * 0.02)
    gen = torch.Generator().manual_seed(seed)
    hist = []
    t0 = time.perf_counter()
    for it in range(iters):
        idx = torch.randint(0, ntr, (batch,), generator=gen).to(X.device)
        xb, yb = Xn[idx], Yn[idx]
        if augment:
            g = int(torch.randint(0, 8, (1,), generator=gen))
            xb, yb = square_symmetry(xb, g, pseudo=(2,)), square_symmetry(yb, g)
        opt.zero_grad(set_to_none=True)
        loss = ((net(xb) - yb) ** 2).mean()
        loss.backward()
        opt.step()
        sched.step()
        if it % 250 == 0 or it == iters - 1:
            with torch.no_grad():
                vl = 0.0
                for s in range(ntr, n, batch):
                    vl += ((net(Xn[s:s + batch]) - Yn[s:s + batch]) ** 2).sum().item()
                vl /= nval * Yn[0].numel()
            hist.append((it, loss.item(), vl))
            if log:
                log("train it %d loss %.4f val %.4f (%.1f s)" % (it, loss.item(), vl,
                                                                   time.perf_counter() - t0))
    scales = (xs.detach().cpu().numpy().astype(np.float64), ys.detach().cpu().numpy().astype(np.float64))
    return scales, hist


class CoarseRun:
    """Coarse 2D hydro run, optionally with the learned closure: each step
    evaluates the network on the step's initial state and adds its tendency
    (with the coarse-grained forcing) as a constant forcing over the RK2mid
    step.  With ``batch=B`` the run advances B members at once on a batched
    plan (one set of launches for all members, the network on B images).
    ``capture()`` records one step as a graph on the pipeline's stream."""

    def __init__(self, pipe, net=None, scales=None, batch=None, fused=None):
        self.pipe = pipe
        self.batch = int(batch) if batch else None
        if self.batch:
            C = _hip.Plan(pipe.coarse_shape, pipe.cfg["length"], batch=self.batch)
            C.set_stream(pipe.stream.cuda_stream)
            fb = [pipe.cg.coarse.download_k(b) for b in pipe.f_coarse]
            self.forcing = []
            with torch.cuda.stream(pipe.stream):
                for a in fb:
                    buf = C.alloc_k()
                    C.upload_k(buf, np.stack([a] * self.batch))
                    self.forcing.append(buf)
        else:
            C = pipe.cg.coarse
            self.forcing = pipe.f_coarse
        self.C = C
        self.dt = pipe.cfg["dt_fine"] * pipe.cfg["dt_ratio"]
        self.state = [C.alloc_k() for _ in range(2)]
        self.bufs = [[C.alloc_k() for _ in range(2)] for _ in range(2)]
        self.net = net
        self.graph = None
        self.graphs = []
        m = pipe.coarse_shape
        nb = self.batch or 1
        self.npix = nb * m[0] * m[1]
        if net is not None:
            xs, ys = scales
            self.in_scale = [1.0 / float(v) for v in xs]
            self.out_scale = [float(v) for v in ys]
            with torch.cuda.stream(pipe.stream):
                self.inf = Inference(net, nb, m[0], m[1], fused=fused)
            self.xb = [C.alloc_x() for _ in range(3)]
            self.fx = [C.alloc_x() for _ in range(2)]
            # the coarse forcing in x space joins the network tendency at the
            # unpack, so one forward transform gives their sum
            self.fcx = [C.alloc_x() for _ in range(2)]
            with torch.cuda.stream(pipe.stream):
                C.bwd_many(self.forcing, self.fcx)

    def set_state(self, ubar):
        """Copy a coarse velocity (two k buffers of the run's plan, or of the
        pipeline's single coarse plan for a batched run: every member gets it)."""
        if self.batch and ubar and ubar[0].nbytes != self.state[0].nbytes:
            arrs = [self.pipe.cg.coarse.download_k(b) for b in ubar]
            for c in range(2):
                self.C.upload_k(self.state[c], np.stack([arrs[c]] * self.batch))
            return
        for c in range(2):
            self.C.copy(self.state[c], ubar[c])

    def one_step(self):
        C = self.C
        forcing = self.forcing
        if self.net is not None:
            features_x(C, self.state, self.xb)
            if self.inf.fused:
                # features -> network -> tendency + coarse forcing, one launch
                self.inf.run_fields(self.xb, self.in_scale, self.fx, self.out_scale, add=self.fcx)
            else:
                pack([(int(b), True) for b in self.xb], self.npix, self.inf.x.shape[1],
                     scale=self.in_scale, bf16=self.inf.x, f32=False)
                out = self.inf.run()
                unpack(out, 2, [int(b) for b in self.fx], True, scale=self.out_scale,
                       add=[int(b) for b in self.fcx])
            # the x-space sum enters the step's products, whose forward
            # transforms carry it (no transform of its own); the features
            # (u, v, omega) are the first right-hand side's x-space inputs
            C.step(_hip.RK2MID, self.pipe.prm, self.state, self.bufs[0], self.bufs[1], None,
                   1, self.dt, xforcing=self.fx, xin=self.xb)
            return
        C.step(_hip.RK2MID, self.pipe.prm, self.state, self.bufs[0], self.bufs[1], None,
               1, self.dt, forcing=forcing)

    def capture(self, steps_per_graph=10):
        """Record the steps as graphs (after one eager warm-up step): one of
        ``steps_per_graph`` consecutive steps, replayed for the bulk of a run
        so that the per-replay launch cost is paid once per that many steps,
        and one of a single step for the remainder."""
        C = self.C
        C.set_graph(False)   # the step's own graph would nest inside this capture
        k = max(1, int(steps_per_graph))
        graphs = []
        with torch.cuda.stream(self.pipe.stream):
            self.one_step()
            self.pipe.stream.synchronize()
            for n in sorted({1, k}):
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=self.pipe.stream):
                    for _ in range(n):
                        self.one_step()
                graphs.append((n, g))
        self.pipe.stream.synchronize()
        self.graphs = sorted(graphs, key=lambda t: -t[0])
        self.graph = self.graphs[-1][1]
        return self.graph

    def advance(self, nsteps):
        n = int(nsteps)
        with torch.cuda.stream(self.pipe.stream):
            if self.graph is None:
                for _ in range(n):
                    self.one_step()
                return
            for k, g in self.graphs:
                while n >= k:
                    g.replay()
                    n -= k


def spectrum_of(plan, bufs, shape, length):
    """Hermitian shell spectrum of device k buffers (host sync)."""
    return shell_spectrum([plan.download_k(b) for b in bufs], shape, length)[1]


def _window_mean(series, times, t0, t1):
    sel = (times >= t0 - 1e-12) & (times <= t1 + 1e-12)
    return np.mean(np.asarray(series)[sel], axis=0)


def evaluate(pipe, net=None, scales=None, T=10.0, snap_every=0.5, avg_from=None, seeds=None,
             log=None, nets=None, windows=None):
    """Run the fine truth and coarse models (with and without closure) from the
    same coarse-grained states; return spectra, enstrophies and errors.

    seeds: evaluation seeds (default: the configuration's eval_seed), run
    together as members of batched fine and coarse plans.  nets: {name:
    (net, scales)} of closures to evaluate side by side (default: net and
    scales as "closure").  Spectra are Hermitian-weighted shell sums (k =
    0..kmax of the coarse grid); the enstrophy is <omega^2>/2 as
    volume_average.py:191-200 computes it (Hermitian-weighted sum of
    |i kx v - i ky u|^2 / 2 over the coarse modes).  For each averaging
    window (t0, t1) (default one window [avg_from, T], avg_from = T/2):
      err[name]            mean over 1 <= k <= kc (the coarse 2/3 cutoff) of
                           |E_model(k) / E_truth(k) - 1|, spectra averaged
                           over the members and the snapshots in the window;
      enstrophy_err[name]  |Z_model / Z_truth - 1| of the window means;
      member_err[name]     the spectrum error of each member alone."""
    cfg = pipe.cfg
    seeds = list(seeds) if seeds else [cfg["eval_seed"]]
    nb = len(seeds)
    avg_from = T / 2 if avg_from is None else avg_from
    windows = list(windows) if windows else [(avg_from, T)]
    L = cfg["length"]
    cs = pipe.coarse_shape
    ratio = cfg["dt_ratio"]
    dtc = cfg["dt_fine"] * ratio
    nsnap = int(round(T / snap_every))
    csteps = int(round(snap_every / dtc))
    nspin = int(round(cfg["spinup"] / cfg["dt_fine"]))
    cg = CoarseGrainer(cfg["fine_shape"], cfg["factor"], L, stream=pipe.stream.cuda_stream, batch=nb)
    F, C = cg.fine, cg.coarse
    if nets is None:
        nets = {"closure": (net, scales)} if net is not None else {}
    runs = {"none": pipe.coarse_run(None, batch=nb)}
    for name, (nt, sc) in nets.items():
        runs[name] = pipe.coarse_run(nt, sc, batch=nb)

    def member_spectra(plan, bufs):
        arrs = [plan.download_k(b) for b in bufs]
        return np.array([shell_spectrum([a[m] for a in arrs], cs, L)[1] for m in range(nb)])

    def member_enstrophy(plan, bufs):
        out = plan.kreduce_members(bufs[0], 1, bufs[1])
        z = np.empty(nb)
        plan.d2h(z, out)
        return 0.5 * z

    with torch.cuda.stream(pipe.stream):
        f = [F.alloc_k() for _ in range(2)]
        st = [F.alloc_k() for _ in range(2)]
        u0 = [initial_velocity(cfg["fine_shape"], L, sd) for sd in seeds]
        for c in range(2):
            F.upload_k(f[c], np.stack([pipe._fk_host[c]] * nb))
            F.upload_k(st[c], np.stack([u[c] for u in u0]))
        bufs = [[F.alloc_k() for _ in range(2)] for _ in range(2)]

        def advance(n):
            F.step(_hip.RK2MID, pipe.prm, st, bufs[0], bufs[1], None, int(n), cfg["dt_fine"],
                   forcing=f)

        ubar = [C.alloc_k() for _ in range(2)]
        advance(nspin)
        for c in range(2):
            cg.filter_k(st[c], ubar[c])
        for r in runs.values():
            r.set_state(ubar)
            r.capture()
            r.set_state(ubar)
        spec = {k: [] for k in ["truth"] + list(runs)}     # [snap][member][k]
        zeta = {k: [] for k in ["truth"] + list(runs)}     # [snap][member]
        times = []
        for sn in range(nsnap):
            advance(csteps * ratio)
            for c in range(2):
                cg.filter_k(st[c], ubar[c])
            for r in runs.values():
                r.advance(csteps)
            times.append((sn + 1) * csteps * dtc)
            spec["truth"].append(member_spectra(C, ubar))
            zeta["truth"].append(member_enstrophy(C, ubar))
            for name, r in runs.items():
                spec[name].append(member_spectra(r.C, r.state))
                zeta[name].append(member_enstrophy(r.C, r.state))
        C.sync()
    times = np.array(times)
    kc = C.cut[0]
    ks = np.arange(1, kc + 1)
    res = dict(times=times.tolist(), kc=int(kc), seeds=seeds, windows=windows,
               errs=[], enstrophy_errs=[], member_errs=[])
    for (t0, t1) in windows:
        mean = {k: _window_mean(v, times, t0, t1).mean(axis=0) for k, v in spec.items()}
        memb = {k: _window_mean(v, times, t0, t1) for k, v in spec.items()}
        zm = {k: float(_window_mean(v, times, t0, t1).mean()) for k, v in zeta.items()}
        res["errs"].append({k: float(np.mean(np.abs(mean[k][ks] / mean["truth"][ks] - 1.0)))
                            for k in runs})
        res["enstrophy_errs"].append({k: abs(zm[k] / zm["truth"] - 1.0) for k in runs})
        res["member_errs"].append({k: [float(np.mean(np.abs(memb[k][m][ks] / memb["truth"][m][ks] - 1.0)))
                                       for m in range(nb)] for k in runs})
        if (t0, t1) == windows[0]:
            res["spectra"] = {k: v.tolist() for k, v in mean.items()}
    res["err"] = res["errs"][0]
    res["enstrophy_err"] = res["enstrophy_errs"][0]
    res["enstrophy"] = {k: np.asarray(v).mean(axis=1).tolist() for k, v in zeta.items()}
    res["energy"] = {k: [float(np.asarray(e).mean(axis=0).sum()) for e in v] for k, v in spec.items()}
    if log:
        log("evaluate: spectrum error " + ", ".join("%s %.3f" % kv for kv in res["err"].items())
            + "; enstrophy error " + ", ".join("%s %.3f" % kv for kv in res["enstrophy_err"].items()))
    return res


def time_steps(pipe, run, nsteps=200):
    """Mean wall time per replayed coarse step (ms) on the pipeline's stream."""
    with torch.cuda.stream(pipe.stream):
        run.advance(5)
        pipe.stream.synchronize()
        t0 = time.perf_counter()
        run.advance(nsteps)
        pipe.stream.synchronize()
        return (time.perf_counter() - t0) / nsteps * 1e3


def time_fine_batched(pipe, batch, nsteps=50, seed=None):
    """Mean wall time per step (ms) of `batch` plain fine runs advanced
    together on one batched plan (every member the same forced run; one
    graph per step, replayed)."""
    cfg = pipe.cfg
    seed = cfg["eval_seed"] if seed is None else seed
    if batch == 1:
        return time_fine_steps(pipe, nsteps, seed)
    F = _hip.Plan(cfg["fine_shape"], cfg["length"], batch=batch)
    F.set_stream(pipe.stream.cuda_stream)
    u0 = initial_velocity(cfg["fine_shape"], cfg["length"], seed)
    with torch.cuda.stream(pipe.stream):
        f = [F.alloc_k() for _ in range(2)]
        st = [F.alloc_k() for _ in range(2)]
        for c in range(2):
            F.upload_k(f[c], np.stack([pipe._fk_host[c]] * batch))
            F.upload_k(st[c], np.stack([u0[c]] * batch))
        bufs = [[F.alloc_k() for _ in range(2)] for _ in range(2)]

        def adv(n):
            F.step(_hip.RK2MID, pipe.prm, st, bufs[0], bufs[1], None, int(n), cfg["dt_fine"],
                   forcing=f)
        adv(5)
        pipe.stream.synchronize()
        t0 = time.perf_counter()
        adv(nsteps)
        pipe.stream.synchronize()
        return (time.perf_counter() - t0) / nsteps * 1e3


def time_fine_steps(pipe, nsteps=200, seed=None):
    """Mean wall time per fine step (ms): one graph per step, replayed."""
    seed = pipe.cfg["eval_seed"] if seed is None else seed
    with torch.cuda.stream(pipe.stream):
        st, bufs = pipe.fine_state(seed)
        pipe.fine_advance(st, bufs, 5)
        pipe.stream.synchronize()
        t0 = time.perf_counter()
        pipe.fine_advance(st, bufs, nsteps)
        pipe.stream.synchronize()
        return (time.perf_counter() - t0) / nsteps * 1e3
