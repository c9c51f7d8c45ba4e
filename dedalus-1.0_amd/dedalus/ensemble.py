"""Ensembles of independent 2D runs on batched plans.

An ensemble is a list of member groups, one per grid size; each group is a
batched plan (``ddx_plan_create_batched``) whose every call advances all of
its members, and ``capture()`` records one RK2mid step of every group as a
single graph, so a step of the whole ensemble (several grid sizes, hundreds
of members) is one graph launch.

Each member is exactly a dedalus IncompressibleHydro 2D run
(dedalus/physics/physics.py:527-599 with RK2mid, time_step.py:275-309): the
batched plan applies the same kernels per member as a plain plan does.

Workload ENSEMBLE_768: 256 members each of 64^2, 128^2 and 256^2, FP32 state
(FP64 twiddles, projection and ETD arithmetic), nu = 5e-4, dt = 1e-3, 500
steps; member m (0..767, in group order) starts from seed 1000 + m: a
random-phase vorticity with shell spectrum ~ k^-3 on 2 <= |k| <= 10,
psi = omega / k^2, u = d_y psi, v = -d_x psi (init_cond.py:405-408), scaled
to rms |u| = 1.  dt: RK2's growth 1 + (|u| k dt)^4 / 8 per step at the 2/3
cutoff k_c = 85 of 256^2 stays below the viscous damping nu k_c^2 dt for
dt < (8 nu / (|u|^4 k_c^2))^(1/3) = 1.9e-3 at |u| = 3.
"""

import numpy as np

from . import _hip
from .data_objects.kgrid import dealias_mask, wavenumbers
from .init_cond.seeded import hermitian_kx0, mean_square

ENSEMBLE_768 = dict(groups=(((64, 64), 256), ((128, 128), 256), ((256, 256), 256)),
                    nu=5e-4, dt=1e-3, nsteps=500, seed0=1000, precision="f32")


def vorticity_member(shape, length, seed, kmin=2.0, kmax=10.0, slope=-3.0, rms=1.0):
    """(u, v) reference-layout k arrays of one member (see the module docstring)."""
    k, _, ksh = wavenumbers(shape, length)
    kx = np.broadcast_to(k['x'], ksh)
    ky = np.broadcast_to(k['y'], ksh)
    k2 = kx ** 2 + ky ** 2
    kk = np.sqrt(k2)
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(ksh) + 1j * rng.standard_normal(ksh)
    band = (kk >= kmin) & (kk <= kmax)
    with np.errstate(divide='ignore', invalid='ignore'):
        amp = np.where(band, np.sqrt(np.where(band, kk, 1.0) ** slope / (2 * np.pi * np.where(band, kk, 1.0))), 0.0)
    w = w * amp
    w[dealias_mask(shape, length)] = 0.0
    hermitian_kx0(w, shape)
    psi = np.where(k2 > 0, w / np.where(k2 > 0, k2, 1.0), 0.0)
    u = 1j * ky * psi
    v = -1j * kx * psi
    ms = mean_square([u, v], shape)
    s = rms / np.sqrt(ms) if ms > 0 else 0.0
    return [u * s, v * s]


class Group:
    """Members on one grid: a batched plan and its state and work buffers."""

    def __init__(self, shape, nmembers, length, precision, stream=None):
        self.shape = tuple(shape)
        self.n = int(nmembers)
        self.length = tuple(length)
        self.plan = _hip.Plan(self.shape, self.length, precision, batch=self.n)
        if stream:
            self.plan.set_stream(stream)
        self.state = [self.plan.alloc_k() for _ in range(2)]
        self.bufs = [[self.plan.alloc_k() for _ in range(2)] for _ in range(2)]

    @property
    def points(self):
        return self.n * self.shape[0] * self.shape[1]

    def load(self, members):
        """members: list of (u, v) reference-layout arrays, one per member."""
        for c in range(2):
            self.plan.upload_k(self.state[c], np.stack([m[c] for m in members]))

    def read(self):
        """[(u, v)] per member (host)."""
        arrs = [self.plan.download_k(b) for b in self.state]
        return [(arrs[0][m], arrs[1][m]) for m in range(self.n)]

    def step(self, params, nsteps, dt):
        self.plan.step(_hip.RK2MID, params, self.state, self.bufs[0], self.bufs[1], None,
                       int(nsteps), dt)


class Ensemble:
    """Groups of batched members stepped together."""

    def __init__(self, groups, nu, dt, precision="f32", length=(2 * np.pi, 2 * np.pi), stream=None):
        import torch
        _hip.init()
        self.stream = stream or torch.cuda.Stream()
        prec = _hip.F32 if precision == "f32" else _hip.F64
        self.groups = [Group(shape, n, length, prec, self.stream.cuda_stream) for shape, n in groups]
        self.params = _hip.make_params(_hip.HYDRO, nu=nu)
        self.dt = float(dt)
        self.length = tuple(length)
        self.graph = None
        self.graphs = []

    @property
    def points(self):
        return sum(g.points for g in self.groups)

    @property
    def nmembers(self):
        return sum(g.n for g in self.groups)

    def seed_members(self, seed0):
        """Member m (in group order) from seed seed0 + m."""
        m = 0
        for g in self.groups:
            g.load([vorticity_member(g.shape, self.length, seed0 + m + j) for j in range(g.n)])
            m += g.n

    def step_eager(self, nsteps=1):
        for _ in range(int(nsteps)):
            for g in self.groups:
                g.step(self.params, 1, self.dt)

    def capture(self, concurrent=True, steps_per_graph=1):
        """Graphs holding RK2mid steps of every group: one of one step, and
        (steps_per_graph > 1) one of that many consecutive steps, which
        advance() replays for the bulk of a run, so that a run of
        steps_per_graph steps is one graph launch.

        The capture first takes one eager step (it allocates the workspaces,
        which cannot be allocated inside a graph), so every member advances
        by one step; seed or load the members afterwards to start the
        replays from a given state.

        concurrent: the groups are independent, so each group other than the
        first is captured on a side stream forked from and joined back to the
        ensemble's stream (once per step); the graph then lets the small
        groups' launches, which do not fill the GPU, run beside the large
        group's."""
        import torch
        for g in self.groups:
            g.plan.set_graph(False)
        side = [torch.cuda.Stream() for _ in self.groups[1:]] if concurrent else []
        k = max(1, int(steps_per_graph))

        def one_step():
            if not side:
                self.step_eager(1)
                return
            fork = torch.cuda.Event()
            fork.record(self.stream)
            joins = []
            for g, s in zip(self.groups[1:], side):
                s.wait_event(fork)
                g.plan.set_stream(s.cuda_stream)
                g.step(self.params, 1, self.dt)
                e = torch.cuda.Event()
                e.record(s)
                joins.append(e)
            self.groups[0].step(self.params, 1, self.dt)
            for e in joins:
                self.stream.wait_event(e)

        graphs = []
        with torch.cuda.stream(self.stream):
            self.step_eager(1)                       # workspaces, twiddles
            self.stream.synchronize()
            for n in sorted({1, k}):
                graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(graph, stream=self.stream):
                    for _ in range(n):
                        one_step()
                graphs.append((n, graph))
        for g in self.groups[1:]:
            g.plan.set_stream(self.stream.cuda_stream)
        self.stream.synchronize()
        self.graphs = sorted(graphs, key=lambda t: -t[0])
        self.graph = self.graphs[-1][1]
        self._side = side                            # kept alive with the graphs
        return self.graph

    def autotune(self, steps=4, reps=2):
        """Choose, per group, the kernel variants that run its step fastest:
        the radix plan of the row and column transforms, 16-column tiles
        (256-point groups), the XCD order of the spectral y pass, and the
        fused forward y pass and epilogue (64-point groups, built-in radix).
        The layout choices give bit-identical results in FP32 (FP64 tile
        widths agree to rounding); the radix plan agrees to rounding.  Each
        candidate is timed with HIP events over `steps` eager steps, best of
        `reps`, and the members' state is restored afterwards.  Returns
        {shape: choice}; capture() afterwards to replay the tuned kernels."""
        import itertools
        import torch
        self.graph = None
        self.graphs = []
        self.graphs = []
        chosen = {}
        with torch.cuda.stream(self.stream):
            for g in self.groups:
                g.plan.set_graph(False)
                saved = [g.plan.alloc_k(zero=False) for _ in g.state]
                for b, s in zip(saved, g.state):
                    g.plan.copy(b, s)
                wide = (0, 1) if g.shape[0] == 256 else (-1,)
                yepi = (0, 1) if g.shape[0] == 64 else (-1,)
                knobs = (_hip.TUNE_WIDE_TILES, _hip.TUNE_XCD_ORDER, _hip.TUNE_FUSED_YEPI, _hip.TUNE_RADIX)
                best = None
                for cand in itertools.product(wide, (0, 1, 2), yepi, (0, 1)):
                    if cand[3] == 1 and cand[2] == 1:
                        continue                      # radix 1 runs the y pass and epilogue apart
                    for knob, v in zip(knobs, cand):
                        g.plan.set_tuning(knob, v)
                    g.step(self.params, 1, self.dt)          # first use of the variant
                    t = None
                    for _ in range(reps):
                        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                        e0.record(self.stream)
                        g.step(self.params, steps, self.dt)
                        e1.record(self.stream)
                        e1.synchronize()
                        ms = e0.elapsed_time(e1) / steps
                        t = ms if t is None else min(t, ms)
                    if best is None or t < best[0]:
                        best = (t, cand)
                for knob, v in zip(knobs, best[1]):
                    g.plan.set_tuning(knob, v)
                for b, s in zip(saved, g.state):
                    g.plan.copy(s, b)
                self.stream.synchronize()
                chosen[g.shape] = dict(wide_tiles=best[1][0], xcd_order=best[1][1], fused_yepi=best[1][2],
                                       radix=best[1][3], ms_per_step=best[0])
        return chosen

    def advance(self, nsteps):
        """nsteps RK2mid steps of every member: the captured graphs (the
        longest first), or eager launches before capture()."""
        import torch
        n = int(nsteps)
        with torch.cuda.stream(self.stream):
            if self.graph is None:
                self.step_eager(n)
                return
            for k, g in self.graphs:
                while n >= k:
                    g.replay()
                    n -= k

    def sync(self):
        self.stream.synchronize()

    def _member_reduce(self, op):
        out = []
        for g in self.groups:
            if op == 0:
                a = g.plan.kreduce_members(g.state[0], 0)
                b = g.plan.kreduce_members(g.state[1], 0)
                va, vb = np.empty(g.n), np.empty(g.n)
                g.plan.d2h(va, a)
                g.plan.d2h(vb, b)
                out.append(va + vb)
            else:
                a = g.plan.kreduce_members(g.state[0], 1, g.state[1])
                va = np.empty(g.n)
                g.plan.d2h(va, a)
                out.append(va)
        return np.concatenate(out)

    def energies(self):
        """<|u|^2> / 2 per member, in member order (device reductions, one
        download per group)."""
        return 0.5 * self._member_reduce(0)

    def enstrophies(self):
        """<omega^2> / 2 per member (volume_average.py:191-200 for each member)."""
        return 0.5 * self._member_reduce(1)


def ensemble_768(stream=None):
    cfg = ENSEMBLE_768
    ens = Ensemble(cfg["groups"], cfg["nu"], cfg["dt"], cfg["precision"], stream=stream)
    ens.seed_members(cfg["seed0"])
    return ens
