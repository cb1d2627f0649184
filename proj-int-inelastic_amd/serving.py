"""Continuously batched solve server for many homogeneous and 1D-space problems.

New (the reference solves one problem per Python process).  Problems are
grouped into classes by their collision plan (dimension, grid, quadrature,
kernel exponent, precision); a class keeps the distribution functions of its
active problems resident in one device tensor.  Each scheduler tick advances
every active problem by one outer step of its class's integrator, with the
collision operator of all problems of the class (all spatial cells of all
1D jobs, or all homogeneous problems) evaluated in one batched launch, each
problem with its own restitution e.  Between ticks, finished problems retire
(their f is copied to the host, the slot and memory are freed) and arrived
requests are admitted while the HBM budget allows; a request that does not fit
waits in the queue (back-pressure) instead of failing.  The budget counts each
request's resident arrays (Request.nbytes) and each class's plan scratch, which
the plan keeps between calls: admitting cells into a class reserves what the
library reports the class's next call will hold (ib_plan_scratch_need).  Classes
with no active problem give their plan (and its scratch) back when a request
would not fit otherwise.

Step of one tick, for a class with integrator "TPI2" and parameters
(inner_steps, inner_dt, dt), is the reference's tel_proj_step2
(src/boltzmann.py:130-140) applied to the stacked states; "TPI3" is
tel_proj_step3 (:142-158), "PFE" proj_euler_step (:120-128), "Euler"
euler_step (:82-84).  df of each
problem is its TVD transport increment (1D jobs) plus dt Q(f)/kn.
"""
import itertools
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from . import device as dev
from .collision import FSInelasticVHSCollision, SpectralMesh
from .collision.inelastic_collisions import collide_group


@dataclass
class Request:
    """One solve: run f0 to t_final with the class's integrator.

    dim/n/nr/m/s/gamma/precision select the collision plan; e is this
    problem's restitution (e0 = e); kn the Knudsen number.  For a 1D job
    f0 has shape (nx, n, n) and bc/xlen give the outflow boundaries and the
    length of the x interval; homogeneous problems have f0 of shape n^dim.
    """
    rid: int
    f0: np.ndarray
    t_final: float
    e: float
    dim: int = 2
    n: int = 32
    nr: int = 32
    m: int = 8
    s: float = 3.624654714576
    gamma: float = 1.0
    kn: float = 1.0
    precision: str = "fp64"
    integrator: str = "TPI2"
    inner_steps: tuple = (2, 2)
    inner_dt: tuple = (0.01, 0.05)
    dt: float = 0.25
    xlen: float = 0.0  # > 0: 1D-space job on [0, xlen] with extrapolation boundaries
    arrival: float = 0.0

    def class_key(self):
        # gamma and e are per-cell parameters of one plan: problems with either
        # kernel share a class and its batched launches
        return (self.dim, self.n, self.nr, self.m, self.s) + self.group_key()

    def group_key(self):
        # classes with one integrator schedule step together (grouped launches)
        return (self.precision, self.integrator, tuple(self.inner_steps), tuple(self.inner_dt),
                self.dt)

    @property
    def cells(self):
        return self.f0.shape[0] if self.xlen > 0 else 1

    def nbytes(self):
        # resident f plus per-step temporaries (df, Q, transport) in FP64; the class's plan
        # scratch is reserved per class at admission (Server._admit)
        return 4 * self.f0.size * 8


@dataclass
class _Active:
    req: Request
    start: int      # first cell in the class tensor
    cells: int
    t: float = 0.0
    steps_left: int = 0
    admitted_at: float = 0.0


@dataclass
class Result:
    rid: int
    f: np.ndarray
    t: float
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self):
        return self.finished - self.arrival


_SERIAL = itertools.count()


class _Class:
    """Problems sharing one collision plan and one integrator."""

    def __init__(self, key, proto, autotune=False, scratch_budget=1 << 30):
        t = dev.require_gpu()
        # graphs are keyed by (serial, version): a class built after another was released
        # never matches the released class's graph, whatever id() Python gives it
        self.serial = next(_SERIAL)
        self.key = key
        self.proto = proto
        self.autotune = autotune
        d = proto.dim
        cfg = {"collision-model": {"dim": d, "e": 1.0, "e0": 1.0, "gamma": proto.gamma},
               "velocity-mesh": {"nv": proto.n, "s": proto.s, "nr": proto.nr}}
        if d == 2:
            cfg["circle-quad-rule"] = {"nphi": proto.m}
        else:
            cfg["spherical-design-rule"] = {"ssrule": "lebedev", "nsphr": proto.m}
        self.sm = SpectralMesh(cfg)
        self.coll = FSInelasticVHSCollision(cfg, self.sm, precision=proto.precision)
        # FP32: a class's whole batch fits its plan's scratch, so its pair loops join the
        # grouped launch of its group (admission counts the scratch against the HBM budget);
        # FP64 plans keep the default budget, by which the staged engine sizes its sub-batches
        if proto.precision != "fp64":
            self.coll.set_tuning(scratch_budget=scratch_budget)
        self.vshape = (proto.n,) * d
        self.active = []
        self.f = t.empty((0,) + self.vshape, dtype=t.float64, device="cuda")
        v = self.sm.v_center
        self.v1 = dev.to_device(np.repeat(v[:, None], proto.n, axis=1).reshape(-1)) if d == 2 else None

    def rebuild(self, keep, new):
        """Compact the resident tensor to the kept problems and append new ones."""
        t = dev.torch()
        parts, act, start = [], [], 0
        for a in keep:
            parts.append(self.f[a.start:a.start + a.cells])
            a.start = start
            start += a.cells
            act.append(a)
        # new problems of one kind next to each other, so that df_from_q runs them in one launch
        new = sorted(new, key=lambda af: (af[0].req.xlen > 0, af[0].cells, af[0].req.xlen, af[0].req.kn))
        if new:  # one host-to-device copy for all admitted problems
            parts.append(dev.to_device(np.concatenate(
                [f0.reshape((a.cells,) + self.vshape) for a, f0 in new])))
        for a, _ in new:
            a.start = start
            start += a.cells
            act.append(a)
        self.f = t.cat(parts).contiguous() if parts else self.f[:0]
        self.active = act
        self.version = getattr(self, "version", 0) + 1
        cell_e = np.concatenate([np.full(a.cells, a.req.e) for a in act]) if act else np.zeros(0)
        self.cell_e = cell_e
        self.cell_gamma = np.concatenate([np.full(a.cells, a.req.gamma) for a in act]) if act else np.zeros(0)
        self.cell_kn = dev.to_device(np.concatenate([np.full(a.cells, a.req.kn) for a in act])) if act else None
        if self.autotune and act:
            self._tune()

    def _tune(self):
        """Apply the launch and tile shape of this class's cell-count bucket: timed the first
        time the process meets the bucket, then taken from the autotuner's cache (a class
        whose cell count moves between buckets gets each bucket's own setting back)."""
        from . import autotune as at
        at.tune(self.coll, self.f, with_loss=True)

    def df_from_q(self, f, q, h):
        """Stacked df of every problem: TVD transport (1D jobs) + h Q / kn.

        Adjacent problems of one kind run as one launch per kernel: homogeneous
        problems with one kn, and 1D jobs with one (Nx, x length, kn), whose TVD
        steps are independent slices of one (jobs, Nx, N^2) launch.
        """
        t = dev.torch()
        lib = _lib.load()
        st = dev.stream_handle()
        out = t.empty_like(f)
        npc = f[0].numel() if f.shape[0] else 0

        def key(a):
            r = a.req
            return (r.cells, r.xlen, r.kn) if r.xlen > 0 else (0, 0.0, r.kn)

        runs = []  # [first cell, cells, jobs, key]
        for a in self.active:
            k = key(a)
            if runs and runs[-1][3] == k and runs[-1][0] + runs[-1][1] == a.start:
                runs[-1][1] += a.cells
                runs[-1][2] += 1
            else:
                runs.append([a.start, a.cells, 1, k])
        for c0, nc, jobs, (nx, xlen, kn) in runs:
            sl = slice(c0, c0 + nc)
            adv = None
            if xlen > 0:
                if nx + 2 < self.proto.n:
                    # the reference slices s[LL-1:UL-1] along v_x and fails to broadcast
                    raise ValueError("operands could not be broadcast together (nx + 2 < N)")
                adv = t.empty_like(f[sl])
                _lib.check(lib.ib_tvd_axis(f[sl].data_ptr(), adv.data_ptr(), jobs, nx, npc,
                                           self.v1.data_ptr(), npc, float(h / (xlen / nx)), 2, 0,
                                           1, 1, 0, st), "ib_tvd_axis")
            _lib.check(lib.ib_df_combine(dev.ptr(adv), q[sl].data_ptr(), out[sl].data_ptr(), float(h),
                                         float(kn), out[sl].numel(), st), "ib_df_combine")
        return out


def _axpy(y, x, a):
    _lib.check(_lib.load().ib_axpy(y.data_ptr(), x.data_ptr(), float(a), y.numel(),
                                   dev.stream_handle()), "ib_axpy")


def _axpy_chain(y, x, a, b=1.0, c=1.0):
    """y += a * (b * (c * x)): the telescopic increments df1 = M0 df0, df2 = M1 df1 without
    storing them."""
    _lib.check(_lib.load().ib_axpy_chain(y.data_ptr(), x.data_ptr(), float(a), float(b), float(c),
                                         y.numel(), dev.stream_handle()), "ib_axpy_chain")


INTEGRATORS = ("Euler", "PFE", "TPI2", "TPI3")


class _Group:
    """Classes (grid shapes) sharing one integrator schedule.

    Each df stage evaluates Q of every problem of every member class with one
    grouped launch per engine stage (ib_collide_group).  While the membership
    is unchanged, one whole outer step is replayed from a captured HIP graph.
    """

    def __init__(self, key, use_graph=True):
        self.key = key
        self.classes = []
        self.use_graph = use_graph
        self._graph = None
        self._graph_key = None
        # each group steps on its own stream (default priority), so that the groups of one
        # tick run side by side on the device
        self.stream = dev.torch().cuda.Stream()

    def df(self, h):
        cls = self.classes
        qs = collide_group([c.coll for c in cls], [c.f for c in cls], with_loss=True,
                           es=[c.cell_e for c in cls], gammas=[c.cell_gamma for c in cls])
        return [c.df_from_q(c.f, q, h) for c, q in zip(cls, qs)]

    def _outer_step(self):
        p = self.classes[0].proto
        fs = [c.f for c in self.classes]
        if p.integrator == "Euler":
            for f, d in zip(fs, self.df(p.dt)):
                _axpy(f, d, 1.0)
        elif p.integrator == "PFE":
            M = p.dt / p.inner_dt[0] - p.inner_steps[0]
            ds = None
            for _ in range(p.inner_steps[0]):
                ds = self.df(p.inner_dt[0])
                for f, d in zip(fs, ds):
                    _axpy(f, d, 1.0)
            for f, d in zip(fs, ds):
                _axpy(f, d, M)
        elif p.integrator == "TPI2":
            K, h = p.inner_steps, p.inner_dt
            M0, M1 = h[1] / h[0] - K[0], p.dt / h[1] - K[1]
            d0 = None
            for _ in range(K[1]):
                d0 = self._inner_block(fs, K[0], h[0])
                for f, d in zip(fs, d0):
                    _axpy_chain(f, d, M0)                  # df1 = M0 df0; f += df1
            for f, d in zip(fs, d0):
                _axpy_chain(f, d, M1, M0)                  # f += M1 df1
        elif p.integrator == "TPI3":
            K, h = p.inner_steps, p.inner_dt
            M0, M1, M2 = h[1] / h[0] - K[0], h[2] / h[1] - K[1], p.dt / h[2] - K[2]
            d0 = None
            for _ in range(K[2]):
                for _ in range(K[1]):
                    d0 = self._inner_block(fs, K[0], h[0])
                    for f, d in zip(fs, d0):
                        _axpy_chain(f, d, M0)              # df1 = M0 df0; f += df1
                for f, d in zip(fs, d0):
                    _axpy_chain(f, d, 1.0, M1, M0)         # df2 = M1 df1; f += df2
            for f, d in zip(fs, d0):
                _axpy_chain(f, d, M2, M1, M0)              # f += M2 df2
        else:
            raise NotImplementedError("This time integrator is not implemented.")

    def _inner_block(self, fs, k, h):
        """k Euler steps of h; returns the last increments."""
        d0 = None
        for _ in range(k):
            d0 = self.df(h)
            for f, d in zip(fs, d0):
                _axpy(f, d, 1.0)
        return d0

    def step(self):
        t = dev.torch()
        gkey = tuple((c.serial, c.version) for c in self.classes)
        if self.use_graph and self._graph is not None and self._graph_key == gkey:
            self._graph.replay()
            return
        self._outer_step()
        if not self.use_graph:
            return
        # membership changed: the eager step above sized the scratch and installed
        # the parameter sets; capture the next steps' work for replay
        t.cuda.synchronize()
        g = t.cuda.CUDAGraph()
        with t.cuda.graph(g):
            self._outer_step()
        self._graph, self._graph_key = g, gkey


class Server:
    """Admission by HBM budget, continuous batching, latency bookkeeping."""

    def __init__(self, hbm_budget=64 << 30, max_cells_per_class=None, clock=time.perf_counter,
                 use_graph=True, autotune=False):
        self.use_graph = use_graph
        self.autotune = autotune
        self.groups = {}
        self.budget = int(hbm_budget)
        self.used = 0
        self.max_cells = max_cells_per_class
        self.classes = {}
        self.queue = []
        self.results = {}
        self.clock = clock
        self.t0 = None
        self.reserved = {}   # class key -> plan scratch bytes counted in `used`
        self.ticks = 0
        self.peak_used = 0
        self.tick_s = 2e-3   # shortest tick while requests are still to arrive (seconds)
        self._step_s = 0.0   # running estimate of one outer step's duration

    def submit(self, req):
        """Queue a request.  Requests that no state of the server could admit raise
        here instead of waiting forever: a memory need above the whole budget, more
        cells than a class may hold, or a step that does not divide t_final (the
        reference's evolve_to_time check, src/solver.py:122-128), an integrator
        this server does not run, or a schedule with too few levels for it."""
        if req.integrator not in INTEGRATORS:
            raise ValueError("request %r: unknown time integrator %r (known: %s)"
                             % (req.rid, req.integrator, ", ".join(INTEGRATORS)))
        levels = {"Euler": 0, "PFE": 1, "TPI2": 2, "TPI3": 3}[req.integrator]
        if len(req.inner_steps) < levels or len(req.inner_dt) < levels or \
                any(int(k) < 1 for k in req.inner_steps[:levels]):
            raise ValueError("request %r: %s needs %d inner levels with at least one step each, "
                             "got inner_steps=%r inner_dt=%r" % (req.rid, req.integrator, levels,
                                                                 req.inner_steps, req.inner_dt))
        if req.nbytes() > self.budget:
            raise ValueError("request %r needs %d bytes, above the HBM budget of %d"
                             % (req.rid, req.nbytes(), self.budget))
        if self.max_cells is not None and req.cells > self.max_cells:
            raise ValueError("request %r has %d cells, above max_cells_per_class = %d"
                             % (req.rid, req.cells, self.max_cells))
        n = int(round(req.t_final / req.dt))
        if abs(n * req.dt - req.t_final) > 1e-5 * req.t_final:
            raise Exception("dt does not divide (tend-tstart) and dt is fixed!")
        self.queue.append(req)

    def _now(self):
        return self.clock() - self.t0

    def _class(self, key, proto):
        cls = self.classes.get(key)
        if cls is None:
            cls = self.classes[key] = _Class(key, proto, self.autotune,
                                             scratch_budget=max(1 << 30, self.budget // 4))
            self.reserved[key] = 0
        return cls

    def _scratch_need(self, key, cells):
        """Plan scratch the class will hold when it steps `cells` cells (never below what it
        holds or has reserved: the plan's buffers only grow)."""
        need, held = self.classes[key].coll.scratch_need(cells, grouped=True)
        return max(need, held, self.reserved[key])

    def _evict_idle(self, keep, admitted):
        """Release the plans of classes with no active or newly admitted problem."""
        freed = 0
        for key in [k for k, c in self.classes.items()
                    if k != keep and not c.active and k not in admitted]:
            freed += self.reserved.pop(key, 0)
            del self.classes[key]
        self.used -= freed
        return freed > 0

    def _admit(self, now):
        """Admit arrived requests in arrival order while memory allows."""
        admitted = {}
        waiting = []
        cells = {}  # class key -> resident + newly admitted cells
        for r in sorted(self.queue, key=lambda r: r.arrival):
            if r.arrival > now:
                waiting.append(r)
                continue
            key = r.class_key()
            cls = self._class(key, r)
            if key not in cells:
                cells[key] = sum(a.cells for a in cls.active)
            fits_cls = self.max_cells is None or cells[key] + r.cells <= self.max_cells
            res = self._scratch_need(key, cells[key] + r.cells)
            grow = r.nbytes() + res - self.reserved[key]
            if fits_cls and self.used + grow > self.budget and self._evict_idle(key, admitted):
                res = self._scratch_need(key, cells[key] + r.cells)
                grow = r.nbytes() + res - self.reserved[key]
            if not (fits_cls and self.used + grow <= self.budget):
                busy = any(c.active for c in self.classes.values()) or admitted
                if fits_cls and not busy:
                    raise ValueError("request %r needs %d bytes of resident arrays and plan scratch, "
                                     "above the HBM budget of %d" % (r.rid, self.used + grow, self.budget))
                waiting.append(r)
                continue
            cells[key] += r.cells
            n = int(round(r.t_final / r.dt))  # submit() checked that dt divides t_final
            a = _Active(req=r, start=0, cells=r.cells, steps_left=n, admitted_at=now)
            admitted.setdefault(key, []).append((a, r.f0))
            self.used += grow
            self.reserved[key] = res
            self.peak_used = max(self.peak_used, self.used)
        self.queue = waiting
        return admitted

    def run(self, requests=None, realtime=True):
        """Serve until every request is finished; returns {rid: Result}."""
        for r in requests or []:
            self.submit(r)
        t = dev.require_gpu()
        self.t0 = self.clock()
        while self.queue or any(c.active for c in self.classes.values()):
            now = self._now() if realtime else float("inf")
            new = self._admit(now)
            for key, items in new.items():
                cls = self.classes[key]
                cls.rebuild(cls.active, items)
            busy = [c for c in self.classes.values() if c.active]
            if not busy:
                nxt = min(r.arrival for r in self.queue)
                time.sleep(max(0.0, min(nxt - self._now(), 0.01)))
                continue
            # Steps this tick: up to the first completion (no admission or retirement falls
            # inside a tick); while requests are still to arrive, also about as long as the wait
            # for the next arrival (at least one step), so arrivals are admitted as before.
            steps = min(a.steps_left for c in busy for a in c.active)
            if realtime and self.queue:
                horizon = max(self.tick_s, min(r.arrival for r in self.queue) - self._now())
                steps = min(steps, max(1, int(horizon / max(self._step_s, 1e-5))))
            main = t.cuda.current_stream()
            tick_start = time.perf_counter()
            for gk in sorted({c.proto.group_key() for c in busy}, key=repr):
                grp = self.groups.setdefault(gk, _Group(gk, self.use_graph))
                grp.classes = [c for c in busy if c.proto.group_key() == gk]
                grp.stream.wait_stream(main)  # the classes' tensors were rebuilt on `main`
                with t.cuda.stream(grp.stream):
                    for _ in range(steps):
                        grp.step()
            self.ticks += 1
            for cls in busy:
                for a in cls.active:
                    for _ in range(steps):
                        a.steps_left -= 1
                        a.t += cls.proto.dt
            t.cuda.synchronize()
            dt_tick = (time.perf_counter() - tick_start) / steps
            self._step_s = dt_tick if self._step_s <= 0 else 0.8 * self._step_s + 0.2 * dt_tick
            done_t = self._now()
            for cls in busy:
                done = [a for a in cls.active if a.steps_left <= 0]
                if not done:
                    continue
                # one device-to-host copy per class when several problems finish
                host = cls.f.cpu().numpy() if len(done) > 4 else None
                for a in done:
                    if host is not None:
                        fh = host[a.start:a.start + a.cells].reshape(a.req.f0.shape).copy()
                    else:
                        fh = cls.f[a.start:a.start + a.cells].cpu().numpy().reshape(a.req.f0.shape)
                    self.results[a.req.rid] = Result(a.req.rid, fh, a.t, a.req.arrival,
                                                     a.admitted_at, done_t)
                    self.used -= a.req.nbytes()
                cls.rebuild([a for a in cls.active if a.steps_left > 0], [])
        return self.results


def latency_stats(results):
    lat = np.array([r.latency for r in results.values()])
    span = max(r.finished for r in results.values()) - min(r.arrival for r in results.values())
    return {"p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "solves_per_s": len(lat) / span if span > 0 else float("inf"), "n": len(lat)}


def poisson_arrivals(n, rate, rng):
    """Arrival times of a Poisson process of the given rate (1/s)."""
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
