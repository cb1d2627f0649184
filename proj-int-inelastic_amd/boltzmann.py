"""Boltzmann solvers with device-resident state (reference ``src/boltzmann.py``).

Same attributes and integrators as the reference: ``time_integrator`` in
{Euler, PFE, TPI2, TPI3, CONSP, CONSP_M} (:46-80), ``inner_steps``,
``inner_dt``, ``num_levels``, ``dt``, ``order``, ``limiters``, ``kn``,
``penalty``, ``e``, ``tau``; ``df(state, dt)`` = advection + dt Q(f)/kn
(:160-164) and ``df_advection`` (:177-205).  Every array operation is a HIP
kernel (ib_tvd1d, ib_collide, ib_df_combine, ib_axpy) on the current stream.

New:
  * ``BoltzmannSolver0D`` for space-homogeneous problems (the reference's
    ``df_advection`` raises there, :166-167); f may carry leading batch axes.
  * ``use_graph = True`` captures one whole outer step (all inner projective
    steps) in a HIP graph after one eager warm-up step and replays it.
  * ``BoltzmannSolver2D`` for 2D space (f of shape (nx, ny, N, N)): the
    reference keeps ghost cells for two space dimensions (solver.py:51-105,
    state.py:95-122) but has no 2D ``df_advection``; here it is the
    dimension-split sum df_x + df_y of the 1D step along x (speed v_x,
    ``problem_data["v1"]``) and along y (speed v_y, ``problem_data["v2"]``),
    each with its own boundary conditions.
"""
from math import pi

from . import _lib
from . import device as dev
from .limiters import tvd
from .solver import Solver


def maxwellian(vx, vy, rho, u, T):
    """2D Maxwellian per x-cell (reference boltzmann.py:9-13), device tensors."""
    t = dev.torch()
    vx = dev.to_device(vx)
    vy = dev.to_device(vy)
    rho = dev.to_device(rho)[:, None, None]
    u = dev.to_device(u)[:, None, None]
    T = dev.to_device(T)[:, None, None]
    return rho / (2 * pi * T) * t.exp(-((vx - u) ** 2 + vy ** 2) / (2 * T))


class BoltzmannSolver(Solver):
    def __init__(self, kn=1.0, riemann_solver=None, collision_operator=None):
        self.kn = kn
        self.tau = None
        self.e = None
        self.num_ghost = 2
        self.order = 2
        self.limiters = tvd.minmod
        self.time_integrator = "Euler"
        self.penalty = None
        self.num_levels = 3
        self.inner_steps = [3, 3, 3]
        self.inner_dt = [kn, 5 * kn, 10 * kn]
        self.use_graph = False
        self._graph = None
        self._graph_key = None
        super(BoltzmannSolver, self).__init__(riemann_solver, collision_operator)

    def setup(self, solution):
        super(BoltzmannSolver, self).setup(solution)

    # -- boltzmann.py:46-80
    def step(self, solution, take_one_step, tstart, tend):
        if self.time_integrator == "CONSP" and not self.penalty:
            raise ValueError("The penalty parameter is not given. Please specify it!")
        if self.time_integrator == "CONSP_M" and not (self.penalty and self.e and (self.tau is not None)):
            raise ValueError("The penalty parameter, e or tau are not given. Please specify it!")
        fn = {
            "Euler": self.euler_step,
            "PFE": self.proj_euler_step,
            "TPI2": self.tel_proj_step2,
            "TPI3": self.tel_proj_step3,
            "CONSP": self.convex_splitting_step,
            "CONSP_M": self.convex_splitting_maxwellian_step,
        }.get(self.time_integrator)
        if fn is None:
            raise NotImplementedError("This time integrator is not implemented.")
        if self.use_graph and self.time_integrator in ("Euler", "PFE", "TPI2", "TPI3"):
            self._graph_step(fn, solution, take_one_step, tstart, tend)
        else:
            fn(solution, take_one_step, tstart, tend)

    def _graph_step(self, fn, solution, take_one_step, tstart, tend):
        t = dev.torch()
        state = solution.state
        key = (fn.__name__, state.f.data_ptr(), tuple(state.f.shape), self.dt,
               tuple(self.inner_steps), tuple(self.inner_dt), self.kn)
        if self._graph is None or self._graph_key != key:
            fn(solution, take_one_step, tstart, tend)  # eager warm-up: buffers sized
            t.cuda.synchronize()
            g = t.cuda.CUDAGraph()
            with t.cuda.graph(g):
                fn(solution, take_one_step, tstart, tend)
            # capture does not execute; the warm-up already advanced f by one step
            self._graph, self._graph_key = g, key
            return
        self._graph.replay()

    @staticmethod
    def _axpy(y, x, a):
        _lib.check(_lib.load().ib_axpy(y.data_ptr(), x.data_ptr(), float(a), y.numel(),
                                       dev.stream_handle()), "ib_axpy")

    # -- boltzmann.py:82-84
    def euler_step(self, solution, take_one_step, tstart, tend):
        state = solution.state
        self._axpy(state.f, self.df(state, self.dt), 1.0)

    # -- boltzmann.py:86-94
    def convex_splitting_step(self, solution, take_one_step, tstart, tend):
        state = solution.state
        if self.collision is not None:
            q = self.collision(state.f)
            self._axpy(state.f, q, self.dt / (self.kn + self.penalty * self.dt))
        adv = self.df_advection(state, self.dt)
        if adv is not None:
            self._axpy(state.f, adv, 1.0)

    # -- boltzmann.py:96-118
    def convex_splitting_maxwellian_step(self, solution, take_one_step, tstart, tend):
        t = dev.torch()
        state = solution.state
        if self.collision is not None:
            cc = state.c_centers
            vx, vy = cc[-2], cc[-1]
            if state.f.dim() == 3:
                vx, vy = vx[0], vy[0]
            rho, u, T = state.rho, state.u[0], state.T
            rho, u, T = [x.reshape(-1) for x in (rho, u, T)]
            e2 = 1 - self.e ** 2
            T_new = (T - 8 * self.tau / e2) * t.exp(-e2 / self.kn / 4 * rho * self.dt) + 8 * self.tau / e2
            dm = (maxwellian(vx, vy, rho, u, T_new) - maxwellian(vx, vy, rho, u, T)).reshape(state.f.shape)
            c1 = self.dt / (self.kn + self.penalty * self.dt)
            c2 = self.penalty * self.dt / (self.kn + self.penalty * self.dt)
            upd = c1 * self.collision(state.f) + c2 * dm
            self._axpy(state.f, upd, 1.0)
        adv = self.df_advection(state, self.dt)
        if adv is not None:
            self._axpy(state.f, adv, 1.0)

    # -- boltzmann.py:120-128
    def proj_euler_step(self, solution, take_one_step, tstart, tend):
        state = solution.state
        M = self.dt / self.inner_dt[0] - self.inner_steps[0]
        df = None
        for _ in range(self.inner_steps[0]):
            df = self.df(state, self.inner_dt[0])
            self._axpy(state.f, df, 1.0)
        self._axpy(state.f, df, M)

    @staticmethod
    def _axpy_chain(y, x, a, b=1.0, c=1.0):
        """y += a * (b * (c * x)) in one HIP kernel, rounding as those products do."""
        _lib.check(_lib.load().ib_axpy_chain(y.data_ptr(), x.data_ptr(), float(a), float(b), float(c),
                                             y.numel(), dev.stream_handle()), "ib_axpy_chain")

    # -- boltzmann.py:130-140 (df1 = M0 df0 as written there).  df1 is never stored: the last
    # df0 and the multipliers go to one chained update per use
    def tel_proj_step2(self, solution, take_one_step, tstart, tend):
        state = solution.state
        M1 = self.dt / self.inner_dt[1] - self.inner_steps[1]
        df0 = M0 = None
        for _ in range(self.inner_steps[1]):
            M0 = self.inner_dt[1] / self.inner_dt[0] - self.inner_steps[0]
            for _ in range(self.inner_steps[0]):
                df0 = self.df(state, self.inner_dt[0])
                self._axpy(state.f, df0, 1.0)
            self._axpy_chain(state.f, df0, M0)             # df1 = M0 df0; f += df1
        self._axpy_chain(state.f, df0, M1, M0)             # f += M1 df1

    # -- boltzmann.py:142-158
    def tel_proj_step3(self, solution, take_one_step, tstart, tend):
        state = solution.state
        M2 = self.dt / self.inner_dt[2] - self.inner_steps[2]
        df0 = M0 = M1 = None
        for _ in range(self.inner_steps[2]):
            M1 = self.inner_dt[2] / self.inner_dt[1] - self.inner_steps[1]
            for _ in range(self.inner_steps[1]):
                M0 = self.inner_dt[1] / self.inner_dt[0] - self.inner_steps[0]
                for _ in range(self.inner_steps[0]):
                    df0 = self.df(state, self.inner_dt[0])
                    self._axpy(state.f, df0, 1.0)
                self._axpy_chain(state.f, df0, M0)         # df1 = M0 df0; f += df1
            self._axpy_chain(state.f, df0, 1.0, M1, M0)    # df2 = M1 df1; f += df2
        self._axpy_chain(state.f, df0, M2, M1, M0)         # f += M2 df2

    # -- boltzmann.py:160-164
    def df(self, state, dt):
        t = dev.torch()
        adv = self.df_advection(state, dt)
        if self.collision is None:
            return adv if adv is not None else t.zeros_like(state.f)
        q = self.collision(state.f)
        out = t.empty_like(state.f)
        _lib.check(_lib.load().ib_df_combine(dev.ptr(adv), q.data_ptr(), out.data_ptr(), float(dt),
                                             float(self.kn), out.numel(), dev.stream_handle()),
                   "ib_df_combine")
        return out

    def df_advection(self, state, dt):
        raise NotImplementedError("You must subclass the BoltzmannSolver.")


class BoltzmannSolver0D(BoltzmannSolver):
    """Space-homogeneous problems: no transport, df = dt Q(f) / kn."""

    def __init__(self, kn=1.0, riemann_solver=None, collision_operator=None):
        self.num_dim_x = 0
        super(BoltzmannSolver0D, self).__init__(kn, riemann_solver, collision_operator)

    def df_advection(self, state, dt):
        return None


class BoltzmannSolver1D(BoltzmannSolver):
    def __init__(self, kn=1, riemann_solver=None, collision_operator=None):
        self.num_dim_x = 1
        self._v1_dev = None
        super(BoltzmannSolver1D, self).__init__(kn, riemann_solver, collision_operator)

    # -- boltzmann.py:177-205 fused into one stencil kernel
    def df_advection(self, state, dt):
        t = dev.torch()
        rp = getattr(self, "rp", None)
        if rp is None or getattr(rp, "kernel_rp", getattr(rp, "__name__", None)) != "advection_1D":
            raise NotImplementedError("df_advection runs the advection_1D Riemann solver on the device")
        lo, hi = self._bc_codes(0)
        f = state.f
        grid = state.grid
        dtdx = dt / grid.delta[0]
        if self._v1_dev is None or self._v1_src is not state.problem_data["v1"]:
            self._v1_src = state.problem_data["v1"]
            self._v1_dev = dev.to_device(self._v1_src).reshape(-1)
        nx = f.shape[0]
        nv = f[0].numel()
        vx_rows = f.shape[1] if f.dim() > 1 else 1
        df = t.empty_like(f)
        _lib.check(_lib.load().ib_tvd1d(f.data_ptr(), df.data_ptr(), nx, nv, self._v1_dev.data_ptr(),
                                        float(dtdx), int(self.order), tvd.limiter_code(self.limiters),
                                        int(lo), int(hi), int(vx_rows), dev.stream_handle()),
                   "ib_tvd1d")
        return df


class BoltzmannSolver2D(BoltzmannSolver):
    def __init__(self, kn=1, riemann_solver=None, collision_operator=None):
        self.num_dim_x = 2
        self._vdev = {}
        super(BoltzmannSolver2D, self).__init__(kn, riemann_solver, collision_operator)

    def _speed(self, state, key):
        src = state.problem_data[key]
        hit = self._vdev.get(key)
        if hit is None or hit[0] is not src:
            hit = (src, dev.to_device(src).reshape(-1))
            self._vdev[key] = hit
        return hit[1]

    def df_advection(self, state, dt):
        """df_x + df_y, each the TVD step of boltzmann.py:177-205 along its space axis."""
        t = dev.torch()
        rp = getattr(self, "rp", None)
        if rp is None or getattr(rp, "kernel_rp", getattr(rp, "__name__", None)) != "advection_1D":
            raise NotImplementedError("df_advection runs the advection_1D Riemann solver on the device")
        f = state.f
        if f.dim() < 3:
            raise ValueError("2D-space f must have shape (nx, ny, velocity...)")
        nx, ny = f.shape[0], f.shape[1]
        nvel = f[0, 0].numel()
        vx, vy = self._speed(state, "v1"), self._speed(state, "v2")
        if vx.numel() != nvel or vy.numel() != nvel:
            raise ValueError("v1 and v2 must hold one speed per velocity point")
        lim = tvd.limiter_code(self.limiters)
        df = t.empty_like(f)
        lib = _lib.load()
        st = dev.stream_handle()
        for axis, (outer, n, inner, v) in enumerate(((1, nx, ny * nvel, vx), (nx, ny, nvel, vy))):
            lo, hi = self._bc_codes(axis)
            dtdx = float(dt / state.grid.delta[axis])
            _lib.check(lib.ib_tvd_axis(f.data_ptr(), df.data_ptr(), outer, n, inner, v.data_ptr(),
                                       nvel, dtdx, int(self.order), lim, int(lo), int(hi), axis,
                                       st), "ib_tvd_axis")
        return df
