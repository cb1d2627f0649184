"""GPU solver step API against the oracle integrators (FP64).

Homogeneous cooling (2D 32^2 inelastic Maxwell molecules, forward Euler),
1D x 2D projective forward Euler (Sod-type Maxwellian Riemann data, hard
spheres), telescopic projective steps on a batch of homogeneous problems with
per-problem restitution, and HIP-graph replay of whole outer steps.
"""
import math
import types

import numpy as np
import pytest

from oracle import collision as OC
from oracle import integrators as OI
from oracle import quadrature as OQ
from oracle import transport as OT

pytestmark = pytest.mark.gpu

S1 = 2.718491035932
S2 = 3.624654714576


def _coll_cfg(N, nr, M, s, e, gamma):
    return {"collision-model": {"dim": 2, "e": e, "e0": e, "gamma": gamma},
            "velocity-mesh": {"nv": N, "s": s, "nr": nr},
            "circle-quad-rule": {"nphi": M}}


def _two_gauss(v, dv, seed):
    rng = np.random.default_rng(seed)
    c1 = np.array([1.0, 0.0]) + rng.normal(0, 0.1, 2)
    c2 = -np.array([1.0, 0.0]) + rng.normal(0, 0.1, 2)
    T = 0.5
    g = lambda c: np.exp(-((v[0] - c[0]) ** 2 + (v[1] - c[1]) ** 2) / (2 * T)) / (2 * math.pi * T)
    f = 0.5 * (g(c1) + g(c2))
    return f / (f.sum() * dv ** 2)


def test_homogeneous_cooling_euler(pkg, need_gpu):
    from proj_int_inelastic_amd.boltzmann import BoltzmannSolver0D
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    from proj_int_inelastic_amd.solution import Solution
    e, N = 0.8, 32
    cfg = _coll_cfg(N, 32, 8, S1, e, 0.0)
    sm = SpectralMesh(cfg)
    coll = FSInelasticVHSCollision(cfg, sm, with_loss=True)
    sol = Solution(types.SimpleNamespace(dim_x=0, dim_v=2, xmin=0.0, xmax=1.0, l=sm.L, nx=1, nv=N))
    op = OC.OracleOperator(2, N, S1, 32, *OQ.circle_rule(8), e=e, e0=e, gamma=0.0)
    v, dv = op.v_grids(), op.delta
    f0 = _two_gauss(v, dv, seed=0)
    sol.state.f = f0.copy()
    solver = BoltzmannSolver0D(kn=1.0, collision_operator=coll)
    solver.dt = 0.01
    temps = [float(sol.state.T)]

    def record(slv, state):
        temps.append(float(state.T))

    solver.before_step = record
    solver.evolve_to_time(sol, 1.0)
    temps.append(float(sol.state.T))
    fg = sol.state.get_f_global()
    # oracle forward Euler, 100 steps
    f = f0.copy()
    tref = [float(OC.temperature(f, v, dv))]
    for _ in range(100):
        f = f + 0.01 * op.collide(f, with_loss=True)
        tref.append(float(OC.temperature(f, v, dv)))
    err = np.abs(fg - f).max() / np.abs(f).max()
    print("cooling f rel Linf", err, "T(1)/T(0)", temps[-1] / temps[0])
    assert err < 1e-12
    assert np.allclose(temps[1:], tref, rtol=1e-12, atol=0)
    assert abs(sol.t - 1.0) < 1e-12
    # Haff: inelastic Maxwell molecules cool as T0 exp(-(1 - e^2) rho t / 4); the
    # M = 8 angular rule resolves the dissipation rate to about 7 %
    rate = -math.log(temps[-1] / temps[0])
    assert abs(rate / ((1 - e * e) / 4) - 1) < 0.1


def test_homogeneous_cooling_projected_conserves(pkg, need_gpu):
    """With the FP64 conservative projection mass and momentum stay at round-off."""
    import torch
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    for prec in ("fp64", "fp32"):
        cfg = _coll_cfg(32, 32, 8, S1, 0.8, 0.0)
        sm = SpectralMesh(cfg)
        coll = FSInelasticVHSCollision(cfg, sm, precision=prec)
        op = OC.OracleOperator(2, 32, S1, 32, *OQ.circle_rule(8), e=0.8, e0=0.8)
        f = torch.from_numpy(_two_gauss(op.v_grids(), op.delta, 1)).cuda()
        m0 = coll.moments(f)
        for _ in range(50):
            q = coll.collide(f, with_loss=True, project=True)
            f += 0.01 * q
        m1 = coll.moments(f)
        drift = (m1 - m0).abs().cpu().numpy()
        assert drift[0] <= 1e-12 * float(m0[0]), (prec, drift)
        assert drift[1:3].max() <= 1e-12 * float(m0[0]) * sm.L, (prec, drift)


def _sod(v, nx):
    x = (np.arange(nx) + 0.5) / nx
    rho = np.where(x < 0.5, 1.0, 0.125)
    T = np.where(x < 0.5, 1.0, 0.8)
    vx, vy = np.meshgrid(v, v, indexing="ij")
    return rho[:, None, None] / (2 * math.pi * T[:, None, None]) * np.exp(
        -(vx ** 2 + vy ** 2) / (2 * T[:, None, None]))


def _solver1d(pkg, coll, nx, N, l, kn, integ, inner_steps, inner_dt, dt):
    from proj_int_inelastic_amd.boltzmann import BoltzmannSolver1D
    from proj_int_inelastic_amd.riemann.advection_1D import advection_1D
    from proj_int_inelastic_amd.solution import Solution
    from proj_int_inelastic_amd.solver import BC
    sol = Solution(types.SimpleNamespace(dim_x=1, dim_v=2, xmin=0.0, xmax=1.0, l=l, nx=nx, nv=N))
    s = BoltzmannSolver1D(kn=kn, riemann_solver=advection_1D, collision_operator=coll)
    s.bc_lower[0] = BC.extrap
    s.bc_upper[0] = BC.extrap
    s.time_integrator = integ
    s.inner_steps = list(inner_steps)
    s.inner_dt = list(inner_dt)
    s.dt = dt
    return sol, s


def test_pfe_1d_riemann_parity(pkg, need_gpu):
    """Projective forward Euler on 200 cells x 32^2 (gain-only Q as the reference's __call__)."""
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    nx, N, e, kn = 200, 32, 0.9, 1e-3
    cfg = _coll_cfg(N, 32, 8, S2, e, 1.0)
    sm = SpectralMesh(cfg)
    coll = FSInelasticVHSCollision(cfg, sm)
    dx = 1.0 / nx
    dt = 0.4 * dx / 8
    sol, s = _solver1d(pkg, coll, nx, N, sm.L, kn, "PFE", [2], [1e-3], dt)
    v = sm.v_center
    f0 = _sod(v, nx)
    sol.state.f = f0.copy()
    s.evolve_to_time(sol)  # one step
    s.evolve_to_time(sol)
    fg = sol.state.get_f_global()
    op = OC.OracleOperator(2, N, S2, 32, *OQ.circle_rule(8), e=e, e0=e, gamma=1.0)
    v1 = np.repeat(v[:, None], N, axis=1)

    def df(f, h):
        d = OT.df_advection(f, v1, h, dx, OT.EXTRAP, OT.EXTRAP, 2, OT.minmod)
        return d + h * op.collide(f) / kn

    f = f0.copy()
    for _ in range(2):
        f = OI.proj_euler_step(f, df, dt, [2], [1e-3])
    err = np.abs(fg - f).max() / np.abs(f).max()
    print("PFE 1D rel Linf", err)
    assert err < 1e-12


def test_tpi2_homogeneous_batch_per_problem_e(pkg, need_gpu):
    import torch
    from proj_int_inelastic_amd.boltzmann import BoltzmannSolver0D
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    from proj_int_inelastic_amd.solution import Solution
    N, es = 16, np.array([0.3, 0.8, 1.0])
    cfg = _coll_cfg(N, 16, 8, 2.5, 1.0, 1.0)
    sm = SpectralMesh(cfg)
    coll = FSInelasticVHSCollision(cfg, sm)
    op = [OC.OracleOperator(2, N, 2.5, 16, *OQ.circle_rule(8), e=x, e0=x, gamma=1.0) for x in es]
    v, dv = op[0].v_grids(), op[0].delta
    f0 = np.stack([_two_gauss(v, dv, seed=s) for s in range(3)])
    sol = Solution(types.SimpleNamespace(dim_x=0, dim_v=2, xmin=0.0, xmax=1.0, l=sm.L, nx=1, nv=N))
    sol.state.f = f0.copy()
    solver = BoltzmannSolver0D(kn=1.0, collision_operator=lambda f: coll.collide(f, with_loss=True, e=es))
    solver.time_integrator = "TPI2"
    solver.inner_steps = [2, 2]
    solver.inner_dt = [0.01, 0.05]
    solver.dt = 0.2
    solver.evolve_to_time(sol, 0.4)
    fg = sol.state.get_f_global()

    def df(f, h):
        return np.stack([h * op[c].collide(f[c], with_loss=True) for c in range(3)])

    f = f0.copy()
    for _ in range(2):
        f = OI.tel_proj_step2(f, df, 0.2, [2, 2], [0.01, 0.05])
    err = np.abs(fg - f).max() / np.abs(f).max()
    assert err < 1e-12, err


def test_tpi3_homogeneous_batch_per_problem_e(pkg, need_gpu):
    """TPI3 (boltzmann.py:142-158) with the reference's default schedule shape: K = (3, 3, 3),
    inner dt (0.01, 0.05, 0.1), so M0 = 2, M1 = -1 (df2 = M1 df1 as written) and M2 = 2 at
    outer dt 0.5; three problems with their own e, against the oracle's tel_proj_step3."""
    from proj_int_inelastic_amd.boltzmann import BoltzmannSolver0D
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    from proj_int_inelastic_amd.solution import Solution
    N, es = 16, np.array([0.25, 0.7, 0.95])
    cfg = _coll_cfg(N, 16, 8, 2.5, 1.0, 1.0)
    sm = SpectralMesh(cfg)
    coll = FSInelasticVHSCollision(cfg, sm)
    op = [OC.OracleOperator(2, N, 2.5, 16, *OQ.circle_rule(8), e=x, e0=x, gamma=1.0) for x in es]
    v, dv = op[0].v_grids(), op[0].delta
    f0 = np.stack([_two_gauss(v, dv, seed=10 + s) for s in range(3)])
    sol = Solution(types.SimpleNamespace(dim_x=0, dim_v=2, xmin=0.0, xmax=1.0, l=sm.L, nx=1, nv=N))
    sol.state.f = f0.copy()
    solver = BoltzmannSolver0D(kn=1.0, collision_operator=lambda f: coll.collide(f, with_loss=True, e=es))
    solver.time_integrator = "TPI3"
    solver.inner_steps = [3, 3, 3]
    solver.inner_dt = [0.01, 0.05, 0.1]
    solver.dt = 0.5
    solver.evolve_to_time(sol, 1.0)
    fg = sol.state.get_f_global()

    def df(f, h):
        return np.stack([h * op[c].collide(f[c], with_loss=True) for c in range(3)])

    f = f0.copy()
    for _ in range(2):
        f = OI.tel_proj_step3(f, df, 0.5, [3, 3, 3], [0.01, 0.05, 0.1])
    err = np.abs(fg - f).max() / np.abs(f).max()
    print("TPI3 rel Linf", err)
    assert err < 1e-12, err


def test_graph_replay_matches_eager(pkg, need_gpu):
    import torch
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    nx, N = 40, 16
    cfg = _coll_cfg(N, 8, 8, 3.0, 0.9, 1.0)
    sm = SpectralMesh(cfg)
    coll = FSInelasticVHSCollision(cfg, sm)
    f0 = _sod(sm.v_center, nx)
    results = []
    for use_graph in (False, True):
        sol, s = _solver1d(pkg, coll, nx, N, sm.L, 1e-2, "PFE", [2], [1e-2], 0.05)
        s.use_graph = use_graph
        sol.state.f = f0.copy()
        snaps = []
        for _ in range(4):
            s.evolve_to_time(sol)
            torch.cuda.synchronize()
            snaps.append(sol.state.get_f_global())
            # eager work between replays
            _ = coll.collide(sol.state.f)
        results.append(snaps)
    for a, b in zip(*results):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("integ", ["CONSP", "CONSP_M"])
def test_convex_splitting_steps(pkg, need_gpu, integ):
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    nx, N, e, kn = 40, 16, 0.8, 1e-2
    cfg = _coll_cfg(N, 16, 8, 3.0, e, 0.0)
    sm = SpectralMesh(cfg)
    coll = FSInelasticVHSCollision(cfg, sm, with_loss=True)
    sol, s = _solver1d(pkg, coll, nx, N, sm.L, kn, integ, [2], [1e-2], 0.01)
    s.penalty = 5.0
    s.e = e
    s.tau = 0.0
    f0 = _sod(sm.v_center, nx)
    sol.state.f = f0.copy()
    s.evolve_to_time(sol, 0.02)
    fg = sol.state.get_f_global()
    op = OC.OracleOperator(2, N, 3.0, 16, *OQ.circle_rule(8), e=e, e0=e, gamma=0.0)
    v = sm.v_center
    v1 = np.repeat(v[:, None], N, axis=1)
    dx = 1.0 / nx
    adv = lambda f, h: OT.df_advection(f, v1, h, dx, OT.EXTRAP, OT.EXTRAP, 2, OT.minmod)
    col = lambda f: op.collide(f, with_loss=True)
    f = f0.copy()
    vx = np.broadcast_to(v1, f.shape)
    vy = np.broadcast_to(v1.T, f.shape)
    for _ in range(2):
        if integ == "CONSP":
            f = OI.convex_splitting_step(f, col, adv, 0.01, kn, 5.0)
        else:
            f = OI.convex_splitting_maxwellian_step(f, col, adv, 0.01, kn, 5.0, e, 0.0, vx, vy, sm.delta)
    err = np.abs(fg - f).max() / np.abs(f).max()
    assert err < 1e-12, err
    with pytest.raises(ValueError):
        s.penalty = None
        s.evolve_to_time(sol)


# ---- 2D space (dimension-split transport), BoltzmannSolver2D

def _solver2d(N, l, nx, ny_eq_nx=True, bcx=(1, 1), bcy=(1, 1), order=2, limiter="minmod",
              coll=None, kn=1.0):
    from proj_int_inelastic_amd.boltzmann import BoltzmannSolver2D
    from proj_int_inelastic_amd.limiters import tvd
    from proj_int_inelastic_amd.riemann.advection_1D import advection_1D
    from proj_int_inelastic_amd.solution import Solution
    sol = Solution(types.SimpleNamespace(dim_x=2, dim_v=2, xmin=0.0, xmax=1.0, l=l, nx=nx, nv=N))
    s = BoltzmannSolver2D(kn=kn, riemann_solver=advection_1D, collision_operator=coll)
    s.order = order
    s.limiters = tvd.minmod if limiter == "minmod" else tvd.mc
    s.bc_lower[0], s.bc_upper[0] = bcx
    s.bc_lower[1], s.bc_upper[1] = bcy
    return sol, s


@pytest.mark.parametrize("order,limiter", [(1, "minmod"), (2, "minmod"), (2, "mc")])
@pytest.mark.parametrize("bcx,bcy", [((1, 1), (2, 2)), ((2, 2), (1, 2))])
def test_2d_space_transport_bit_exact(pkg, need_gpu, order, limiter, bcx, bcy):
    """BoltzmannSolver2D.df_advection (two ib_tvd_axis sweeps) equals the oracle's
    dimension-split increment bit for bit."""
    N, nx, l = 16, 18, 6.0
    sol, s = _solver2d(N, l, nx, bcx=bcx, bcy=bcy, order=order, limiter=limiter)
    rng = np.random.default_rng(order * 7 + bcx[0] + 3 * bcy[0])
    f = rng.random((nx, nx, N, N))
    f[4:7] = 0.25
    sol.state.f = f.copy()
    dt = 0.2 / nx / l
    df = s.df_advection(sol.state, dt).cpu().numpy()
    v1, v2 = sol.state.problem_data["v1"], sol.state.problem_data["v2"]
    lim = OT.minmod if limiter == "minmod" else OT.mc
    ref = OT.df_advection_2d(f, v1, v2, dt, 1.0 / nx, 1.0 / nx, bcx, bcy, order, lim)
    assert np.array_equal(df, ref)


def test_2d_space_euler_with_collisions(pkg, need_gpu):
    """Euler steps on 12 x 12 cells x 16^2 velocities with gain+loss Q against the oracle."""
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    N, nx, kn, e = 16, 14, 0.1, 0.8
    cfg = _coll_cfg(N, 8, 8, 2.0, e, 1.0)
    sm = SpectralMesh(cfg)
    coll = FSInelasticVHSCollision(cfg, sm)
    sol, s = _solver2d(N, sm.L, nx, bcx=(2, 2), bcy=(1, 1),
                       coll=lambda f: coll.collide(f, with_loss=True), kn=kn)
    s.time_integrator = "Euler"
    dt = 0.1 / nx / sm.L
    s.dt = dt
    v = sm.v_center
    vxg, vyg = np.meshgrid(v, v, indexing="ij")
    xs = (np.arange(nx) + 0.5) / nx
    rho = 1.0 + 0.4 * np.exp(-((xs[:, None] - 0.5) ** 2 + (xs[None, :] - 0.4) ** 2) / 0.02)
    f0 = rho[:, :, None, None] / (2 * np.pi) * np.exp(-((vxg - 0.3) ** 2 + vyg ** 2) / 2)[None, None]
    sol.state.f = f0.copy()
    s.evolve_to_time(sol, 2 * dt)
    fg = sol.state.get_f_global()
    op = OC.OracleOperator(2, N, 2.0, 8, *OQ.circle_rule(8), e=e, e0=e, gamma=1.0)
    v1, v2 = sol.state.problem_data["v1"], sol.state.problem_data["v2"]
    f = f0.copy()
    for _ in range(2):
        adv = OT.df_advection_2d(f, v1, v2, dt, 1.0 / nx, 1.0 / nx, (2, 2), (1, 1), 2, OT.minmod)
        f = f + (adv + dt * op.collide(f, with_loss=True) / kn)
    err = np.abs(fg - f).max() / np.abs(f).max()
    print("2D x 2D Euler rel Linf", err)
    assert err < 1e-12
