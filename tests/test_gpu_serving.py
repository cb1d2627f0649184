"""Serving loop: each result is the same alone and in the mix (continuous
batching with admission by memory budget and per-problem restitution)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gauss2(N, L, seed):
    rng = np.random.default_rng(seed)
    dv = 2 * L / N
    v = -L + (np.arange(N) + 0.5) * dv
    vx, vy = np.meshgrid(v, v, indexing="ij")
    c = rng.normal(0, 0.3, 2)
    f = np.exp(-((vx - 1 - c[0]) ** 2 + (vy - c[1]) ** 2)) + np.exp(-((vx + 1) ** 2 + vy ** 2) / 0.8)
    return f / (f.sum() * dv * dv)


def _riemann(nx, N, L):
    dv = 2 * L / N
    v = -L + (np.arange(N) + 0.5) * dv
    vx, vy = np.meshgrid(v, v, indexing="ij")
    x = (np.arange(nx) + 0.5) / nx
    rho = np.where(x < 0.5, 1.0, 0.125)[:, None, None]
    T = np.where(x < 0.5, 1.0, 0.8)[:, None, None]
    return rho / (2 * math.pi * T) * np.exp(-(vx ** 2 + vy ** 2) / (2 * T))


def _requests(Request, L):
    reqs = []
    for i, (e, tf, g) in enumerate([(0.3, 0.5, 0.0), (0.9, 1.0, 1.0), (0.6, 0.75, 1.0), (1.0, 0.25, 0.0)]):
        reqs.append(Request(rid=i, f0=_gauss2(16, L, i), t_final=tf, e=e, gamma=g, n=16, nr=16, m=8, s=3.0,
                            integrator="TPI2", inner_steps=(2, 2), inner_dt=(0.01, 0.05), dt=0.25,
                            arrival=0.0))
    for i, e in enumerate([0.7, 0.95]):
        reqs.append(Request(rid=10 + i, f0=_riemann(40, 16, L), t_final=0.02, e=e, n=16, nr=16, m=8,
                            s=3.0, kn=1e-2, integrator="PFE", inner_steps=(2,), inner_dt=(1e-3,),
                            dt=0.01, xlen=1.0, arrival=0.0))
    return reqs


def test_alone_equals_mix(pkg, need_gpu):
    from proj_int_inelastic_amd import serving
    L = 0.5 * (3 + math.sqrt(2)) * 3.0
    mix = serving.Server().run(_requests(serving.Request, L), realtime=False)
    for r in _requests(serving.Request, L):
        alone = serving.Server().run([r], realtime=False)[r.rid]
        a, b = alone.f, mix[r.rid].f
        err = np.abs(a - b).max() / np.abs(a).max()
        assert err < 1e-12, (r.rid, err)
        assert abs(alone.t - r.t_final) < 1e-12


def _class_need(req, cells):
    """Plan scratch a fresh operator of the request's class reports for `cells` cells."""
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    cfg = {"collision-model": {"dim": req.dim, "e": 1.0, "e0": 1.0, "gamma": req.gamma},
           "velocity-mesh": {"nv": req.n, "s": req.s, "nr": req.nr}}
    if req.dim == 2:
        cfg["circle-quad-rule"] = {"nphi": req.m}
    else:
        cfg["spherical-design-rule"] = {"ssrule": "lebedev", "nsphr": req.m}
    c = FSInelasticVHSCollision(cfg, SpectralMesh(cfg), precision=req.precision)
    return c.scratch_need(cells, grouped=True)[0]


def test_backpressure_queue(pkg, need_gpu):
    """Room for two of four problems (resident arrays and the plan scratch of two cells):
    the others wait, and every result equals the unconstrained server's."""
    from proj_int_inelastic_amd import serving
    L = 0.5 * (3 + math.sqrt(2)) * 3.0
    reqs = _requests(serving.Request, L)[:4]
    budget = 2 * reqs[0].nbytes() + _class_need(reqs[0], 2)
    srv = serving.Server(hbm_budget=budget)
    res = srv.run(reqs, realtime=False)
    assert sorted(res) == [0, 1, 2, 3]
    free = serving.Server().run(_requests(serving.Request, L)[:4], realtime=False)
    for k in res:
        assert np.abs(res[k].f - free[k].f).max() <= 1e-12 * np.abs(free[k].f).max()
    assert srv.peak_used <= budget
    assert srv.used == srv.reserved[reqs[0].class_key()]   # only the class's plan scratch is left


def test_admission_counts_plan_scratch(pkg, need_gpu):
    """Two 3D 32^3 FP32 problems whose resident arrays fit together but whose plan scratch for
    two cells (fused partial planes) does not: the second queues until the first retires;
    results equal the unconstrained run, and a problem that cannot fit alone is refused."""
    import dataclasses
    from proj_int_inelastic_amd import serving
    L = 0.5 * (3 + math.sqrt(2)) * S3
    reqs = [serving.Request(rid=i, f0=_gauss_nd(3, 32, L, 90 + i), t_final=0.25, e=0.5 + 0.3 * i, dim=3, n=32,
                            nr=8, m=14, s=S3, gamma=1.0, precision="fp32", integrator="TPI2",
                            inner_steps=(2, 2), inner_dt=(0.01, 0.05), dt=0.25) for i in range(2)]
    one, two = _class_need(reqs[0], 1), _class_need(reqs[0], 2)
    assert two > one > 2 * reqs[0].nbytes()
    budget = 2 * reqs[0].nbytes() + (one + two) // 2
    srv = serving.Server(hbm_budget=budget)
    res = srv.run([dataclasses.replace(r) for r in reqs], realtime=False)
    free_srv = serving.Server()
    free = free_srv.run([dataclasses.replace(r) for r in reqs], realtime=False)
    assert srv.ticks == 2 * free_srv.ticks            # one after the other
    assert srv.peak_used <= budget
    for k in free:
        assert np.abs(res[k].f - free[k].f).max() <= 1e-12 * np.abs(free[k].f).max()
    with pytest.raises(ValueError, match="above the HBM budget"):
        serving.Server(hbm_budget=reqs[0].nbytes() + one // 2).run([dataclasses.replace(reqs[0])],
                                                                   realtime=False)


def test_graph_replay_and_mixed_shapes(pkg, need_gpu):
    """Grouped launches over mixed shapes with HIP-graph replay equal the eager loop."""
    from proj_int_inelastic_amd import serving
    L = 0.5 * (3 + math.sqrt(2)) * 3.0
    reqs = []
    for i, (n, e, g, tf) in enumerate([(16, 0.4, 0.0, 1.0), (32, 0.8, 1.0, 1.0), (16, 0.9, 1.0, 0.5)]):
        reqs.append(serving.Request(rid=i, f0=_gauss2(n, L, i), t_final=tf, e=e, gamma=g, n=n, nr=n, m=8,
                                    s=3.0, integrator="TPI2", inner_steps=(2, 2), inner_dt=(0.01, 0.05),
                                    dt=0.25))
    a = serving.Server(use_graph=True).run(list(reqs), realtime=False)
    b = serving.Server(use_graph=False).run(list(reqs), realtime=False)
    for k in a:
        assert np.array_equal(a[k].f, b[k].f), k


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotuned_server_matches_default(pkg, need_gpu, tmp_path, monkeypatch, precision):
    """A server that tunes each class's launch and tile shape gives the default server's
    results: to FP64 round-off, and within the FP32 bar in FP32 (a tile shape changes the
    order of the FP32 angular sums)."""
    import dataclasses
    from proj_int_inelastic_amd import autotune, serving
    monkeypatch.setenv("INELASTIC_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_MEM", {})
    L = 0.5 * (3 + math.sqrt(2)) * 3.0
    reqs = [dataclasses.replace(r, precision=precision) for r in _requests(serving.Request, L)]
    a = serving.Server(autotune=True).run(list(reqs), realtime=False)
    b = serving.Server().run(list(reqs), realtime=False)
    assert len(autotune._MEM) >= 2
    for k in b:
        err = np.abs(a[k].f - b[k].f).max() / np.abs(b[k].f).max()
        assert err <= (1e-12 if precision == "fp64" else 1e-5), (k, err)


def test_realtime_arrivals_match_offline(pkg, need_gpu):
    """Requests arriving 5 ms apart on the wall clock (continuous batching admits them into
    classes that are already stepping) give the results of the offline run, where every
    request is admitted at once; latency bookkeeping is consistent."""
    import dataclasses
    from proj_int_inelastic_amd import serving
    L = 0.5 * (3 + math.sqrt(2)) * 3.0
    reqs = [dataclasses.replace(r, arrival=0.005 * i) for i, r in enumerate(_requests(serving.Request, L))]
    live = serving.Server().run(list(reqs), realtime=True)
    offline = serving.Server().run(list(reqs), realtime=False)
    assert sorted(live) == sorted(offline)
    for k in offline:
        a, b = live[k], offline[k]
        assert np.abs(a.f - b.f).max() <= 1e-12 * np.abs(b.f).max(), k
        assert a.finished >= a.admitted >= a.arrival >= 0.0
    st = serving.latency_stats(live)
    assert st["n"] == len(reqs) and 0 < st["p50_s"] <= st["p99_s"]


def _gauss_nd(d, n, L, seed):
    rng = np.random.default_rng(seed)
    dv = 2 * L / n
    v = -L + (np.arange(n) + 0.5) * dv
    g = np.meshgrid(*([v] * d), indexing="ij")
    c = rng.normal(0, 0.1, (2, d))
    f = sum(np.exp(-sum((g[a] - sg * (1.0 if a == 0 else 0.0) - c[k, a]) ** 2 for a in range(d)) / (2 * 0.5))
            for k, sg in enumerate((1, -1)))
    return f / (f.sum() * dv ** d)


S3 = 3.624654714576
# the uneven batch's shapes: (dim, N, M, s); Nr = N
UNEVEN = [(2, 16, 8, 3.0), (2, 32, 8, S3), (2, 64, 8, 3.0), (3, 16, 38, 3.0), (3, 32, 38, S3)]


def _uneven_reqs(Request, precision, t_final=0.25):
    reqs = []
    rng = np.random.default_rng(21)
    for i, (d, n, m, s) in enumerate(UNEVEN):
        L = 0.5 * (3 + math.sqrt(2)) * s
        reqs.append(Request(rid=i, f0=_gauss_nd(d, n, L, 40 + i), t_final=t_final,
                            e=float(rng.uniform(0.2, 1.0)), dim=d, n=n, nr=n, m=m, s=s,
                            gamma=float(i % 2), precision=precision, integrator="TPI2",
                            inner_steps=(2, 2), inner_dt=(0.01, 0.05), dt=0.25))
    return reqs


def test_uneven_tpi2_mix_against_oracle(pkg, need_gpu):
    """The uneven batch's five shapes (2D 16^2 / 32^2 / 64^2, 3D 16^3 / 32^3, Nr = N, per-problem
    e and gamma) advanced one TPI2 outer step together through the server: FP64 against the
    oracle's tel_proj_step2 (1e-12), FP32 within 2e-5 of FP64 (relative L-inf of f)."""
    from oracle import collision as OC
    from oracle import integrators as OI
    from oracle import quadrature as OQ
    from proj_int_inelastic_amd import serving
    r64 = serving.Server().run(_uneven_reqs(serving.Request, "fp64"), realtime=False)
    r32 = serving.Server().run(_uneven_reqs(serving.Request, "fp32"), realtime=False)
    for req in _uneven_reqs(serving.Request, "fp64"):
        d, n, m, s = req.dim, req.n, req.m, req.s
        rule = OQ.circle_rule(m) if d == 2 else OQ.lebedev_rule(m)
        op = OC.OracleOperator(d, n, s, n, *rule, e=req.e, e0=req.e, gamma=req.gamma)
        df = lambda f, h: h * op.collide(f, with_loss=True, workers=8)
        ref = OI.tel_proj_step2(req.f0.copy(), df, 0.25, [2, 2], [0.01, 0.05])
        scale = np.abs(ref).max()
        e64 = np.abs(r64[req.rid].f - ref).max() / scale
        e32 = np.abs(r32[req.rid].f - r64[req.rid].f).max() / scale
        print("uneven", (d, n), "fp64 vs oracle", e64, "fp32 vs fp64", e32, flush=True)
        assert e64 < 1e-12, ((d, n), e64)
        assert e32 < 2e-5, ((d, n), e32)


def _serving_mix(Request, precision):
    """Serving-workload shapes: 1D x 2D 32^2 jobs with Nx in {64, 128, 256}, eps in {1e-2, 1e-3},
    hard spheres, Sod-type data, PFE; homogeneous 2D 64^2 and 3D 16^3 / 32^3 TPI2 members."""
    reqs = []
    Ls = 0.5 * (3 + math.sqrt(2)) * S3
    for i, (nx, kn) in enumerate([(64, 1e-2), (128, 1e-3), (256, 1e-2), (256, 1e-3)]):
        reqs.append(Request(rid=i, f0=_riemann(nx, 32, Ls), t_final=1e-3, e=0.6 + 0.1 * i, dim=2,
                            n=32, nr=32, m=8, s=S3, gamma=1.0, kn=kn, precision=precision,
                            integrator="PFE", inner_steps=(2,), inner_dt=(6e-5,), dt=2.5e-4, xlen=1.0))
    for j, (d, n, m, s, tf) in enumerate([(2, 64, 8, 3.0, 0.5), (3, 16, 38, 3.0, 0.75), (3, 32, 38, S3, 0.25),
                                          (3, 16, 38, 3.0, 0.25)]):
        L = 0.5 * (3 + math.sqrt(2)) * s
        reqs.append(Request(rid=10 + j, f0=_gauss_nd(d, n, L, 70 + j), t_final=tf, e=0.35 + 0.15 * j,
                            dim=d, n=n, nr=n, m=m, s=s, gamma=float(j % 2), precision=precision,
                            integrator="TPI2", inner_steps=(2, 2), inner_dt=(0.01, 0.05), dt=0.25))
    return reqs


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_serving_workload_shapes_alone_equals_mix(pkg, need_gpu, precision):
    from proj_int_inelastic_amd import serving
    mix = serving.Server().run(_serving_mix(serving.Request, precision), realtime=False)
    for r in _serving_mix(serving.Request, precision):
        alone = serving.Server().run([r], realtime=False)[r.rid]
        a, b = alone.f, mix[r.rid].f
        err = np.abs(a - b).max() / np.abs(a).max()
        assert err < 1e-12, (precision, r.rid, err)
        assert abs(alone.t - r.t_final) < 1e-12


def test_server_tpi3_matches_solver(pkg, need_gpu):
    """A TPI3 class of the server equals the solver's tel_proj_step3 on the same problems."""
    import types
    from proj_int_inelastic_amd import serving
    from proj_int_inelastic_amd.boltzmann import BoltzmannSolver0D
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    from proj_int_inelastic_amd.solution import Solution
    L = 0.5 * (3 + math.sqrt(2)) * 3.0
    es = [0.4, 0.85]
    reqs = [serving.Request(rid=i, f0=_gauss2(16, L, 50 + i), t_final=1.0, e=e, n=16, nr=16, m=8, s=3.0,
                            gamma=1.0, integrator="TPI3", inner_steps=(3, 3, 3), inner_dt=(0.01, 0.05, 0.1),
                            dt=0.5) for i, e in enumerate(es)]
    res = serving.Server().run(list(reqs), realtime=False)
    cfg = {"collision-model": {"dim": 2, "e": 1.0, "e0": 1.0, "gamma": 1.0},
           "velocity-mesh": {"nv": 16, "s": 3.0, "nr": 16}, "circle-quad-rule": {"nphi": 8}}
    sm = SpectralMesh(cfg)
    coll = FSInelasticVHSCollision(cfg, sm)
    sol = Solution(types.SimpleNamespace(dim_x=0, dim_v=2, xmin=0.0, xmax=1.0, l=sm.L, nx=1, nv=16))
    sol.state.f = np.stack([r.f0 for r in reqs])
    solver = BoltzmannSolver0D(kn=1.0, collision_operator=lambda f: coll.collide(f, with_loss=True,
                                                                                 e=np.array(es)))
    solver.time_integrator = "TPI3"
    solver.inner_steps = [3, 3, 3]
    solver.inner_dt = [0.01, 0.05, 0.1]
    solver.dt = 0.5
    solver.evolve_to_time(sol, 1.0)
    fg = sol.state.get_f_global()
    for i in range(2):
        err = np.abs(res[i].f - fg[i]).max() / np.abs(fg[i]).max()
        assert err < 1e-12, (i, err)
