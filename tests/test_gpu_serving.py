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


def test_backpressure_queue(pkg, need_gpu):
    from proj_int_inelastic_amd import serving
    L = 0.5 * (3 + math.sqrt(2)) * 3.0
    reqs = _requests(serving.Request, L)[:4]
    budget = 2 * reqs[0].nbytes()
    srv = serving.Server(hbm_budget=budget)
    res = srv.run(reqs, realtime=False)
    assert sorted(res) == [0, 1, 2, 3]
    free = serving.Server().run(_requests(serving.Request, L)[:4], realtime=False)
    for k in res:
        assert np.abs(res[k].f - free[k].f).max() <= 1e-12 * np.abs(free[k].f).max()
    assert srv.used == 0


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
    """A server that tunes each class's launch shape gives the default server's results."""
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
        assert err <= 1e-12, (k, err)


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
