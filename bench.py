"""Benchmark: Q(f,f) cell-evaluations/s on the 3D 32^3 batch (one MI355X).

Workload (the 3D velocity batch): 1024 independent cells, 32^3 velocity grid on
[-8, 8]^3 (s = 3.624654714576), Nr = 32 Gauss-Legendre radial nodes, M = 38
Lebedev angles, inelastic hard spheres (gamma = 1), per-cell restitution
e ~ U[0.5, 1] (e0 = e), anisotropic Gaussian initial data (per-axis T ~
U[0.5, 2], mean velocity ~ N(0, 0.3^2)), gain + loss, followed by the FP64
conservative projection.  Synthetic data, seeded.

One step = one Q evaluation of the whole batch (inputs resident in HBM).
``value`` is the FP32 form's cell-evals/s; the plain FP64 form is timed on the
same batch and reported beside it with the speed-up; the FP32 error is
measured against the FP64 oracle on a sample of cells.  ``roofline`` is for
the dominant kernel of the FP32 form, k_fused3 (vector-ALU-bound FFT work, no
MFMA), on the work it executes: its real-data transforms cover the kx in
[0, 16] half spectrum, Nr M (5 n log2 n + 8 n) flop per cell-eval; the C2C
count (2 Nr M + Nr + 2) 5 n log2 n + (16 Nr M + 7 Nr) n is reported beside it
as ``c2c_equivalent``.  ``counters`` and ``traffic`` are read from the committed
rocprofv3 summaries of the same kernel (marked ``static``).
``cpu_baseline`` times the FP64 numpy/scipy oracle (8 FFT worker threads) on a
bounded sample of the same workload, the reference's two 2D configurations on
the oracle (``cpu_baseline.reference_configs``) and an estimate of the uneven
batch's solves/s on it (``cpu_baseline.uneven_batch``).

Side measurements in the same JSON line: the Kahan-compensated FP32 form, the
sustained FMA rate of the vector ALU (``valu_fma_peak``), the HBM-bound kernels of
the step loop at large sizes (``hbm_kernels``), three more grid shapes with the
roofline of their fused kernels (``shape_2d_32``, ``shape_3d_16``,
``shape_2d_64``), the reference's two 2D configurations through the solver API
(``reference_configs_fp64``), the uneven batch of homogeneous cooling problems
and the Poisson serving loop (p50/p99).  Each has a switch to skip it.
"""
import argparse
import json
import math
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

S3 = 3.624654714576
N, NR, M, D = 32, 32, 38, 3


def flops_per_cell(n=N ** D, nr=NR, m=M):
    """C2C algorithmic flop of one Q(f,f) cell-eval with the loss term."""
    return (2 * nr * m + nr + 2) * 5 * n * math.log2(n) + (16 * nr * m + 7 * nr) * n


def flops_per_cell_half(n=N ** D, nr=NR, m=M):
    """Flop the real-data forms (k_fused3 at 3D 32^3, k_fused2 at 2D 32^2) execute per
    cell-eval: per quadrature pair one inverse and one forward real-data transform (the
    half spectrum: 5 n log2 n together) and the half-spectrum phase, product and
    accumulate passes (8 n)."""
    return nr * m * (5 * n * math.log2(n) + 8 * n)


def make_batch(nb, L, seed=0):
    rng = np.random.default_rng(seed)
    dv = 2 * L / N
    v = -L + (np.arange(N) + 0.5) * dv
    g = np.meshgrid(v, v, v, indexing="ij")
    f = np.empty((nb, N, N, N))
    for c in range(nb):
        T = rng.uniform(0.5, 2.0, 3)
        u = rng.normal(0.0, 0.3, 3)
        x = np.exp(-sum((g[a] - u[a]) ** 2 / (2 * T[a]) for a in range(3)))
        f[c] = x / (x.sum() * dv ** 3)
    e = rng.uniform(0.5, 1.0, nb)
    return f, e


def cfg3(prec_e=1.0):
    return {"collision-model": {"dim": 3, "e": prec_e, "e0": prec_e, "gamma": 1.0},
            "velocity-mesh": {"nv": N, "s": S3, "nr": NR},
            "spherical-design-rule": {"ssrule": "lebedev", "nsphr": M}}


def _load_pkg():
    import importlib.util
    name = "proj_int_inelastic_amd"
    if name in sys.modules:
        return sys.modules[name]
    pkg = ROOT / "proj-int-inelastic_amd"
    spec = importlib.util.spec_from_file_location(name, pkg / "__init__.py",
                                                  submodule_search_locations=[str(pkg)])
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def time_path(coll, f_dev, e, steps, warmup, timed_kernel=False):
    import torch
    for _ in range(warmup):
        coll.collide(f_dev, with_loss=True, e=e, project=True)
    torch.cuda.synchronize()
    if timed_kernel:
        coll.set_timing(True)
    st = torch.cuda.current_stream()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(st)
    q = None
    for _ in range(steps):
        q = coll.collide(f_dev, with_loss=True, e=e, project=True)
    ev1.record(st)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return ev0.elapsed_time(ev1) / 1e3 / steps, wall / steps, q


def cpu_baseline(nb_sample, seed):
    """FP64 oracle (scipy.fft) on a bounded sample of cells of the same workload.

    One operator (one restitution value) serves every sampled cell: the work of
    a Q evaluation does not depend on e.  Precomputation is excluded, as the
    reference's plan building is.  One untimed warm-up cell, then every sampled
    cell is timed on its own; the rate is 1 / median per-cell time.  Timed with
    8 FFT worker threads (the reference's pyfftw NUM_THREADS = 8) and, for the
    record, with every core of the process's affinity mask.
    """
    from oracle import collision as OC
    from oracle import quadrature as OQ
    sig, ws = OQ.lebedev_rule(M)
    L = 0.5 * (3 + math.sqrt(2)) * S3
    f, e = make_batch(nb_sample + 1, L, seed)
    op = OC.OracleOperator(3, N, S3, NR, sig, ws, e=e[0], e0=e[0], gamma=1.0)
    ncpu = len(os.sched_getaffinity(0))
    out = {}
    plan = ((8, nb_sample),) if ncpu == 8 else ((8, (nb_sample + 1) // 2), (ncpu, nb_sample // 2))
    for workers, half in plan:
        op.collide(f[0], with_loss=True, workers=workers)
        times = []
        for c in range(1, half + 1):
            t0 = time.perf_counter()
            op.collide(f[c], with_loss=True, workers=workers)
            times.append(time.perf_counter() - t0)
        out[workers] = (1.0 / float(np.median(times)), len(times), float(sum(times)))
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return out, ncpu, model


def cpu_reference_configs(c1_steps=20, workers=8):
    """The reference's two 2D configurations on the FP64 oracle (host cores, 8 FFT workers), on
    the same data as reference_configs: C1 forward Euler steps of 2D 32^2 homogeneous cooling
    (gain + loss), and one projective outer step of C2 (Nx = 200, two inner steps, each the
    oracle's minmod TVD increment plus dt Q / eps with the reference's gain-only operator)."""
    from oracle import collision as OC
    from oracle import integrators as OI
    from oracle import quadrature as OQ
    from oracle import transport as OT
    sig, ws = OQ.circle_rule(8)
    out = {}
    s1 = 2.718491035932
    op = OC.OracleOperator(2, 32, s1, 32, sig, ws, e=0.8, e0=0.8, gamma=0.0)
    L = 0.5 * (3 + math.sqrt(2)) * s1
    dv = 2 * L / 32
    v = -L + (np.arange(32) + 0.5) * dv
    vx, vy = np.meshgrid(v, v, indexing="ij")
    rng = np.random.default_rng(0)
    cs = [np.array([sg, 0.0]) + rng.normal(0, 0.1, 2) for sg in (1.0, -1.0)]
    f = sum(np.exp(-((vx - c[0]) ** 2 + (vy - c[1]) ** 2) / 1.0) for c in cs)
    f /= f.sum() * dv * dv
    op.collide(f, with_loss=True, workers=workers)
    t0 = time.perf_counter()
    for _ in range(c1_steps):
        OI.euler_step(f, lambda g, h: h * op.collide(g, with_loss=True, workers=workers), 0.01)
    out["config1"] = {"steps_per_s": c1_steps / (time.perf_counter() - t0), "steps": c1_steps}
    nx = 200
    op2 = OC.OracleOperator(2, 32, S3, 32, sig, ws, e=0.9, e0=0.9, gamma=1.0)
    vx2, vy2 = np.meshgrid(op2.v_center(), op2.v_center(), indexing="ij")
    x = (np.arange(nx) + 0.5) / nx
    rho = np.where(x < 0.5, 1.0, 0.125)[:, None, None]
    T = np.where(x < 0.5, 1.0, 0.8)[:, None, None]
    f2 = rho / (2 * math.pi * T) * np.exp(-(vx2 ** 2 + vy2 ** 2) / (2 * T))
    dx = 1.0 / nx

    def df(g, h):
        return OT.df_advection(g, vx2, h, dx) + h * op2(g, workers=workers) / 1e-3

    t0 = time.perf_counter()
    OI.proj_euler_step(f2, df, 0.4 * dx / 8, [2], [1e-3])
    out["config2"] = {"outer_steps_per_s": 1.0 / (time.perf_counter() - t0), "outer_steps": 1}
    out["workers"] = workers
    return out


def cpu_uneven_estimate(q_sec_3d32, workers=8):
    """CPU solves/s of the uneven batch's mix, estimated from the FP64 oracle: one timed Q
    evaluation per 2D 16^2 / 32^2 / 64^2 and 3D 16^3 shape (after one warm-up; the 3D 32^3 time
    is the cpu_baseline sample's), times the 32 Q evaluations of a TPI2 solve to t = 2 (8 outer
    steps of 2 x 2 inner steps); the shapes are equally likely, so the mix's rate is
    5 / (sum of the per-shape solve times)."""
    from oracle import collision as OC
    from oracle import quadrature as OQ
    per = {}
    for d, n, m in [(2, 16, 8), (2, 32, 8), (2, 64, 8), (3, 16, 38)]:
        sig, ws = OQ.circle_rule(m) if d == 2 else OQ.lebedev_rule(m)
        s = 3.0 if n < 32 else S3
        op = OC.OracleOperator(d, n, s, n, sig, ws, e=0.6, e0=0.6, gamma=1.0)
        g = np.meshgrid(*([op.v_center()] * d), indexing="ij")
        f = np.exp(-sum(x ** 2 for x in g) / 2.0)
        op.collide(f, with_loss=True, workers=workers)
        t0 = time.perf_counter()
        op.collide(f, with_loss=True, workers=workers)
        per["%dd%d" % (d, n)] = time.perf_counter() - t0
    per["3d32"] = q_sec_3d32
    solve = {k: 32 * v for k, v in per.items()}
    return {"solves_per_s_estimate": len(solve) / sum(solve.values()), "q_seconds": per,
            "q_per_solve": 32, "workers": workers}


def counters_from_profile():
    """k_fused3 counter figures (VALU busy, LDS bandwidth, bank conflicts, occupancy) from the
    committed SQ summary, or None."""
    p = ROOT / "profiles" / "fused3_sq_summary.txt"
    if not p.exists():
        return None
    out = {}
    for line in p.read_text().splitlines():
        parts = line.split()
        if len(parts) == 2 and not line.startswith("#"):
            out[parts[0]] = float(parts[1])
    return out or None


def traffic_from_profile():
    """HBM bytes per fused launch from the committed PMC summary, or None."""
    p = ROOT / "profiles" / "fused3_pmc_traffic.txt"
    if not p.exists():
        return None
    for line in p.read_text().splitlines():
        if line.startswith("traffic_bytes_per_cell"):
            return float(line.split()[1])
    return None


def serve_sweep_requests(precision, nreq, rate, seed, full_time=False):
    """Requests of the swept-rate serving run, shaped like the serving workload: Poisson
    arrivals mixing homogeneous cooling solves distributed as the uneven batch (2D 16^2 /
    32^2 / 64^2, 3D 16^3 / 32^3, TPI2, t_final in {0.5, 0.75, .., 3}) with 1D x 2D jobs
    (Nx in {64, 128, 256}, 32^2 velocity, eps in {1e-2, 1e-3}, Sod-type Maxwellian data,
    hard spheres, e ~ U[0.5, 1]).  The 1D jobs use a projective forward Euler schedule
    that is stable at every Nx (2 inner steps of 6e-5, outer 2.5e-4, inner CFL <= 0.12)
    and final times 2.5e-4 x U{200..1200}, i.e. 0.05 .. 0.3 (full_time: 2.5e-4 x U{2000..12000},
    the workload's U[0.5, 3])."""
    from proj_int_inelastic_amd import serving
    rng = np.random.default_rng(seed)
    arrivals = serving.poisson_arrivals(nreq, rate, rng)
    shapes = [(2, 16, 8), (2, 32, 8), (2, 64, 8), (3, 16, 38), (3, 32, 38)]
    reqs = []
    for i, t_arr in enumerate(arrivals):
        if rng.random() < 0.75:
            d, n, m = shapes[int(rng.integers(0, len(shapes)))]
            s = 3.0 if n < 32 else S3
            L = 0.5 * (3 + math.sqrt(2)) * s
            dv = 2 * L / n
            v = -L + (np.arange(n) + 0.5) * dv
            g = np.meshgrid(*([v] * d), indexing="ij")
            c = rng.normal(0, 0.1, (2, d))
            f0 = sum(np.exp(-sum((g[a] - sg * (1.0 if a == 0 else 0.0) - c[k, a]) ** 2
                                 for a in range(d)) / (2 * 0.5)) for k, sg in enumerate((1, -1)))
            f0 /= f0.sum() * dv ** d
            reqs.append(serving.Request(rid=i, f0=f0, t_final=0.25 * int(rng.integers(2, 13)),
                                        e=float(rng.uniform(0.2, 1.0)), dim=d, n=n, nr=n, m=m, s=s,
                                        gamma=float(rng.integers(0, 2)), precision=precision,
                                        integrator="TPI2", inner_steps=(2, 2), inner_dt=(0.01, 0.05),
                                        dt=0.25, arrival=float(t_arr)))
        else:
            nx, n, s = int(rng.choice([64, 128, 256])), 32, S3
            L = 0.5 * (3 + math.sqrt(2)) * s
            v = -L + (np.arange(n) + 0.5) * 2 * L / n
            vx, vy = np.meshgrid(v, v, indexing="ij")
            x = (np.arange(nx) + 0.5) / nx
            rho = np.where(x < 0.5, 1.0, 0.125)[:, None, None]
            T = np.where(x < 0.5, 1.0, 0.8)[:, None, None]
            f0 = rho / (2 * math.pi * T) * np.exp(-(vx ** 2 + vy ** 2) / (2 * T))
            steps = int(rng.integers(2000, 12001)) if full_time else int(rng.integers(200, 1201))
            reqs.append(serving.Request(rid=i, f0=f0, t_final=2.5e-4 * steps,
                                        e=float(rng.uniform(0.5, 1.0)), dim=2, n=n, nr=n, m=8, s=s,
                                        gamma=1.0, kn=float(rng.choice([1e-2, 1e-3])),
                                        precision=precision, integrator="PFE", inner_steps=(2,),
                                        inner_dt=(6e-5,), dt=2.5e-4, xlen=1.0, arrival=float(t_arr)))
    return reqs


def serve_sweep(rates, nreq, with_fp64, full_time=False):
    from proj_int_inelastic_amd import serving
    out = []
    for k, rate in enumerate(rates):
        row = {"rate_per_s": rate, "requests": nreq}
        for prec in (("fp32", "fp64") if with_fp64 else ("fp32",)):
            srv = serving.Server(hbm_budget=32 << 30)
            res = srv.run(serve_sweep_requests(prec, nreq, rate, seed=100 + k, full_time=full_time),
                          realtime=True)
            row[prec] = serving.latency_stats(res)
            print("serve sweep rate %g %s %s" % (rate, prec, row[prec]), file=sys.stderr, flush=True)
        out.append(row)
    return out


def valu_peak():
    """Sustained vector-ALU FMA rate (TFLOP/s) in FP32 and FP64, from the probe library's
    FMA-loop kernel built with the collision library's flags (no packed FP32)."""
    import ctypes
    lib = ctypes.CDLL(str(ROOT / "proj-int-inelastic_amd" / "libinelastic_probe.so"))
    fn = lib.ibp_fma_peak
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_int, ctypes.c_int,
                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    out = {}
    for dt, name in ((0, "fp32"), (1, "fp64")):
        fl, ms = ctypes.c_double(), ctypes.c_double()
        rc = fn(dt, 256 * 64, 512, 5, ctypes.byref(fl), ctypes.byref(ms))
        if rc != 0:
            raise RuntimeError("ibp_fma_peak failed (%d)" % rc)
        out[name + "_tflops"] = fl.value / (ms.value * 1e-3) / 1e12
    return out


def hbm_kernels():
    """Bandwidth of the HBM-bound kernels of the step loop at large sizes, in algorithmic
    bytes (FP64 reads + writes each kernel must make) per second, beside a device copy."""
    import torch
    lib = _load_pkg()._lib.load()
    st = torch.cuda.current_stream()

    def timed(fn, reps=5):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            fn()
        e1.record(st)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / reps

    h = st.cuda_stream
    npt = 1 << 27  # 1 GiB of FP64 per array
    a = torch.rand(npt, dtype=torch.float64, device="cuda")
    b = torch.rand(npt, dtype=torch.float64, device="cuda")
    c = torch.empty_like(a)
    res = {}
    sec = timed(lambda: c.copy_(a))
    copy_gbs = 16 * npt / sec / 1e9
    res["device_copy_GBps"] = copy_gbs

    def rec(name, nbytes, sec, what):
        gbs = nbytes / sec / 1e9
        res[name] = {"GBps": gbs, "frac_of_copy": gbs / copy_gbs, "frac_of_8TBps": gbs / 8000.0,
                     "ms": sec * 1e3, "bytes_alg": nbytes, "what": what}

    def chk(rc):
        if rc != 0:
            raise RuntimeError(_load_pkg()._lib.last_error())

    sec = timed(lambda: chk(lib.ib_axpy(c.data_ptr(), a.data_ptr(), 0.5, npt, h)))
    rec("ib_axpy", 24 * npt, sec, "y += a x, 2^27 points (24 B/point)")
    sec = timed(lambda: chk(lib.ib_df_combine(a.data_ptr(), b.data_ptr(), c.data_ptr(), 1e-3, 1e-2, npt, h)))
    rec("ib_df_combine", 24 * npt, sec, "adv + dt q / kn, 2^27 points (24 B/point)")
    nx, nv = npt // 1024, 1024
    v1 = torch.linspace(-8, 8, 32, dtype=torch.float64, device="cuda").repeat_interleave(32)
    sec = timed(lambda: chk(lib.ib_tvd1d(a.data_ptr(), c.data_ptr(), nx, nv, v1.data_ptr(), 0.01, 2, 0,
                                         1, 1, 32, h)))
    rec("ib_tvd1d", 16 * npt, sec, "2nd-order minmod TVD step, Nx = %d cells x 32^2 (16 B/point)" % nx)
    nb = npt // 32768
    mom = torch.empty((nb, 5), dtype=torch.float64, device="cuda")
    sec = timed(lambda: chk(lib.ib_moments(a.data_ptr(), mom.data_ptr(), nb, 3, 32, 8.0, h)))
    rec("ib_moments", 8 * npt, sec, "mass/momentum/energy of %d cells of 32^3 (8 B/point)" % nb)
    sec = timed(lambda: chk(lib.ib_project_conservative(c.data_ptr(), nb, 3, 32, 8.0, h)))
    rec("ib_project_conservative", 16 * npt, sec, "in-place projection of %d cells of 32^3 (16 B/point)" % nb)
    del a, b, c
    return res


def reference_configs(c2_steps=20):
    """The reference's own two FP64 configurations through the solver step API, timed on the
    device (eager and with each outer step replayed from a captured HIP graph):
    C1 space-homogeneous 2D inelastic Maxwell molecules, 32^2 on [-6,6]^2, Nr = 32, M = 8,
    e = 0.8, seeded two-Gaussian data (T = 0.5), forward Euler dt = 0.01, 100 steps (with
    gain + loss, the conservative form of Q); C2 Nx = 200 x 32^2 on [-8,8]^2, hard spheres,
    e = 0.9, eps = 1e-3, Sod-type Maxwellian data, minmod TVD, PFE with K = 2 inner steps of
    eps and outer dt = 0.4 dx / 8 (the reference's gain-only __call__), c2_steps outer steps."""
    import types
    import torch
    from proj_int_inelastic_amd.boltzmann import BoltzmannSolver0D, BoltzmannSolver1D
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    from proj_int_inelastic_amd.riemann.advection_1D import advection_1D
    from proj_int_inelastic_amd.solution import Solution
    from proj_int_inelastic_amd.solver import BC
    out = {}
    # C1
    s1 = 2.718491035932
    cfg = {"collision-model": {"dim": 2, "e": 0.8, "e0": 0.8, "gamma": 0.0},
           "velocity-mesh": {"nv": 32, "s": s1, "nr": 32}, "circle-quad-rule": {"nphi": 8}}
    sm = SpectralMesh(cfg)
    dv = 2 * sm.L / 32
    v = -sm.L + (np.arange(32) + 0.5) * dv
    vx, vy = np.meshgrid(v, v, indexing="ij")
    rng = np.random.default_rng(0)
    cs = [np.array([sg, 0.0]) + rng.normal(0, 0.1, 2) for sg in (1.0, -1.0)]
    f0 = sum(np.exp(-((vx - c[0]) ** 2 + (vy - c[1]) ** 2) / 1.0) for c in cs)
    f0 /= f0.sum() * dv * dv
    c1 = {"workload": "C1: 2D 32^2 homogeneous cooling, Maxwell molecules, e=0.8, Euler dt=0.01, 100 steps"}
    for graph in (False, True):
        coll = FSInelasticVHSCollision(cfg, sm, with_loss=True)
        sol = Solution(types.SimpleNamespace(dim_x=0, dim_v=2, xmin=0.0, xmax=1.0, l=sm.L, nx=1, nv=32))
        sol.state.f = f0.copy()
        slv = BoltzmannSolver0D(kn=1.0, collision_operator=coll)
        slv.dt = 0.01
        slv.use_graph = graph
        slv.evolve_to_time(sol, 0.05)  # warm-up: plans, scratch, graph capture
        sol.state.f.copy_(torch.from_numpy(f0))  # in place: the captured graph stays valid
        sol.t = 0.0
        T0 = float(sol.state.T)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        slv.evolve_to_time(sol, 1.0)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        c1[("graph" if graph else "eager") + "_steps_per_s"] = 100 / sec
        c1["T1_over_T0"] = float(sol.state.T) / T0
    out["config1"] = c1
    # C2
    nx, n, s2 = 200, 32, S3
    cfg = {"collision-model": {"dim": 2, "e": 0.9, "e0": 0.9, "gamma": 1.0},
           "velocity-mesh": {"nv": n, "s": s2, "nr": 32}, "circle-quad-rule": {"nphi": 8}}
    sm = SpectralMesh(cfg)
    v = sm.v_center
    vx, vy = np.meshgrid(v, v, indexing="ij")
    x = (np.arange(nx) + 0.5) / nx
    rho = np.where(x < 0.5, 1.0, 0.125)[:, None, None]
    T = np.where(x < 0.5, 1.0, 0.8)[:, None, None]
    f0 = rho / (2 * math.pi * T) * np.exp(-(vx ** 2 + vy ** 2) / (2 * T))
    dt = 0.4 * (1.0 / nx) / 8
    c2 = {"workload": "C2: Nx=200 x 2D 32^2 Sod-type Riemann data, hard spheres, e=0.9, eps=1e-3, "
                      "minmod TVD, PFE K=2 inner 1e-3, outer 0.4 dx/8, %d outer steps" % c2_steps}
    for graph in (False, True):
        coll = FSInelasticVHSCollision(cfg, sm)
        sol = Solution(types.SimpleNamespace(dim_x=1, dim_v=2, xmin=0.0, xmax=1.0, l=sm.L, nx=nx, nv=n))
        slv = BoltzmannSolver1D(kn=1e-3, riemann_solver=advection_1D, collision_operator=coll)
        slv.bc_lower[0] = BC.extrap
        slv.bc_upper[0] = BC.extrap
        slv.time_integrator = "PFE"
        slv.inner_steps = [2]
        slv.inner_dt = [1e-3]
        slv.dt = dt
        slv.use_graph = graph
        sol.state.f = f0.copy()
        slv.evolve_to_time(sol, 2 * dt)  # warm-up
        sol.state.f.copy_(torch.from_numpy(f0))
        sol.t = 0.0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        slv.evolve_to_time(sol, c2_steps * dt)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        c2[("graph" if graph else "eager") + "_steps_per_s"] = c2_steps / sec
        c2[("graph" if graph else "eager") + "_cell_evals_per_s"] = c2_steps * 2 * nx / sec
    out["config2"] = c2
    return out


def uneven_requests(nprob, precision, seed=9):
    """The uneven batch of homogeneous cooling problems.

    2D N in {16, 32, 64}^2 (M = 8) and 3D N in {16, 32}^3 (M = 38 Lebedev),
    Nr = N, Maxwell or hard-sphere kernel at random, e ~ U[0.2, 1], seeded
    two-Gaussian data, TPI2 with 2 levels of K = 2 inner steps
    (dt0 = 0.01, dt1 = 0.05, outer dt = 0.25) to t = 2.
    """
    from proj_int_inelastic_amd import serving
    rng = np.random.default_rng(seed)
    shapes = [(2, 16, 8), (2, 32, 8), (2, 64, 8), (3, 16, 38), (3, 32, 38)]
    reqs = []
    for i in range(nprob):
        d, n, m = shapes[int(rng.integers(0, len(shapes)))]
        s = 3.0 if n < 32 else S3
        L = 0.5 * (3 + math.sqrt(2)) * s
        dv = 2 * L / n
        v = -L + (np.arange(n) + 0.5) * dv
        g = np.meshgrid(*([v] * d), indexing="ij")
        c = rng.normal(0, 0.1, (2, d))
        f0 = sum(np.exp(-sum((g[a] - sg * (1.0 if a == 0 else 0.0) - c[k, a]) ** 2
                             for a in range(d)) / (2 * 0.5)) for k, sg in enumerate((1, -1)))
        f0 /= f0.sum() * dv ** d
        reqs.append(serving.Request(rid=i, f0=f0, t_final=2.0, e=float(rng.uniform(0.2, 1.0)), dim=d,
                                    n=n, nr=n, m=m, s=s, gamma=float(rng.integers(0, 2)),
                                    precision=precision, integrator="TPI2", inner_steps=(2, 2),
                                    inner_dt=(0.01, 0.05), dt=0.25, arrival=0.0))
    return reqs


def uneven_run(precision, nprob, autotune=False):
    """The uneven batch through the server.  With autotune, an untimed pass of the
    same problems to one outer step picks every class's launch shape first (the
    autotuner keeps it for the process); the timed pass then starts from a new
    server (plans, scratch and graphs are built inside the timed region)."""
    import dataclasses
    import torch
    from proj_int_inelastic_amd import serving
    reqs = uneven_requests(nprob, precision)
    out = {"autotune": bool(autotune)}
    if autotune:
        t0 = time.perf_counter()
        short = [dataclasses.replace(r, t_final=r.dt) for r in reqs]
        serving.Server(hbm_budget=200 << 30, autotune=True).run(short, realtime=False)
        torch.cuda.synchronize()
        out["tune_pass_s"] = time.perf_counter() - t0
    srv = serving.Server(hbm_budget=200 << 30, autotune=autotune)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = srv.run(reqs, realtime=False)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out.update({"solves": len(res), "seconds": dt, "solves_per_s": len(res) / dt})
    return out, res


SIDE_SHAPES = {
    # name: (dim, N, Nr, M, s, fused kernel, what); k_fused2 at 32^2 is a real-data form
    "2d32": (2, 32, 32, 8, S3, "k_fused2",
             "2D 32^2, Nr=32, M=8 (the homogeneous-cooling and 1D x 2D shape)"),
    "2d64": (2, 64, 64, 8, 3.0, "k_fusedw",
             "2D 64^2, Nr=64, M=8, s=3 (the 2D 64^2 shape of the uneven batch)"),
    "3d16": (3, 16, 16, 38, 3.0, "k_fusedc3",
             "3D 16^3, Nr=16, M=38 Lebedev, s=3 (the 3D shape of the uneven batch and serving mix)"),
}


def run_shape(name, nb, steps, warmup, precs=("fp32", "fp64")):
    """Side measurement on one more shape: nb cells of seeded anisotropic Gaussians
    (per-axis T ~ U[0.5, 2]), hard spheres, e = 0.9, gain + loss, in the listed precisions;
    roofline of the fused FP32 kernel from its live HIP-event time, on the work it executes:
    the C2C count for the 3D 16^3 and 2D 64^2 kernels (they transform the full complex
    spectrum), the half-spectrum count for the real-data 2D 32^2 kernel (its C2C-equivalent
    rate reported beside it)."""
    import torch
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    d, n, nr, m, s, kname, what = SIDE_SHAPES[name]
    cfg = {"collision-model": {"dim": d, "e": 0.9, "e0": 0.9, "gamma": 1.0},
           "velocity-mesh": {"nv": n, "s": s, "nr": nr}}
    if d == 2:
        cfg["circle-quad-rule"] = {"nphi": m}
    else:
        cfg["spherical-design-rule"] = {"ssrule": "lebedev", "nsphr": m}
    sm = SpectralMesh(cfg)
    rng = np.random.default_rng(5)
    dv = 2 * sm.L / n
    v = -sm.L + (np.arange(n) + 0.5) * dv
    grids = np.meshgrid(*([v] * d), indexing="ij")
    T = rng.uniform(0.5, 2.0, (nb, d))
    f = np.exp(-sum(grids[a][None] ** 2 / (2 * T[:, a].reshape((nb,) + (1,) * d)) for a in range(d)))
    f /= f.sum(axis=tuple(range(1, d + 1)), keepdims=True) * dv ** d
    f_dev = torch.from_numpy(f).cuda()
    res = {"workload": "%s, hard spheres, e=0.9, gain+loss, %d cells" % (what, nb)}
    flop = flops_per_cell(n ** d, nr, m)
    real = name == "2d32"
    flop_exec = flops_per_cell_half(n ** d, nr, m) if real else flop
    out = {}
    for prec in precs:
        c = FSInelasticVHSCollision(cfg, sm, precision=prec)
        for _ in range(max(warmup, 1)):
            c.collide(f_dev, with_loss=True)
        torch.cuda.synchronize()
        if prec == "fp32":
            c.set_timing(True)
        st = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(steps):
            out[prec] = c.collide(f_dev, with_loss=True)
        e1.record(st)
        torch.cuda.synchronize()
        sec = e0.elapsed_time(e1) / 1e3 / steps
        res[prec + "_cell_evals_per_s"] = nb / sec
        if prec == "fp32":
            kms, kl, ku = c.get_timing()
            kavg = kms / max(kl, 1)
            ach = flop_exec * (ku / max(kl, 1)) / (kavg * 1e-3) / 1e12
            rl = {"bound": "valu", "achieved": ach, "peak": 157.3, "unit": "TFLOP/s",
                  "frac": ach / 157.3, "kernel_avg_ms": kavg, "cells_per_launch": ku / max(kl, 1),
                  "flop_per_cell_executed": flop_exec}
            if real:
                c2c = flop * (ku / max(kl, 1)) / (kavg * 1e-3) / 1e12
                rl["c2c_equivalent"] = {"flop_per_cell": flop, "achieved": c2c, "frac": c2c / 157.3,
                                        "note": "the C2C algorithmic count charged to the same time"}
            res["roofline_" + kname] = rl
    if "fp64" in out and "fp32" in out:
        a, b = out["fp32"], out["fp64"]
        axes = tuple(range(1, d + 1))
        res["fp32_over_fp64"] = res["fp32_cell_evals_per_s"] / res["fp64_cell_evals_per_s"]
        res["fp32_vs_fp64_rel_linf_max"] = float(((a - b).abs().amax(dim=axes) /
                                                  b.abs().amax(dim=axes)).max())
    return res


_T0 = time.perf_counter()


def _progress(msg):
    """One line on stderr per finished stage."""
    print("bench: %s done (%.0f s)" % (msg, time.perf_counter() - _T0), file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=1024)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--check-cells", type=int, default=2)
    ap.add_argument("--cpu-cells", type=int, default=36)
    ap.add_argument("--mode", choices=["bench", "profile", "profile2d", "profile2d32fp64", "profile2d64",
                                       "profile3d16", "profile3d32fp64", "uneven", "copy"],
                    default="bench",
                    help="profile: FP32 form only on --cells cells (for rocprofv3); "
                         "profile2d: the 2D 32^2 FP32 form only on --cells-2d cells; "
                         "profile2d32fp64: the 2D 32^2 FP64 form only on --cells-2d cells; "
                         "profile2d64: the 2D 64^2 FP32 form only on --cells-2d64 cells; "
                         "profile3d16: the 3D 16^3 FP32 form only on --cells-3d16 cells; "
                         "profile3d32fp64: the plain FP64 form of the 3D 32^3 batch only on --cells cells; "
                         "uneven: the FP32 uneven batch only (--uneven-problems, default shapes, "
                         "no graph replay; --per-plan-fused for one fused launch per plan); "
                         "copy: a plain 1 GiB device copy (PMC byte-counter calibration)")
    ap.add_argument("--serve-autotune", action="store_true",
                    help="serving: classes take their autotuned launch and tile shapes")
    ap.add_argument("--per-plan-fused", action="store_true",
                    help="uneven mode: one fused launch per plan instead of the grouped launch")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-kahan", action="store_true", help="skip the fp32k side measurement")
    ap.add_argument("--serve-requests", type=int, default=32,
                    help="requests of the Poisson-arrival serving run (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=10.0, help="arrival rate (1/s)")
    ap.add_argument("--serve-seed", type=int, default=100, help="seed of the serving request mix")
    ap.add_argument("--serve-sweep", default="",
                    help="comma-separated arrival rates: run only the swept-rate serving workload")
    ap.add_argument("--serve-sweep-requests", type=int, default=32)
    ap.add_argument("--serve-full-time", action="store_true",
                    help="1D jobs of the sweep run to t_final in U[0.5, 3] instead of 0.05 .. 0.3")
    ap.add_argument("--cells-2d", type=int, default=4096,
                    help="cells of the 2D 32^2 side measurement (0 = skip)")
    ap.add_argument("--cells-3d16", type=int, default=2048,
                    help="cells of the 3D 16^3 side measurement (0 = skip)")
    ap.add_argument("--cells-2d64", type=int, default=1024,
                    help="cells of the 2D 64^2 side measurement (0 = skip)")
    ap.add_argument("--uneven-problems", type=int, default=4096,
                    help="problems of the uneven homogeneous batch (0 = skip)")
    ap.add_argument("--no-hbm", action="store_true", help="skip the HBM-bound kernel measurements")
    ap.add_argument("--no-ref-configs", action="store_true",
                    help="skip the step-loop timing of the reference's 2D configurations")
    ap.add_argument("--no-uneven-autotune", action="store_true",
                    help="run the FP32 uneven batch with the default launch shapes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the Q(f,f) batch is sharded by cells with no "
              "exchange, so a multi-GPU figure would be N x the one-GPU value; run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    import torch
    if args.mode == "copy":
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda").fill_(1.0)
        torch.cuda.synchronize()
        b = a.clone()
        torch.cuda.synchronize()
        print(json.dumps({"mode": "copy", "bytes_read": a.numel() * 8, "bytes_written": b.numel() * 8}))
        return
    pkg = _load_pkg()
    from proj_int_inelastic_amd.collision import FSInelasticVHSCollision, SpectralMesh
    if args.serve_sweep:
        rates = [float(x) for x in args.serve_sweep.split(",") if x]
        print(json.dumps({"mode": "serve_sweep", "full_time": args.serve_full_time,
                          "sweep": serve_sweep(rates, args.serve_sweep_requests, not args.no_fp64,
                                               args.serve_full_time)}))
        return
    if args.mode == "profile2d":
        print(json.dumps({"mode": "profile2d", **run_shape("2d32", args.cells_2d, args.steps, args.warmup,
                                                           ("fp32",))}))
        return
    if args.mode == "profile2d32fp64":
        print(json.dumps({"mode": "profile2d32fp64", **run_shape("2d32", args.cells_2d, args.steps,
                                                                 args.warmup, ("fp64",))}))
        return
    if args.mode == "profile2d64":
        print(json.dumps({"mode": "profile2d64", **run_shape("2d64", args.cells_2d64, args.steps,
                                                             args.warmup, ("fp32",))}))
        return
    if args.mode == "uneven":
        from proj_int_inelastic_amd import serving
        from proj_int_inelastic_amd.collision.inelastic_collisions import set_group_mode
        set_group_mode(not args.per_plan_fused)
        reqs = uneven_requests(args.uneven_problems, "fp32")
        srv = serving.Server(hbm_budget=200 << 30, use_graph=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = srv.run(reqs, realtime=False)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        print(json.dumps({"mode": "uneven", "grouped_fused": not args.per_plan_fused, "solves": len(res),
                          "seconds": dt, "solves_per_s": len(res) / dt}))
        return
    if args.mode == "profile3d16":
        print(json.dumps({"mode": "profile3d16", **run_shape("3d16", args.cells_3d16, args.steps,
                                                             args.warmup, ("fp32",))}))
        return

    cfg = cfg3()
    sm = SpectralMesh(cfg)
    f, e = make_batch(args.cells, sm.L, seed=0)
    f_dev = torch.from_numpy(f).cuda()
    c32 = FSInelasticVHSCollision(cfg, sm, precision="fp32")
    if args.mode in ("profile", "profile3d32fp64"):
        cp = c32 if args.mode == "profile" else FSInelasticVHSCollision(cfg, sm, precision="fp64")
        for _ in range(args.warmup + args.steps):
            cp.collide(f_dev, with_loss=True, e=e, project=True)
        torch.cuda.synchronize()
        print(json.dumps({"mode": args.mode, "cells": args.cells, "launch_steps": args.warmup + args.steps}))
        return
    t32, w32, q32 = time_path(c32, f_dev, e, args.steps, args.warmup, timed_kernel=True)
    kms, klaunch, kcells = c32.get_timing()
    c32.set_timing(False)
    res = {"fp32_s_per_step": t32}
    _progress("FP32 3D 32^3 timed steps")
    # PCIe-inclusive rate: numpy in, numpy out (host copies inside the timed region)
    nh = min(256, args.cells)
    fh = np.ascontiguousarray(f[:nh])
    c32.collide(fh, with_loss=True, e=e[:nh], project=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    c32.collide(fh, with_loss=True, e=e[:nh], project=True)
    torch.cuda.synchronize()
    res["pcie_inclusive_cell_evals_per_s"] = nh / (time.perf_counter() - t0)
    if not args.no_fp64:
        c64 = FSInelasticVHSCollision(cfg, sm, precision="fp64")
        t64, w64, q64 = time_path(c64, f_dev, e, args.fp64_steps, 1 if args.warmup else 0)
        res["fp64_s_per_step"] = t64
        d = (q32 - q64).abs().amax(dim=(1, 2, 3)) / q64.abs().amax(dim=(1, 2, 3))
        res["fp32_vs_fp64_gpu_rel_linf_max"] = float(d.max())
        _progress("FP64 3D 32^3")
    if not args.no_kahan:
        ck = FSInelasticVHSCollision(cfg, sm, precision="fp32k")
        tk, wk, qk = time_path(ck, f_dev, e, 1, 1 if args.warmup else 0)
        res["fp32k"] = {"cell_evals_per_s": args.cells / tk}
        if not args.no_fp64:
            dk = (qk - q64).abs().amax(dim=(1, 2, 3)) / q64.abs().amax(dim=(1, 2, 3))
            res["fp32k"]["vs_fp64_gpu_rel_linf_max"] = float(dk.max())
        del qk
        _progress("FP32 Kahan 3D 32^3")
    # error of the FP32 form against the FP64 oracle on a sample of cells
    from oracle import collision as OC
    from oracle import quadrature as OQ
    sig, ws = OQ.lebedev_rule(M)
    errs = []
    for c in range(min(args.check_cells, args.cells)):
        op = OC.OracleOperator(3, N, S3, NR, sig, ws, e=e[c], e0=e[c], gamma=1.0)
        ref = OC.project_conservative(op.collide(f[c], with_loss=True, workers=8), op.v_grids())
        qc = q32[c].cpu().numpy()
        errs.append(float(np.abs(qc - ref).max() / np.abs(ref).max()))
    # conservation of the projected FP32 result (FP64 moments)
    mom = c32.moments(q32)
    fm = c32.moments(f_dev)
    mass_drift = float((mom[:, 0].abs() / fm[:, 0].abs()).max())
    _progress("oracle check of %d cells" % len(errs))
    cells_s = args.cells / t32
    flop = flops_per_cell()
    # dominant kernel (the fused FP32 kernel), timed live with HIP events on its stream
    kernel_avg_ms = kms / max(klaunch, 1)
    cells_per_launch = kcells / max(klaunch, 1)
    flop_exec = flops_per_cell_half()
    achieved = flop_exec * cells_per_launch / (kernel_avg_ms * 1e-3) / 1e12
    achieved_c2c = flop * cells_per_launch / (kernel_avg_ms * 1e-3) / 1e12
    peak = 157.3
    tr = traffic_from_profile()
    out = {
        "metric": "Q(f,f) cell-evals/s at N=32^3 (FP32 fast form, rel Linf vs FP64 oracle <= 1e-5)",
        "value": cells_s,
        "unit": "cell-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (FFTs, angular quadrature sums) + fp64 (radial sum, final transform, loss, projection)",
        "data": "synthetic (seeded anisotropic Gaussians, e ~ U[0.5,1])",
        "config": {"workload": "3D velocity batch: %d cells x 32^3, Nr=32, M=38 Lebedev, "
                               "hard spheres, per-cell e, gain+loss+projection" % args.cells,
                   "cells": args.cells, "n": N, "nr": NR, "m": M},
        "fp32_rel_linf_vs_oracle_max": max(errs) if errs else None,
        "fp32_mass_rel_drift_max": mass_drift,
        # compute-bound on the vector ALU (FFTs are no dense contraction: no MFMA instruction);
        # peak = the published MI355X FP32 figure, which counts packed dual-issue FP32 (kept off
        # in this build); frac_of_attainable compares with the measured non-packed FMA rate
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     "traffic": tr * cells_per_launch if tr is not None else None,
                     "traffic_source": "static: profiles/fused3_pmc_traffic.txt (rocprofv3 --pmc "
                                       "FETCH_SIZE, WRITE_SIZE of k_fused3, copy-calibrated)",
                     "kernel": "k_fused3", "kernel_avg_ms": kernel_avg_ms,
                     "cells_per_launch": cells_per_launch,
                     "flop_per_cell_executed": flop_exec,
                     "c2c_equivalent": {"flop_per_cell": flop, "achieved": achieved_c2c,
                                        "frac": achieved_c2c / peak,
                                        "note": "the C2C algorithmic count charged to the same time"},
                     "kernel_share_of_step": kms / (t32 * 1e3 * args.steps)},
    }
    out["pcie_inclusive_cell_evals_per_s"] = res["pcie_inclusive_cell_evals_per_s"]
    if "fp32k" in res:
        out["fp32_kahan"] = dict(res["fp32k"], note="Kahan-compensated angular sums (precision fp32k)")
    if "fp64_s_per_step" in res:
        out["fp64_plain_cell_evals_per_s"] = args.cells / res["fp64_s_per_step"]
        out["fp32_over_fp64"] = res["fp64_s_per_step"] / t32
        out["fp32_vs_fp64_gpu_rel_linf_max"] = res["fp32_vs_fp64_gpu_rel_linf_max"]
    ctr = counters_from_profile()
    out["roofline"]["counters"] = dict(ctr, source="static: profiles/fused3_sq_summary.txt "
                                                   "(rocprofv3 --pmc SQ counters of k_fused3)") if ctr else None
    pk = valu_peak()
    out["roofline"]["measured_valu_fma_peak_tflops"] = pk["fp32_tflops"]
    out["roofline"]["frac_of_attainable"] = achieved / pk["fp32_tflops"]
    out["valu_fma_peak"] = dict(pk, note="FMA-loop probe, no packed FP32 (the library's build flags)")
    if not args.no_hbm:
        out["hbm_kernels"] = hbm_kernels()
        _progress("HBM-bound kernels")
    precs = ("fp32",) if args.no_fp64 else ("fp32", "fp64")
    if args.cells_2d > 0:
        out["shape_2d_32"] = run_shape("2d32", args.cells_2d, args.steps, args.warmup, precs)
        _progress("2D 32^2 shape")
    if args.cells_3d16 > 0:
        out["shape_3d_16"] = run_shape("3d16", args.cells_3d16, args.steps, args.warmup, precs)
        _progress("3D 16^3 shape")
    if args.cells_2d64 > 0:
        out["shape_2d_64"] = run_shape("2d64", args.cells_2d64, args.steps, args.warmup, precs)
        _progress("2D 64^2 shape")
    if not args.no_ref_configs:
        out["reference_configs_fp64"] = reference_configs()
        _progress("reference configurations")
    if args.uneven_problems > 0:
        u32, r32 = uneven_run("fp32", args.uneven_problems, autotune=not args.no_uneven_autotune)
        _progress("uneven batch FP32")
        out["uneven_batch"] = {"workload": "homogeneous cooling, 2D 16^2/32^2/64^2 + 3D 16^3/32^3, "
                                           "TPI2 to t=2, grouped launches (one launch of the fused "
                                           "pair loops of all FP32 shapes per df stage)", "fp32": u32}
        from proj_int_inelastic_amd.collision.inelastic_collisions import set_group_mode
        set_group_mode(False)
        try:
            u32p, _ = uneven_run("fp32", args.uneven_problems, autotune=not args.no_uneven_autotune)
        finally:
            set_group_mode(True)
        out["uneven_batch"]["fp32_one_fused_launch_per_plan"] = u32p
        _progress("uneven batch FP32, one fused launch per plan")
        if not args.no_fp64:
            u64, r64 = uneven_run("fp64", args.uneven_problems)
            out["uneven_batch"]["fp64"] = u64
            _progress("uneven batch FP64")
            out["uneven_batch"]["fp32_over_fp64"] = u32["solves_per_s"] / u64["solves_per_s"]
            errs = [np.abs(r32[k].f - r64[k].f).max() / np.abs(r64[k].f).max() for k in r64]
            out["uneven_batch"]["fp32_vs_fp64_final_f_rel_linf_max"] = float(max(errs))
    if args.serve_requests > 0:
        from proj_int_inelastic_amd import serving
        out["serving"] = {"workload": "Poisson arrivals (rate %g/s) of %d solves: 3/4 homogeneous TPI2 "
                                      "cooling (2D 16^2/32^2/64^2, 3D 16^3/32^3, t_final in {0.5 .. 3}), "
                                      "1/4 1D x 2D 32^2 PFE jobs (Nx in {64, 128, 256}, eps in "
                                      "{1e-2, 1e-3}, t_final = 2.5e-4 x U{2000..12000})"
                                      % (args.serve_rate, args.serve_requests)}
        sres = {}
        for prec in (("fp32", "fp64") if not args.no_fp64 else ("fp32",)):
            srv = serving.Server(hbm_budget=32 << 30, autotune=args.serve_autotune)
            sres[prec] = srv.run(serve_sweep_requests(prec, args.serve_requests, args.serve_rate,
                                                      seed=args.serve_seed, full_time=True),
                                 realtime=True)
            out["serving"][prec] = serving.latency_stats(sres[prec])
            print("serving %s %s" % (prec, out["serving"][prec]), file=sys.stderr, flush=True)
        if "fp64" in sres:
            errs = [np.abs(sres["fp32"][k].f - sres["fp64"][k].f).max() / np.abs(sres["fp64"][k].f).max()
                    for k in sres["fp64"]]
            out["serving"]["fp32_vs_fp64_final_f_rel_linf_max"] = float(max(errs))
        out["p99_solve_latency_s"] = out["serving"]["fp32"]["p99_s"]
    if not args.no_cpu:
        rates, ncpu, model = cpu_baseline(args.cpu_cells, seed=1)
        r8, n8, t8 = rates[8]
        ra, na, ta = rates[ncpu]
        out["cpu_baseline"] = {"value": r8, "unit": "cell-evals/s", "cores": 8, "kind": "port",
                               "sample": "%d cells of the same workload, one at a time (median), FP64 "
                                         "oracle with 8 FFT threads, %.1f s" % (n8, t8),
                               "all_cores": {"value": ra, "cores": ncpu, "cells": na, "seconds": ta},
                               "host": model}
        if not args.no_ref_configs:
            out["cpu_baseline"]["reference_configs"] = cpu_reference_configs()
        if args.uneven_problems > 0:
            out["cpu_baseline"]["uneven_batch"] = cpu_uneven_estimate(1.0 / r8)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
