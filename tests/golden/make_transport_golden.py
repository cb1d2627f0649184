"""Write golden vectors of the reference's TVD transport step.

Runs ONLY in the build container where the reference is mounted: imports the
reference's numpy-only modules (src/boltzmann.py, src/solution.py,
src/riemann/advection_1D.py, src/limiters/tvd.py) and records
BoltzmannSolver1D.df_advection and tvd.limit outputs for seeded inputs into
transport_golden.npz (inputs and outputs only).  These pin the masked-array
ratio, the zero-wave handling and the index arithmetic of
boltzmann.py:177-205, which no closed form fixes bit for bit.  It also records
the flux-form scheme of src/riemann/upwind.py (grad_vf2, grad_vf1, F_p_2,
F_m_2), whose EPS-regularised van Leer ratio no closed form fixes either.
"""
import pathlib
import sys
import types

import numpy as np

REF = pathlib.Path("/root/reference")
OUT = pathlib.Path(__file__).resolve().parent / "transport_golden.npz"


def main():
    sys.path.insert(0, str(REF))
    from src.boltzmann import BoltzmannSolver1D
    from src.limiters import tvd
    from src.riemann.advection_1D import advection_1D
    from src.solution import Solution
    from src.solver import BC

    rng = np.random.default_rng(2024)
    out = {}
    nx, nv, l = 24, 8, 8.0
    cfg = types.SimpleNamespace(dim_x=1, dim_v=2, xmin=0.0, xmax=1.0, l=l, nx=nx, nv=nv)
    base = rng.random((nx, nv, nv))
    base[5:9] = 0.7            # flat region: zero waves
    base[14] = base[13]        # one zero wave
    base[18, :, :3] = 1e-300   # tiny values near the masking threshold
    out["f"] = base
    case = 0
    for bc in (BC.extrap, BC.periodic):
        for order in (1, 2):
            for dt in (2.5e-4, 1.0e-3):
                sol = Solution(cfg)
                sol.state.f = base.copy()
                s = BoltzmannSolver1D(kn=1.0, riemann_solver=advection_1D)
                s.order = order
                s.bc_lower[0] = bc
                s.bc_upper[0] = bc
                s._allocate_bc_arrays(sol.state)
                df = s.df_advection(sol.state, dt)
                out["case%d_params" % case] = np.array([bc, order, dt])
                out["case%d_df" % case] = df
                case += 1
    out["v1"] = Solution(cfg).state.problem_data["v1"]
    out["dx"] = np.array([1.0 / nx])
    # limiter alone
    wave = rng.normal(size=(12, 3))
    wave[3] = 0.0
    wave[7, 1] = 1e-310
    s = np.array([-1.0, 0.5, 2.0])
    out["limit_wave"] = wave
    out["limit_s"] = s
    out["limit_out"] = tvd.limit(wave.copy(), s, tvd.minmod, 0.1)
    # flux-form upwind scheme (separate generator: the arrays above stay as they were)
    from src.riemann import upwind
    rng2 = np.random.default_rng(77)
    uf = rng2.random((14, 3, 4))
    uf[4:7] = 0.3                  # flat rows: ratios of EPS-sized denominators
    uf[10, 1] = uf[9, 1]
    uv = rng2.normal(size=(14, 3, 4))
    uv[2] = 0.0
    uv1 = rng2.normal(size=(14, 1, 1))
    udx, udt = 0.05, 0.01
    out["up_f"], out["up_v"], out["up_v1"] = uf, uv, uv1
    out["up_dxdt"] = np.array([udx, udt])
    out["up_grad_vf2"] = upwind.grad_vf2(uf, uv, udx, udt)
    out["up_grad_vf2_b"] = upwind.grad_vf2(uf, uv1, udx, udt)
    out["up_grad_vf1"] = upwind.grad_vf1(uf, uv)
    vp, vm = np.maximum(uv, 0.0), np.minimum(uv, 0.0)
    out["up_F_p_2"] = upwind.F_p_2(uf, vp[1:-1], udx, udt)
    out["up_F_m_2"] = upwind.F_m_2(uf, vm[1:-1], udx, udt)
    np.savez(OUT, **out)
    print("wrote", OUT, "cases", case)


if __name__ == "__main__":
    main()
