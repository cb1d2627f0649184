"""FP64 restatement of the reference time integrators.

Follows reference ``src/boltzmann.py``:
  * euler_step      :82-84    f += df(dt)
  * proj_euler_step :120-128  K0 inner steps of dt0, then f += M df_last,
                              M = dt/dt0 - K0
  * tel_proj_step2  :130-140  (df1 = M0 df0, reproduced as written)
  * tel_proj_step3  :142-158
  * df              :160-164  df_advection(dt) + dt Q(f)/kn
and ``src/solver.py:107-143`` for ``evolve_to_time`` (fixed dt, raises if dt
does not divide the interval).

``df`` here is any callable df(f, dt) -> increment; the oracle composes it from
oracle.transport and oracle.collision.  Pins: the scalar test equation
f' = lam f has closed-form amplification factors for every scheme
(tests/test_oracle_integrators.py).
"""


def euler_step(f, df, dt):
    f += df(f, dt)
    return f


def proj_euler_step(f, df, dt, inner_steps, inner_dt):
    M = dt / inner_dt[0] - inner_steps[0]
    d = None
    for _ in range(inner_steps[0]):
        d = df(f, inner_dt[0])
        f += d
    f += M * d
    return f


def tel_proj_step2(f, df, dt, inner_steps, inner_dt):
    M1 = dt / inner_dt[1] - inner_steps[1]
    df1 = None
    for _ in range(inner_steps[1]):
        M0 = inner_dt[1] / inner_dt[0] - inner_steps[0]
        df0 = None
        for _ in range(inner_steps[0]):
            df0 = df(f, inner_dt[0])
            f += df0
        df1 = M0 * df0
        f += df1
    f += M1 * df1
    return f


def tel_proj_step3(f, df, dt, inner_steps, inner_dt):
    M2 = dt / inner_dt[2] - inner_steps[2]
    df2 = None
    for _ in range(inner_steps[2]):
        M1 = inner_dt[2] / inner_dt[1] - inner_steps[1]
        df1 = None
        for _ in range(inner_steps[1]):
            M0 = inner_dt[1] / inner_dt[0] - inner_steps[0]
            df0 = None
            for _ in range(inner_steps[0]):
                df0 = df(f, inner_dt[0])
                f += df0
            df1 = M0 * df0
            f += df1
        df2 = M1 * df1
        f += df2
    f += M2 * df2
    return f


STEPS = {
    "Euler": lambda f, df, dt, K, h: euler_step(f, df, dt),
    "PFE": proj_euler_step,
    "TPI2": tel_proj_step2,
    "TPI3": tel_proj_step3,
}


def num_steps(t0, tend, dt):
    """solver.py:122-128."""
    n = int((tend - t0 + 1e-10) / dt)
    if abs(n * dt - (tend - t0)) > 1e-5 * (tend - t0):
        raise Exception("dt does not divide (tend-tstart) and dt is fixed!")
    return n


def evolve(f, df, integrator, dt, t0, tend, inner_steps=(3, 3, 3), inner_dt=None,
           callback=None):
    """Fixed-step loop of solver.py:107-143; returns (f, t)."""
    step = STEPS[integrator]
    n = num_steps(t0, tend, dt)
    t = t0
    for i in range(n):
        f = step(f, df, dt, inner_steps, inner_dt)
        t = t0 + (i + 1) * dt
        if callback is not None:
            callback(i, t, f)
    return f, t


def convex_splitting_step(f, collision, df_advection, dt, kn, penalty):
    """boltzmann.py:86-94: f += dt/(kn + p dt) Q(f); f += df_advection(f, dt)."""
    f += dt / (kn + penalty * dt) * collision(f)
    f += df_advection(f, dt)
    return f


def maxwellian_2d(vx, vy, rho, u, T):
    """boltzmann.py:9-13 (per x-cell, u along v_x)."""
    import numpy as np
    rho = rho[:, None, None]
    u = u[:, None, None]
    T = T[:, None, None]
    return rho / (2 * np.pi * T) * np.exp(-((vx - u) ** 2 + vy ** 2) / (2 * T))


def convex_splitting_maxwellian_step(f, collision, df_advection, dt, kn, penalty, e, tau, vx, vy, dv):
    """boltzmann.py:96-118 with state.py:31-52 moments (2D velocity, f (Nx, N, N))."""
    import numpy as np
    sum_f = lambda g: np.sum(g, axis=(-1, -2)) * dv * dv
    rho = sum_f(f)
    u = [sum_f(f * c) / rho for c in (vx, vy)]
    E = 0.5 * sum_f(f * (vx ** 2 + vy ** 2))
    T = E / rho - 0.5 * (u[0] ** 2 + u[1] ** 2)
    T_new = (T - 8 * tau / (1 - e ** 2)) * np.exp(-(1 - e ** 2) / kn / 4 * rho * dt) + 8 * tau / (1 - e ** 2)
    dm = maxwellian_2d(vx, vy, rho, u[0], T_new) - maxwellian_2d(vx, vy, rho, u[0], T)
    f += dt / (kn + penalty * dt) * collision(f) + penalty * dt / (kn + penalty * dt) * dm
    f += df_advection(f, dt)
    return f
