"""FP64 restatement of the 1D-space TVD upwind transport step.

Follows the reference:
  * ghost cells     ``src/solver.py:51-105`` and ``src/state.py:109-122``
                    (g = 2; BC.extrap copies the first/last interior cell,
                    BC.periodic wraps)
  * Riemann solver  ``src/riemann/advection_1D.py:4-21``
                    (wave = f_r - f_l, A+dQ = max(s,0) wave, A-dQ = min(s,0) wave)
  * limiter         ``src/limiters/tvd.py:6-40`` (masked-array ratio, minmod)
  * wave propagation``src/boltzmann.py:177-205`` (LeVeque 2nd-order correction)
  * flux-form upwind ``src/riemann/upwind.py:6-50`` (van Leer, unwired there)

New: the MC limiter phi = max(0, min((1+theta)/2, 2, 2 theta)), and the
dimension-split 2D-space increment df_x + df_y (df_advection_2d), pinned by
reduction to the 1D step for data constant along one axis and by exact
transport of piecewise constant data at CFL 1 along either axis.

Pins (tests/test_oracle_transport_integrators.py, tests/test_golden_transport.py):
exact transport of piecewise constant data at CFL 1, discrete conservation
with periodic ghosts, total variation non-increase, and golden vectors written
from the reference's own numpy-only modules (tests/golden/) for the TVD step,
the limiter and the flux-form scheme.
"""
import numpy as np

EXTRAP = 1
PERIODIC = 2


def fill_ghosts(f, bc_lower, bc_upper, num_ghost=2, fbc=None):
    """f (Nx, ...) -> fbc (Nx + 2g, ...) with ghost cells (solver.py:59-105)."""
    g = num_ghost
    if fbc is None:
        fbc = np.zeros((f.shape[0] + 2 * g,) + f.shape[1:])
    fbc[g:-g] = f
    if bc_lower == EXTRAP:
        for i in range(g):
            fbc[i] = fbc[g]
    elif bc_lower == PERIODIC:
        fbc[:g] = fbc[-2 * g:-g]
    elif bc_lower is None:
        raise Exception("Lower boundary condition not specified.")
    else:
        raise NotImplementedError("Boundary condition %s not implemented" % bc_lower)
    if bc_upper == EXTRAP:
        for i in range(g):
            fbc[-i - 1] = fbc[-g - 1]
    elif bc_upper == PERIODIC:
        fbc[-g:] = fbc[g:2 * g]
    elif bc_upper is None:
        raise Exception("Upper boundary condition not specified.")
    else:
        raise NotImplementedError("Boundary condition %s not implemented" % bc_upper)
    return fbc


def advection_1D(q_l, q_r, problem_data):
    """Upwind Riemann solver (advection_1D.py:4-21)."""
    wave = q_r - q_l
    s = np.empty(q_l.shape[1:])
    s[:] = problem_data["v1"]
    apdq = np.maximum(s, 0.0) * wave
    amdq = np.minimum(s, 0.0) * wave
    return wave, s, amdq, apdq


def minmod(r, cfl):
    """phi = max(0, min(1, r)) (tvd.py:30-40); cfl unused."""
    return np.maximum(0.0, np.minimum(1.0, r))


def mc(r, cfl):
    """Monotonized-central limiter (new): max(0, min((1+r)/2, 2, 2r))."""
    return np.maximum(0.0, np.minimum(np.minimum(0.5 * (1.0 + r), 2.0), 2.0 * r))


def limit(wave, s, limiter, dtdx):
    """Wave limiting (tvd.py:6-27), masked division restated with np.ma."""
    wave = wave.copy()
    wave_norm2 = wave ** 2
    wave_zero_mask = np.array((wave == 0), dtype=float)
    wave_nonzero_mask = 1.0 - wave_zero_mask
    dotls = wave[1:] * wave[:-1]
    spos = np.array(s > 0.0, dtype=float)
    r = np.ma.array((spos * dotls[:-1] + (1 - spos) * dotls[1:]))
    r /= np.ma.array(wave_norm2[1:-1])
    r.fill_value = 0
    r = r.filled()
    cfl = np.abs(s * dtdx)
    wlimitr = limiter(r, cfl)
    wave[1:-1] = (
        wave[1:-1] * wave_zero_mask[1:-1]
        + wlimitr * wave[1:-1] * wave_nonzero_mask[1:-1]
    )
    return wave


def df_advection(f, v1, dt, dx, bc_lower=EXTRAP, bc_upper=EXTRAP, order=2,
                 limiter=minmod, num_ghost=2):
    """Increment of the advection step, already scaled by dt (boltzmann.py:177-205).

    f: (Nx, N, N) (any trailing velocity shape); v1: v_x on the velocity grid,
    broadcastable to f.shape[1:].  Returns (Nx, ...).
    """
    fbc = fill_ghosts(f, bc_lower, bc_upper, num_ghost)
    nx = f.shape[0]
    dtdx = dt / dx
    df = np.zeros(fbc.shape)
    f_l = fbc[:-1]
    f_r = fbc[1:]
    wave, s, amdf, apdf = advection_1D(f_l, f_r, {"v1": v1})
    LL = num_ghost - 1
    UL = nx + num_ghost + 1
    df[LL:UL] = -dtdx * (amdf[LL:UL] + apdf[LL - 1:UL - 1])
    if order == 2:
        F = np.zeros(fbc.shape)
        wave = limit(wave, s, limiter, dtdx)
        sabs = np.abs(s[LL - 1:UL - 1])
        om = 1.0 - sabs * dtdx
        F[LL:UL] = 0.5 * sabs * om * wave[LL - 1:UL - 1]
        df[LL:UL - 1] -= dtdx * (F[LL + 1:UL] - F[LL:UL - 1])
    return df[num_ghost:-num_ghost]


# ---- flux-form upwind of src/riemann/upwind.py (not called by any reference solver)
EPS = 1e-8


def van_leer_limiter(r):
    """phi(r) = (r + |r|) / (1 + |r|) (upwind.py:16-17)."""
    return (r + np.abs(r)) / (1.0 + np.abs(r))


def F_p_2(f, vp, dx, dt, limiter=van_leer_limiter):
    """upwind.py:22-27."""
    r = (f[1:-1] - f[:-2]) / (f[2:] - f[1:-1] + EPS)
    phi = limiter(r)
    F = f.copy()
    F[1:-1] += 0.5 * phi * (1.0 - vp * dt / dx) * (f[2:] - f[1:-1])
    return F[2:-2] - F[1:-3]


def F_m_2(f, vm, dx, dt, limiter=van_leer_limiter):
    """upwind.py:32-37."""
    r = (f[2:] - f[1:-1]) / (f[1:-1] - f[:-2] + EPS)
    phi = limiter(r)
    F = f.copy()
    F[1:-1] += 0.5 * phi * (-1.0 - vm * dt / dx) * (f[1:-1] - f[:-2])
    return F[3:-1] - F[2:-2]


def grad_vf2(f, v, dx, dt):
    """upwind.py:40-44."""
    vp = np.maximum(v, 0.0)
    vm = np.minimum(v, 0.0)
    return vp[2:-2] * F_p_2(f, vp[1:-1], dx, dt) + vm[2:-2] * F_m_2(f, vm[1:-1], dx, dt)


def grad_vf1(f, v):
    """upwind.py:47-50."""
    vp = np.maximum(v, 0)
    vm = np.minimum(v, 0)
    up, dn = vp * f, vm * f
    return (up[1:-1] - up[:-2]) + (dn[2:] - dn[1:-1])


def df_advection_2d(f, vx, vy, dt, dx, dy, bcx=(EXTRAP, EXTRAP), bcy=(EXTRAP, EXTRAP), order=2,
                    limiter=minmod):
    """New (the reference has no 2D df_advection): dimension-split increment
    df_x + df_y of f (nx, ny, N, N), each the 1D step of boltzmann.py:177-205
    along its space axis with speed vx (resp. vy) on the velocity grid.
    Computed line by line, so the reference's s[LL-1:UL-1] slice acts on the
    velocity axis as in 1D (harmless for nx + 2 >= N and ny + 2 >= N)."""
    nx, ny = f.shape[0], f.shape[1]
    dfx = np.empty_like(f)
    for j in range(ny):
        dfx[:, j] = df_advection(f[:, j], vx, dt, dx, bcx[0], bcx[1], order, limiter)
    dfy = np.empty_like(f)
    for i in range(nx):
        dfy[i] = df_advection(f[i], vy, dt, dy, bcy[0], bcy[1], order, limiter)
    return dfx + dfy
