"""FP64 restatement of the reference's fast-sweeping iteration (test infrastructure).

Follows reference ``src/iter_solver.py:38-62`` (``BoltzmannFastSweep.step``):
after the gain term g = Q+(f) and the density rho of the current iterate,
every velocity row with v_x > 0 is swept forward over the interior cells,
    f_i <- (nu g_i + (v_x/dx) f_{i-1}) / (nu rho_i + v_x/dx),   i = 1..nx-2,
and every row with v_x < 0 backward,
    f_i <- (nu g_i - (v_x/dx) f_{i+1}) / (nu rho_i - v_x/dx),   i = nx-2..1,
each using the value just updated (Gauss-Seidel); nu = 1/(sqrt(pi/2) kn).
The return value is max |f_new - f_old|.

Pins (tests/test_oracle_iter_solver.py): each row's sweep is the solve of a
bidiagonal linear system, checked against numpy.linalg.solve of that system;
a Maxwellian-like state with g = rho f, constant in x, is a fixed point.
"""
import math

import numpy as np


def sweep(f, gain, rho, vx, dx, nu):
    """In-place sweep of f (nx, N, N) with given gain (nx, N, N), rho (nx,), vx (N,)."""
    f_old = f.copy()
    nx = f.shape[0]
    a = vx / dx
    pos, neg = vx > 0, vx < 0
    r = rho[:, None, None]
    for i in range(1, nx - 1):
        f[i, pos] = (nu * gain[i, pos] + a[pos][:, None] * f[i - 1, pos]) / (nu * r[i] + a[pos][:, None])
    for i in range(nx - 2, 0, -1):
        f[i, neg] = (nu * gain[i, neg] - a[neg][:, None] * f[i + 1, neg]) / (nu * r[i] - a[neg][:, None])
    return np.max(np.abs(f - f_old))


def fast_sweep_step(f, gain_fn, vx, dx, dv, kn):
    """One BoltzmannFastSweep.step: rho = sum f dv^2 (state.py:91-92), gain = gain_fn(f)."""
    nu = 1 / np.sqrt(math.pi / 2) / kn
    rho = np.sum(f, axis=(-1, -2)) * dv * dv
    gain = gain_fn(f)
    return sweep(f, gain, rho, vx, dx, nu)
