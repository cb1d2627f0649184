"""Pins of the oracle's fast-sweeping iteration (src/iter_solver.py:38-62), CPU only."""
import math

import numpy as np

from oracle import iter_solver as IS


def _inputs(nx=9, N=6, seed=0):
    rng = np.random.default_rng(seed)
    f = rng.random((nx, N, N)) + 0.1
    gain = rng.random((nx, N, N))
    rho = rng.random(nx) + 0.5
    vx = np.linspace(-2.5, 2.5, N)  # includes no zero row for even N
    vx[1] = 0.0                      # one row with v_x = 0 stays untouched
    return f, gain, rho, vx


def test_sweep_is_the_bidiagonal_solve():
    """Forward rows solve (nu rho_i + a) f_i - a f_{i-1} = nu g_i, i = 1..nx-2, with f_0 fixed;
    backward rows (nu rho_i - a) f_i + a f_{i+1} = nu g_i with f_{nx-1} fixed."""
    f, gain, rho, vx = _inputs()
    dx, nu = 0.1, 1 / np.sqrt(math.pi / 2) / 0.3
    out = f.copy()
    IS.sweep(out, gain, rho, vx, dx, nu)
    nx, N = f.shape[0], f.shape[1]
    m = nx - 2
    for r in range(N):
        a = vx[r] / dx
        for c in range(N):
            if vx[r] == 0:
                assert np.array_equal(out[:, r, c], f[:, r, c])
                continue
            A = np.zeros((m, m))
            b = nu * gain[1:-1, r, c].copy()
            for k in range(m):
                i = k + 1
                if a > 0:
                    A[k, k] = nu * rho[i] + a
                    if k > 0:
                        A[k, k - 1] = -a
                    else:
                        b[k] += a * f[0, r, c]
                else:
                    A[k, k] = nu * rho[i] - a
                    if k < m - 1:
                        A[k, k + 1] = a
                    else:
                        b[k] -= a * f[nx - 1, r, c]
            x = np.linalg.solve(A, b)
            np.testing.assert_allclose(out[1:-1, r, c], x, rtol=1e-13, atol=0)
            assert out[0, r, c] == f[0, r, c] and out[-1, r, c] == f[-1, r, c]


def test_sweep_fixed_point_and_error():
    """g = rho f with f constant in x is a fixed point (error 0); a perturbed iterate reports
    max |f_new - f_old|."""
    nx, N = 7, 4
    prof = np.random.default_rng(3).random((N, N)) + 0.2
    f = np.broadcast_to(prof, (nx, N, N)).copy()
    rho = np.full(nx, 1.7)
    gain = rho[:, None, None] * f
    vx = np.array([-1.5, -0.5, 0.5, 1.5])
    err = IS.sweep(f, gain, rho, vx, 0.2, 3.0)
    np.testing.assert_allclose(f, np.broadcast_to(prof, (nx, N, N)), rtol=1e-15)
    assert err < 1e-15
    g = f.copy()
    g[3] += 0.25
    before = g.copy()
    err = IS.sweep(g, gain, rho, vx, 0.2, 3.0)
    assert err == np.max(np.abs(g - before)) and err > 0


def test_planted_jacobi_mistake_is_caught():
    """A Jacobi sweep (using the old f_{i-1}) differs from the Gauss-Seidel solve."""
    f, gain, rho, vx = _inputs(seed=5)
    dx, nu = 0.1, 2.0
    gs = f.copy()
    IS.sweep(gs, gain, rho, vx, dx, nu)
    jac = f.copy()
    a = vx / dx
    pos = vx > 0
    for i in range(1, f.shape[0] - 1):
        jac[i, pos] = (nu * gain[i, pos] + a[pos][:, None] * f[i - 1, pos]) / (nu * rho[i] + a[pos][:, None])
    assert not np.allclose(gs[:, pos], jac[:, pos], rtol=1e-6)
