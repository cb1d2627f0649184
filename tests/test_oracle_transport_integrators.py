"""Pins of oracle/transport.py and oracle/integrators.py — CPU only."""
import numpy as np
import pytest

from oracle import integrators as I
from oracle import transport as T


def _v1(N, L=8.0):
    dv = 2 * L / N
    v = -L + (np.arange(N) + 0.5) * dv
    return np.repeat(v[:, None], N, axis=1)  # axis -2 <-> v_x


def test_ghost_fill_extrap_periodic():
    f = np.arange(5.0)[:, None] * np.ones((5, 3))
    e = T.fill_ghosts(f, T.EXTRAP, T.EXTRAP)
    assert np.array_equal(e[:, 0], [0, 0, 0, 1, 2, 3, 4, 4, 4])
    p = T.fill_ghosts(f, T.PERIODIC, T.PERIODIC)
    assert np.array_equal(p[:, 0], [3, 4, 0, 1, 2, 3, 4, 0, 1])
    with pytest.raises(Exception):
        T.fill_ghosts(f, None, T.EXTRAP)


def test_first_order_exact_transport_cfl1():
    # s = +1 everywhere, dt/dx = 1: first-order upwind shifts the data by one cell
    nx = 12
    f = np.random.default_rng(0).random((nx, 1, 1))
    v1 = np.ones((1, 1))
    df = T.df_advection(f, v1, dt=0.1, dx=0.1, bc_lower=T.PERIODIC, bc_upper=T.PERIODIC, order=1)
    assert np.allclose(f + df, np.roll(f, 1, axis=0), atol=1e-15)
    # second order at CFL 1: the correction vanishes (1 - |s| dt/dx = 0)
    df2 = T.df_advection(f, v1, dt=0.1, dx=0.1, bc_lower=T.PERIODIC, bc_upper=T.PERIODIC, order=2)
    assert np.allclose(df2, df, atol=1e-15)


@pytest.mark.parametrize("limiter", [T.minmod, T.mc])
def test_conservation_and_tvd(limiter):
    nx, N = 40, 8
    rng = np.random.default_rng(1)
    f = rng.random((nx, N, N))
    f[10:20] += 2.0
    v1 = _v1(N, 4.0)
    dx = 1.0 / nx
    dt = 0.4 * dx / 4.0
    df = T.df_advection(f, v1, dt, dx, T.PERIODIC, T.PERIODIC, 2, limiter)
    assert abs(df.sum()) < 1e-13 * np.abs(f).sum()
    g = f + df
    tv = lambda a: np.abs(np.diff(np.concatenate([a, a[:1]]), axis=0)).sum(axis=0)
    assert np.all(tv(g) <= tv(f) + 1e-12)


def test_limiter_closed_forms():
    r = np.array([-1.0, 0.0, 0.5, 1.0, 3.0])
    assert np.array_equal(T.minmod(r, None), [0, 0, 0.5, 1, 1])
    assert np.array_equal(T.mc(r, None), [0, 0, 0.75, 1, 2])
    # limit(): zero waves stay zero and masked ratios become 0
    wave = np.array([0.0, 1.0, 0.0, 2.0, 1.0])[:, None]
    s = np.array([1.0])
    out = T.limit(wave, s, T.minmod, 0.1)
    assert np.array_equal(out[:, 0], [0.0, 0.0, 0.0, 0.0, 1.0])


def test_scalar_amplification_factors():
    lam = -3.0
    df = lambda f, dt: lam * dt * f
    dt, h = 0.1, [0.01, 0.03, 0.06]
    K = [2, 2, 2]
    a = 1 + lam * h[0]
    f = I.proj_euler_step(np.array([1.0]), df, dt, K, h)
    # f_K = a^K, last increment lam h a^(K-1): f = a^(K-1) (a + M lam h)
    assert abs(f[0] - a * (a + (dt / h[0] - 2) * lam * h[0])) < 1e-15
    # TPI2 as written (df1 = M0 df0)
    g = 1.0
    for _ in range(K[1]):
        for _ in range(K[0]):
            d0 = lam * h[0] * g
            g += d0
        d1 = (h[1] / h[0] - K[0]) * d0
        g += d1
    g += (dt / h[1] - K[1]) * d1
    f = I.tel_proj_step2(np.array([1.0]), df, dt, K, h)
    assert abs(f[0] - g) < 1e-15
    f = I.euler_step(np.array([1.0]), df, dt)
    assert abs(f[0] - (1 + lam * dt)) < 1e-15


def test_evolve_steps_and_divisibility():
    assert I.num_steps(0.0, 1.0, 0.01) == 100
    with pytest.raises(Exception):
        I.num_steps(0.0, 1.0, 0.03)
    f, t = I.evolve(np.array([1.0]), lambda f, dt: -dt * f, "Euler", 0.25, 0.0, 1.0)
    assert t == 1.0 and abs(f[0] - 0.75 ** 4) < 1e-15


def test_oracle_2d_split_reduces_to_1d_and_transports_exactly():
    """df_advection_2d: data constant in y with v_y = 0 gives the 1D increment on every
    y line; first order at CFL 1 along y (v_x = 0) shifts piecewise constant data by
    one cell with periodic ghosts."""
    from oracle import transport as T
    rng = np.random.default_rng(4)
    nx, ny, N = 10, 8, 4
    v = np.linspace(-1.5, 1.5, N)
    vx = np.broadcast_to(v[:, None], (N, N)).copy()
    prof = rng.random((nx, N, N))
    f = np.broadcast_to(prof[:, None], (nx, ny, N, N)).copy()
    dfy0 = T.df_advection_2d(f, vx, np.zeros((N, N)), 0.01, 0.1, 0.1, (T.PERIODIC,) * 2,
                             (T.EXTRAP,) * 2)
    ref = T.df_advection(prof, vx, 0.01, 0.1, T.PERIODIC, T.PERIODIC)
    for j in range(ny):
        assert np.array_equal(dfy0[:, j], ref)
    # exact shift along y at CFL 1 (first order, v_y = +1 everywhere, v_x = 0)
    g = rng.random((nx, ny, N, N))
    dy = 0.1
    df = T.df_advection_2d(g, np.zeros((N, N)), np.ones((N, N)), dy, 0.1, dy, (T.PERIODIC,) * 2,
                           (T.PERIODIC,) * 2, order=1)
    np.testing.assert_allclose(g + df, np.roll(g, 1, axis=1), rtol=0, atol=1e-15)
