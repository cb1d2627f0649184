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


def _tpi3_factor(lam, dt, K, h):
    """Amplification of one TPI3 step on f' = lam f, in closed form.

    Inner Euler factor a = 1 + lam h0; the last inner increment of a level-0 block is
    c0 = lam h0 a^(K0-1) times the block's starting value.  With df1 = M0 df0 and
    df2 = M1 df1 as the reference writes them (boltzmann.py:142-158), a level-1 step
    multiplies by A1 = a^K0 + M0 c0, a level-2 step by A2 = A1^K1 + M1 M0 c0 A1^(K1-1),
    and the step by A2^K2 + M2 M1 M0 c0 A1^(K1-1) A2^(K2-1), where M0 = h1/h0 - K0,
    M1 = h2/h1 - K1, M2 = dt/h2 - K2.
    """
    a = 1 + lam * h[0]
    M0, M1, M2 = h[1] / h[0] - K[0], h[2] / h[1] - K[1], dt / h[2] - K[2]
    c0 = lam * h[0] * a ** (K[0] - 1)
    A1 = a ** K[0] + M0 * c0
    A2 = A1 ** K[1] + M1 * M0 * c0 * A1 ** (K[1] - 1)
    return A2 ** K[2] + M2 * M1 * M0 * c0 * A1 ** (K[1] - 1) * A2 ** (K[2] - 1)


@pytest.mark.parametrize("lam,dt,K,h", [
    (-3.0, 0.2, (2, 2, 2), (0.01, 0.03, 0.06)),
    (-1.0, 1.0, (3, 3, 3), (0.01, 0.05, 0.1)),      # the reference defaults: M1 = 10/5 - 3 = -1
    (-20.0, 0.07, (1, 2, 3), (0.004, 0.01, 0.02)),
    (0.5, 0.3, (4, 1, 2), (0.02, 0.1, 0.12)),
])
def test_tpi3_amplification_factor(lam, dt, K, h):
    df = lambda f, step: lam * step * f
    f = I.tel_proj_step3(np.array([1.0, -2.0]), df, dt, K, h)
    g = _tpi3_factor(lam, dt, K, h)
    np.testing.assert_allclose(f, [g, -2.0 * g], rtol=1e-13, atol=0)


def test_tpi_levels_aligned_reduce_to_euler():
    """With h1 = K0 h0, h2 = K1 h1, dt = K2 h2 every extrapolation factor is 0 and TPI2 / TPI3 /
    PFE are plain Euler steps of h0."""
    lam, h0 = -2.5, 0.01
    df = lambda f, step: lam * step * f
    K = (3, 2, 4)
    h = (h0, K[0] * h0, K[0] * K[1] * h0)
    dt3 = K[2] * h[2]
    a = 1 + lam * h0
    f3 = I.tel_proj_step3(np.array([1.0]), df, dt3, K, h)
    assert abs(f3[0] - a ** (K[0] * K[1] * K[2])) < 1e-14
    f2 = I.tel_proj_step2(np.array([1.0]), df, h[2], K, h)
    assert abs(f2[0] - a ** (K[0] * K[1])) < 1e-14
    f1 = I.proj_euler_step(np.array([1.0]), df, h[1], K, h)
    assert abs(f1[0] - a ** K[0]) < 1e-14


def test_tpi3_planted_df1_mistake_fails():
    """Using the full level-1 increment instead of df1 = M0 df0 changes the factor: the pin sees it."""
    lam, dt, K, h = -3.0, 0.2, (2, 2, 2), (0.01, 0.03, 0.06)
    a = 1 + lam * h[0]
    M0 = h[1] / h[0] - K[0]
    wrong_A1 = a ** K[0] * (1 + M0 * lam * h[0])     # extrapolating the whole block instead
    right_A1 = a ** K[0] + M0 * lam * h[0] * a ** (K[0] - 1)
    assert abs(wrong_A1 - right_A1) > 1e-4


def test_convex_splitting_is_the_penalised_implicit_step():
    """CONSP (boltzmann.py:86-94): without transport, f1 solves
    kn (f1 - f0) / dt = Q(f0) - p (f1 - f0) -- the penalty taken implicitly; with Q = 0 the
    step is the transport step alone."""
    rng = np.random.default_rng(11)
    A = rng.normal(size=(6, 6))
    Q = lambda f: A @ f
    f0 = rng.normal(size=6)
    dt, kn, p = 0.03, 0.2, 7.0
    f1 = I.convex_splitting_step(f0.copy(), Q, lambda f, step: 0 * f, dt, kn, p)
    np.testing.assert_allclose(kn * (f1 - f0) / dt, Q(f0) - p * (f1 - f0), rtol=0, atol=1e-13)
    # p = 0: forward Euler of Q / kn
    f1 = I.convex_splitting_step(f0.copy(), Q, lambda f, step: 0 * f, dt, kn, 0.0)
    np.testing.assert_allclose(f1, f0 + dt / kn * Q(f0), rtol=0, atol=1e-14)
    # transport applied to the post-collision state (sequential split)
    adv = lambda f, step: -step * np.roll(f, 1)
    f1 = I.convex_splitting_step(f0.copy(), lambda f: 0 * f, adv, dt, kn, p)
    np.testing.assert_allclose(f1, f0 - dt * np.roll(f0, 1), rtol=0, atol=1e-15)


def _maxwell_grid(N=64, L=11.0):
    dv = 2 * L / N
    v = -L + (np.arange(N) + 0.5) * dv
    vx, vy = np.meshgrid(v, v, indexing="ij")
    return vx, vy, dv


def _moments(f, vx, vy, dv):
    s = lambda g: np.sum(g, axis=(-1, -2)) * dv * dv
    rho = s(f)
    u = s(f * vx) / rho
    T = 0.5 * s(f * (vx ** 2 + vy ** 2)) / rho - 0.5 * u ** 2 - 0.5 * (s(f * vy) / rho) ** 2
    return rho, u, T


def test_maxwellian_2d_moments():
    vx, vy, dv = _maxwell_grid()
    rho, u, T = np.array([1.0, 0.3]), np.array([0.0, 0.7]), np.array([1.0, 1.6])
    f = I.maxwellian_2d(vx, vy, rho, u, T)
    r, uu, TT = _moments(f, vx, vy, dv)
    np.testing.assert_allclose(r, rho, rtol=1e-12)
    np.testing.assert_allclose(uu, u, rtol=0, atol=1e-12)
    np.testing.assert_allclose(TT, T, rtol=1e-12)


def test_convex_splitting_maxwellian_haff_cooling_and_fixed_point():
    """CONSP_M (boltzmann.py:96-118) with Q = 0 and a very large penalty takes f to the
    Maxwellian of temperature T_new: for tau = 0 Haff's law T exp(-(1 - e^2) rho dt / (4 kn)), the
    solution of dT/dt = -(1 - e^2) rho T / (4 kn); for tau > 0 the steady temperature
    8 tau / (1 - e^2) is a fixed point (f unchanged)."""
    vx, vy, dv = _maxwell_grid()
    rho, u, T = np.array([1.0, 0.5]), np.array([0.2, -0.1]), np.array([1.2, 0.9])
    f0 = I.maxwellian_2d(vx, vy, rho, u, T)
    zero = lambda f: 0 * f
    noadv = lambda f, step: 0 * f
    e, kn, dt = 0.8, 0.5, 0.05
    f1 = I.convex_splitting_maxwellian_step(f0.copy(), zero, noadv, dt, kn, 1e13, e, 0.0, vx, vy, dv)
    r1, u1, T1 = _moments(f1, vx, vy, dv)
    np.testing.assert_allclose(r1, rho, rtol=1e-11)
    np.testing.assert_allclose(u1, u, rtol=0, atol=1e-11)
    np.testing.assert_allclose(T1, T * np.exp(-(1 - e ** 2) * rho * dt / (4 * kn)), rtol=1e-10)
    # two half steps compose to one full step (exponential law)
    fh = I.convex_splitting_maxwellian_step(f0.copy(), zero, noadv, dt / 2, kn, 1e13, e, 0.0, vx, vy, dv)
    fh = I.convex_splitting_maxwellian_step(fh, zero, noadv, dt / 2, kn, 1e13, e, 0.0, vx, vy, dv)
    np.testing.assert_allclose(_moments(fh, vx, vy, dv)[2], T1, rtol=1e-10)
    tau = 0.01
    Ts = 8 * tau / (1 - e ** 2)
    fs = I.maxwellian_2d(vx, vy, rho, u, np.full(2, Ts))
    # the grid moments of a Maxwellian equal its parameters to ~1e-14, so T_new = T up to that
    f2 = I.convex_splitting_maxwellian_step(fs.copy(), zero, noadv, dt, kn, 3.0, e, tau, vx, vy, dv)
    np.testing.assert_allclose(f2, fs, rtol=0, atol=1e-12 * fs.max())
