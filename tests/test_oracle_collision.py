"""Pins of the FP64 oracle (oracle/collision.py) — CPU only.

Each check is chosen so that a plausible mistake in the restated
reference algorithm (inelastic_collisions.py:57-137, 208-230) fails:
brute force catches FFT sequencing/normalisation/index-order mistakes,
BKW catches kernel-normalisation and loss-term mistakes, Haff catches a
wrong restitution factor a_j, the invariants catch batch/axis mix-ups.
"""
import math

import numpy as np
import pytest

from oracle import brute
from oracle import collision as C
from oracle import quadrature as Qd

S_C1 = 2.718491035932  # L = 6
S_C3 = 3.624654714576  # L = 8


def _two_gauss(op, seed=0, T=0.5):
    rng = np.random.default_rng(seed)
    v = op.v_grids()
    c1 = np.array([1.0, 0.0] + [0.0] * (op.d - 2)) + rng.normal(0, 0.1, op.d)
    c2 = -np.array([1.0, 0.0] + [0.0] * (op.d - 2)) + rng.normal(0, 0.1, op.d)

    def g(c):
        return np.exp(-sum((v[a] - c[a]) ** 2 for a in range(op.d)) / (2 * T)) / (2 * math.pi * T) ** (op.d / 2)

    f = 0.5 * (g(c1) + g(c2))
    return f / (f.sum() * op.delta ** op.d)


@pytest.mark.parametrize("d,N,nr,M", [(2, 8, 4, 5), (3, 4, 3, 14)])
def test_brute_force_matches(d, N, nr, M):
    rng = np.random.default_rng(1)
    if d == 2:
        sig, ws = Qd.circle_rule(M)
        ws = ws * (1 + 0.1 * np.arange(M))  # non-uniform weights exercise the per-node path
    else:
        sig, ws = Qd.lebedev_rule(M)
    op = C.OracleOperator(d, N, 2.0, nr, sig, ws, e=0.6, e0=0.95, gamma=1.0)
    f = rng.random((N,) * d)
    gp, gl = op.gain_loss(f)
    bg, bl = brute.brute_collide(f, d, N, op.L, op.R, op.r, op.wr, op.sigma, op.wsigma, 0.6, 0.95, 1.0)
    assert np.abs(gp - bg).max() <= 1e-13 * np.abs(bg).max()
    assert np.abs(gl - bl).max() <= 1e-13 * np.abs(bl).max()


def _bkw2(v2, K):
    return 1 / (2 * math.pi * K * K) * np.exp(-v2 / (2 * K)) * (2 * K - 1 + (1 - K) * v2 / (2 * K))


def test_bkw_2d_elastic_maxwell():
    # B = 1/(2 pi), K = 1 - exp(-t/8)/2 ; Q = df/dt
    sig, ws = Qd.circle_rule(8)
    op = C.OracleOperator(2, 32, S_C1, 32, sig, ws, e=1.0, e0=1.0, gamma=0.0)
    v = op.v_grids()
    v2 = v[0] ** 2 + v[1] ** 2
    t = 1.0
    K = 1 - math.exp(-t / 8) / 2
    h = 1e-5
    dfdt = (_bkw2(v2, K + h) - _bkw2(v2, K - h)) / (2 * h) * math.exp(-t / 8) / 16
    Q = op.collide(_bkw2(v2, K), with_loss=True)
    err = np.abs(Q - dfdt).max() / np.abs(dfdt).max()
    assert err < 1e-4, err


def test_bkw_3d_elastic_maxwell():
    # B = 1/(4 pi), K = 1 - exp(-t/6)
    sig, ws = Qd.lebedev_rule(38)
    op = C.OracleOperator(3, 24, 3.0, 24, sig, ws, e=1.0, e0=1.0, gamma=0.0)
    v = op.v_grids()
    v2 = sum(x * x for x in v)

    def bkw(K):
        return (2 * math.pi * K) ** -1.5 * np.exp(-v2 / (2 * K)) * ((5 * K - 3) / (2 * K) + (1 - K) * v2 / (2 * K * K))

    t = 5.5
    K = 1 - math.exp(-t / 6)
    h = 1e-5
    dfdt = (bkw(K + h) - bkw(K - h)) / (2 * h) * math.exp(-t / 6) / 6
    Q = op.collide(bkw(K), with_loss=True)
    err = np.abs(Q - dfdt).max() / np.abs(dfdt).max()
    assert err < 1e-3, err


@pytest.mark.parametrize("M,tol", [(8, 0.08), (64, 3e-3)])
def test_haff_energy_identity(M, tol):
    # inelastic Maxwell molecules: dT/dt = -(1 - e^2) rho T / 4 (any d);
    # reference boltzmann.py:103-105 encodes the same rate.
    e = 0.8
    sig, ws = Qd.circle_rule(M)
    op = C.OracleOperator(2, 32, S_C1, 32, sig, ws, e=e, e0=e, gamma=0.0)
    f = _two_gauss(op)
    v, dv = op.v_grids(), op.delta
    Q = op.collide(f, with_loss=True)
    rho, _, _ = C.moments(f, v, dv)
    _, qm, qe = C.moments(Q, v, dv)
    T = C.temperature(f, v, dv)
    rate = qe / rho  # dT/dt = 2 dE/dt / (d rho) with d = 2 and momentum conserved
    assert abs(rate / (-(1 - e * e) * rho * T / 4) - 1) < tol


def test_maxwellian_elastic_equilibrium():
    sig, ws = Qd.circle_rule(16)
    op = C.OracleOperator(2, 32, S_C1, 32, sig, ws, e=1.0, e0=1.0, gamma=1.0)
    v = op.v_grids()
    m = np.exp(-(v[0] ** 2 + v[1] ** 2) / 2) / (2 * math.pi)
    assert np.abs(op.collide(m, with_loss=True)).max() < 1e-5 * m.max()


def test_invariants_scaling_shift_batch():
    sig, ws = Qd.circle_rule(8)
    op = C.OracleOperator(2, 16, 2.0, 8, sig, ws, e=0.7, e0=0.9, gamma=1.0)
    rng = np.random.default_rng(3)
    f = rng.random((16, 16))
    q = op.collide(f, with_loss=True)
    assert np.allclose(op.collide(2.5 * f, with_loss=True), 6.25 * q, rtol=0, atol=1e-13 * np.abs(q).max() * 6.25)
    fr = np.roll(f, (3, -5), axis=(0, 1))
    qr = op.collide(fr, with_loss=True)
    assert np.abs(qr - np.roll(q, (3, -5), axis=(0, 1))).max() < 1e-13 * np.abs(q).max()
    fb = np.stack([f, 2 * f, fr])
    qb = op.collide(fb, with_loss=True)
    assert np.abs(qb[0] - q).max() < 1e-14 * np.abs(q).max()
    assert np.abs(qb[2] - qr).max() < 1e-14 * np.abs(q).max()


def test_gain_only_default_and_heat_bath():
    sig, ws = Qd.circle_rule(8)
    op = C.OracleOperator(2, 16, 2.0, 8, sig, ws, e=0.7, e0=0.7)
    f = np.random.default_rng(4).random((16, 16))
    gp, gl = op.gain_loss(f)
    assert np.array_equal(op.collide(f), gp)
    assert np.allclose(op.collide(f, with_loss=True), gp - gl, atol=1e-14)
    # Laplacian of a periodic plane wave
    v = op.v_grids()
    k = 3
    w = np.cos(math.pi * k * (v[0] + op.L) / op.L)
    lap = op.laplacian(w)
    assert np.abs(lap + (math.pi * k / op.L) ** 2 * w).max() < 1e-12
    out = op(f, heat_bath=0.3)
    assert np.allclose(out, gp + 0.3 * op.laplacian(f), atol=1e-14)


def test_projection_conserves_to_roundoff():
    sig, ws = Qd.lebedev_rule(14)
    op = C.OracleOperator(3, 8, 2.0, 4, sig, ws, e=0.5, e0=0.5, gamma=1.0)
    f = np.random.default_rng(5).random((2, 8, 8, 8))
    q = op.collide(f, with_loss=True)
    qc = C.project_conservative(q, op.v_grids())
    v = op.v_grids()
    scale = np.abs(q).sum()
    assert abs(qc.sum()) < 1e-14 * scale
    for a in range(3):
        assert abs((qc * v[a]).sum()) < 1e-14 * scale * op.L
    # energy is left free: projection of an energy-only perturbation keeps it
    assert np.abs(C.project_conservative(qc, v) - qc).max() < 1e-15 * np.abs(q).max() * 10


def test_mass_momentum_quadrature_accuracy():
    sig, ws = Qd.circle_rule(64)
    op = C.OracleOperator(2, 32, S_C1, 32, sig, ws, e=0.8, e0=0.8)
    f = _two_gauss(op, seed=7)
    q = op.collide(f, with_loss=True)
    rho, mom, _ = C.moments(q, op.v_grids(), op.delta)
    assert abs(rho) < 1e-7 and np.abs(mom).max() < 1e-7, (rho, mom)


def test_planted_mistake_is_caught(monkeypatch):
    """Plant a sign error in the translate kernel; the BKW pin must fail."""
    sig, ws = Qd.circle_rule(8)
    op = C.OracleOperator(2, 32, S_C1, 32, sig, ws, e=1.0, e0=1.0, gamma=0.0)
    op.exp = np.conj(op.exp)
    v = op.v_grids()
    v2 = v[0] ** 2 + v[1] ** 2
    K = 1 - math.exp(-1 / 8) / 2
    h = 1e-5
    dfdt = (_bkw2(v2, K + h) - _bkw2(v2, K - h)) / (2 * h) * math.exp(-1 / 8) / 16
    Q = op.collide(_bkw2(v2, K), with_loss=True)
    assert np.abs(Q - dfdt).max() / np.abs(dfdt).max() > 1e-2


def test_symmetrized_operator_equals_c2c_for_band_limited_data():
    """The Hermitian-symmetrised operator (what the real-data FP32 form
    computes) equals the reference C2C operator when fhat vanishes on every
    Nyquist plane, and differs at first order in that content otherwise."""
    sig, ws = Qd.lebedev_rule(14)
    op = C.OracleOperator(3, 8, 2.0, 4, sig, ws, e=0.6, e0=0.9, gamma=1.0)
    sy = op.symmetrized()
    rng = np.random.default_rng(8)
    f = rng.random((8, 8, 8))
    fh = np.fft.fftn(f)
    fh[4, :, :] = 0
    fh[:, 4, :] = 0
    fh[:, :, 4] = 0
    fb = np.fft.ifftn(fh).real
    q = op.collide(fb, with_loss=True)
    assert np.abs(sy.collide(fb, with_loss=True) - q).max() < 1e-14 * np.abs(q).max()
    q = op.collide(f, with_loss=True)
    assert np.abs(sy.collide(f, with_loss=True) - q).max() > 1e-6 * np.abs(q).max()


@pytest.mark.parametrize("N,nr,M,seed", [(8, 3, 14, 2), (16, 2, 6, 3)])
def test_nyquist_term_completes_symmetrized(N, nr, M, seed):
    """The anti-Hermitian (Nyquist-plane) term the FP32 3D form adds to its real-data
    operator: (i) restated through the plane factorisation it equals the direct sum
    G_A * fft(f * ifft(E_A fhat)) built from the oracle's own (Nr, M, n) tables, and
    (ii) the symmetrised operator plus it is the reference's C2C operator, to round-off,
    on data with O(1) Nyquist content (e != e0, hard spheres, Lebedev weights)."""
    sig, ws = Qd.lebedev_rule(M)
    op = C.OracleOperator(3, N, 2.5, nr, sig, ws, e=0.55, e0=0.9, gamma=1.0)
    rng = np.random.default_rng(seed)
    f = rng.random((N,) * 3) + 0.3
    term = op.nyquist_term(f)
    F = np.fft.fftn(f)
    direct = np.zeros_like(term)
    for j in range(nr):
        for m in range(op.M):
            E, G = op.exp[j, m], op.gain[j, m]
            EA = 0.5 * (E - np.conj(op._mirror(E)))
            GA = 0.5 * (G - np.conj(op._mirror(G)))
            direct += GA * np.fft.fftn(f * np.fft.ifftn(EA * F))
    assert np.abs(term - direct).max() <= 1e-13 * np.abs(direct).max()
    p = N // 2
    off = np.ones((N,) * 3, dtype=bool)
    off[p, :, :] = off[:, p, :] = off[:, :, p] = False
    assert np.abs(term[off]).max() == 0.0
    # every plane and the corner carry part of the term on this data
    assert np.abs(term[p, 1, 2]) > 0 and np.abs(term[1, p, 2]) > 0 and np.abs(term[1, 2, p]) > 0
    qc = op.collide(f)
    qs = op.symmetrized().collide(f)
    qfix = qs + (np.fft.ifftn(term) / op.sphr_fac).real
    assert np.abs(qs - qc).max() > 1e-4 * np.abs(qc).max()  # the term matters here
    assert np.abs(qfix - qc).max() <= 1e-14 * np.abs(qc).max()


@pytest.mark.parametrize("N,nr,M,seed", [(8, 3, 8, 4), (32, 2, 5, 5)])
def test_nyquist_term_2d_completes_symmetrized(N, nr, M, seed):
    """The d = 2 (Nyquist-line) term of the FP32 2D 32^2 real-data form: equal to the direct
    sum G_A * fft(f * ifft(E_A fhat)) from the oracle's own tables, zero off the two lines,
    and the symmetrised operator plus it is the reference's C2C operator to round-off on
    data with O(1) Nyquist content (e != e0, hard spheres, odd angle count)."""
    sig, ws = Qd.circle_rule(M)
    op = C.OracleOperator(2, N, 2.5, nr, sig, ws, e=0.45, e0=0.85, gamma=1.0)
    rng = np.random.default_rng(seed)
    f = rng.random((N, N)) + 0.3
    term = op.nyquist_term(f)
    F = np.fft.fft2(f)
    direct = np.zeros_like(term)
    for j in range(nr):
        for m in range(op.M):
            E, G = op.exp[j, m], op.gain[j, m]
            EA = 0.5 * (E - np.conj(op._mirror(E)))
            GA = 0.5 * (G - np.conj(op._mirror(G)))
            direct += GA * np.fft.fft2(f * np.fft.ifft2(EA * F))
    assert np.abs(term - direct).max() <= 1e-13 * np.abs(direct).max()
    p = N // 2
    off = np.ones((N, N), dtype=bool)
    off[p, :] = off[:, p] = False
    assert np.abs(term[off]).max() == 0.0
    assert np.abs(term[p, 1]) > 0 and np.abs(term[1, p]) > 0 and np.abs(term[p, p]) > 0
    qc = op.collide(f)
    qs = op.symmetrized().collide(f)
    qfix = qs + (np.fft.ifft2(term) / op.sphr_fac).real
    assert np.abs(qs - qc).max() > 1e-4 * np.abs(qc).max()
    assert np.abs(qfix - qc).max() <= 1e-14 * np.abs(qc).max()


def test_restitution_profile_values():
    """e(r) = (e0 - e)/2 tanh(r - R/2) + (e0 + e)/2 (inelastic_collisions.py:77-79), pinned by its
    values: the midpoint of the blend at r = R/2, the odd symmetry of tanh about it, and the limits
    e (r << R/2) and e0 (r >> R/2)."""
    from oracle.collision import e_profile
    R, e, e0 = 6.0, 0.35, 0.95
    assert e_profile(np.array([R / 2]), R, e, e0)[0] == 0.5 * (e + e0)
    t = np.array([0.1, 0.7, 2.5])
    np.testing.assert_allclose(e_profile(R / 2 + t, R, e, e0) + e_profile(R / 2 - t, R, e, e0), e + e0,
                               rtol=0, atol=1e-15)
    assert abs(e_profile(np.array([R / 2 + 30.0]), R, e, e0)[0] - e0) < 1e-15
    assert abs(e_profile(np.array([R / 2 - 30.0]), R, e, e0)[0] - e) < 1e-15
    # constant e when e0 = e (the 3D configuration sets both)
    assert np.all(e_profile(np.linspace(0, R, 7), R, 0.6, 0.6) == 0.6)
    # slope at the midpoint: d/dr = (e0 - e)/2
    hstep = 1e-6
    slope = (e_profile(np.array([R / 2 + hstep]), R, e, e0)[0]
             - e_profile(np.array([R / 2 - hstep]), R, e, e0)[0]) / (2 * hstep)
    assert abs(slope - 0.5 * (e0 - e)) < 1e-8


def test_spherical_mean_kernel_values():
    """sp = J0 (2D) and sin x / x (3D) (inelastic_collisions.py:42-47), pinned by known values."""
    from oracle.collision import sp_func
    j0_zero = 2.404825557695773
    assert abs(sp_func(2, np.array([j0_zero]))[0]) < 1e-15
    assert sp_func(2, np.array([0.0]))[0] == 1.0
    assert abs(sp_func(2, np.array([1.0]))[0] - 0.7651976865579666) < 1e-15
    assert sp_func(3, np.array([0.0]))[0] == 1.0
    assert abs(sp_func(3, np.array([np.pi]))[0]) < 1e-16
    assert abs(sp_func(3, np.array([1.0]))[0] - np.sin(1.0)) < 1e-16
