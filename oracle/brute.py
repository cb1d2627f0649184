"""Direct evaluation of the discrete gain/loss sums without any FFT library.

Used only to pin ``oracle.collision.OracleOperator.collide`` at tiny N.
The discrete operator of reference ``src/collision/inelastic_collisions.py``
:208-227 written as explicit sums (numpy FFT conventions, forward
unnormalised, inverse 1/n):

  fhat(k)  = sum_x f(x) exp(-2 pi i k.x / N)
  Q+hat(k) = 1/|S| sum_{j,m} G_jm(k) (1/n) sum_l E_jm(l) fhat(l) fhat(k - l)
  Q-(x)    = f(x) sum_j l_j (1/n) sum_k S_j(k) fhat(k) exp(2 pi i k.x / N)

with G, E, S, l from the closed forms of :57-114 evaluated here directly:
nothing is taken from the operator under test.  Its kernel tables are built
from other formulas of the same functions, so a slip in the oracle's tables
cannot cancel: the restitution profile as the logistic blend
e + (e0 - e) / (1 + exp(-2 (r - R/2))) (equal to the tanh form of :77-79),
J0 from its integral (1/pi) int_0^pi cos(x sin t) dt by the trapezoidal rule
(exact to round-off for this periodic integrand at 64 nodes and |x| < 40),
sin(x)/x with its x = 0 limit, |S^1| = 2 pi, |S^2| = 4 pi, and the frequency
of FFT index q as q - N [q >= N/2].  k - l is taken modulo N per axis.
"""
import math

import numpy as np


def _restitution(r, R, e, e0):
    return e + (e0 - e) / (1.0 + np.exp(-2.0 * (np.asarray(r, dtype=np.float64) - 0.5 * R)))


def _j0(x):
    t = (np.arange(64) + 0.5) * (math.pi / 64)   # midpoint rule on [0, pi] = trapezoid of the even extension
    x = np.asarray(x, dtype=np.float64)
    return np.cos(x[..., None] * np.sin(t)).mean(axis=-1)


def _spherical_mean(d, x):
    x = np.asarray(x, dtype=np.float64)
    if d == 2:
        return _j0(x)
    safe = np.where(x == 0, 1.0, x)
    return np.where(x == 0, 1.0, np.sin(safe) / safe)


def _area(d):
    return {2: 2 * math.pi, 3: 4 * math.pi}[d]


def _freq(N):
    q = np.arange(N)
    return np.where(q >= N // 2, q - N, q)


def _dft_mats(N):
    x = np.arange(N)
    k = np.arange(N)  # natural FFT output order
    F = np.exp(-2j * math.pi * np.outer(k, x) / N)
    Fi = np.exp(2j * math.pi * np.outer(x, k) / N) / N
    return F, Fi


def _apply_axes(mat, a, d):
    """Apply a 1D matrix along each of the last d axes of a (shape N^d)."""
    out = a
    for ax in range(d):
        out = np.moveaxis(np.tensordot(mat, np.moveaxis(out, ax, 0), axes=(1, 0)), 0, ax)
    return out


def brute_collide(f, d, N, L, R, r, wr, sigma, wsigma, e, e0, gamma):
    """Return (Q+, Q-) real parts for one cell f of shape N^d."""
    F, Fi = _dft_mats(N)
    fh = _apply_axes(F, f.astype(np.complex128), d)
    kint = _freq(N).astype(np.float64)
    grids = np.meshgrid(*([kint] * d), indexing="ij")
    knorm = np.sqrt(sum(g * g for g in grids))
    S = _area(d)
    n = N ** d
    e_r = _restitution(r, R, e, e0)
    a = math.pi * (1 + e_r) / (4 * L)
    qh = np.zeros((N,) * d, dtype=np.complex128)
    for j in range(len(r)):
        for m in range(sigma.shape[0]):
            sk = sum(sigma[m, ax] * grids[ax] for ax in range(d))
            E = np.exp(-1j * math.pi * r[j] * sk / L)
            G = (S * r[j] ** (gamma + d - 1) * wr[j] * wsigma[m]
                 * np.exp(1j * a[j] * r[j] * sk) * _spherical_mean(d, a[j] * r[j] * knorm))
            A = E * fh
            conv = np.zeros_like(qh)
            # conv(k) = sum_l A(l) fh(k - l): shift fh by l for every l
            for lidx in np.ndindex(*((N,) * d)):
                conv += A[lidx] * np.roll(fh, shift=lidx, axis=tuple(range(d)))
            qh += G * conv / n
    gain = _apply_axes(Fi, qh, d) / S
    lsym = np.zeros((N,) * d)
    for j in range(len(r)):
        lsym += S * r[j] ** (gamma + d - 1) * wr[j] * _spherical_mean(d, math.pi * r[j] * knorm / L)
    loss = f * _apply_axes(Fi, lsym * fh, d)
    return gain.real, loss.real
