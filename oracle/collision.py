"""FP64 restatement of the fast spectral inelastic VHS collision operator.

Follows reference ``src/collision/inelastic_collisions.py``:
  * load_parameters          :30-52   e, e0 (default 1), gamma, sp = J0 / sinc,
                                      |S^{d-1}| = 2 pi^{d/2} / Gamma(d/2)
  * perform_precomputation   :57-137  k = fftshift(arange(-N/2, N/2)), |k|,
                                      Laplacian symbol, tanh e(r) profile,
                                      gain / translate / loss kernels
  * collide                  :208-227 C2C FFTs over the last d axes, gain is
                                      divided by |S|, ``.real`` at the end,
                                      gain only unless ``with_loss``
  * laplacian                :229-230
and ``src/collision/base.py:32-58`` for ``__call__`` (gain only + heat bath).

Additions the reference lacks (new, FP64): per-node sphere weights (Lebedev),
moments for any d, and the L2 conservative projection of Q onto
{sum Q = 0, sum Q v_a = 0}.

Pins (tests/test_oracle_collision.py): brute-force discrete convolution
(oracle/brute.py, explicit DFT sums) at N = 8; exact invariants (quadratic
scaling, periodic shift equivariance, batch invariance); the elastic
Maxwell BKW closed form (2D, B = 1/(2 pi)); the Haff energy identity of
inelastic Maxwell molecules dT/dt = -(1 - e^2) rho T / 4; mass and momentum
conservation of gain - loss to quadrature accuracy.
"""
import math

import numpy as np

try:
    import scipy.fft as _sfft
    from scipy import special as _special
except ImportError:  # pragma: no cover - scipy is part of the image
    _sfft = None
    _special = None

from . import quadrature


def sphere_area(d):
    """|S^{d-1}| (inelastic_collisions.py:50-52)."""
    return 2 * math.pi ** (0.5 * d) / math.gamma(0.5 * d)


def sp_func(d, x):
    """Spherical mean kernel: J0 in 2D, sin(x)/x in 3D (:42-47)."""
    if d == 2:
        return _special.jv(0, x)
    if d == 3:
        return np.sinc(x / np.pi)
    raise ValueError("Dimension must be 2 or 3.")


def e_profile(r, R, e, e0):
    """e_j = (e0 - e)/2 tanh(r_j - R/2) + (e0 + e)/2 (:77-79)."""
    return 0.5 * (e0 - e) * np.tanh(r - 0.5 * R) + 0.5 * (e0 + e)


def spectral_index(N):
    """FFT-ordered integer frequencies, Nyquist = -N/2 (:62)."""
    return np.fft.fftshift(np.arange(-int(N / 2), int(N / 2)))


def index_norm(N, d):
    """|k| on the d-dimensional FFT grid (:64-67)."""
    idx = spectral_index(N)
    sq = idx ** 2
    acc = 0
    for i in range(d):
        acc = acc + sq[(...,) + (None,) * i]
    return np.sqrt(acc)


def _fftn(x, d, workers):
    axes = tuple(range(-d, 0))
    if _sfft is not None:
        return _sfft.fftn(x, axes=axes, workers=workers)
    return np.fft.fftn(x, axes=axes)


def _ifftn(x, d, workers):
    axes = tuple(range(-d, 0))
    if _sfft is not None:
        return _sfft.ifftn(x, axes=axes, workers=workers)
    return np.fft.ifftn(x, axes=axes)


class OracleOperator:
    """One collision operator: (d, N, L, R, radial rule, sphere rule, e, e0, gamma).

    ``sigma``/``wsigma`` are the sphere nodes (M, d) and per-node weights (M,).
    The reference uses one scalar weight for all nodes; passing a constant
    array reproduces it exactly.
    """

    def __init__(self, d, N, s, nr, sigma, wsigma, e=1.0, e0=1.0, gamma=0.0):
        self.d = int(d)
        self.N = int(N)
        self.L, self.R = quadrature.velocity_box(s)
        self.r, self.wr = quadrature.radial_rule(nr, self.R)
        self.nr = int(nr)
        self.sigma = np.asarray(sigma, dtype=np.float64)
        self.wsigma = np.broadcast_to(
            np.asarray(wsigma, dtype=np.float64), (self.sigma.shape[0],)
        ).copy()
        self.M = self.sigma.shape[0]
        self.e_in, self.e0, self.gamma = float(e), float(e0), float(gamma)
        self.sphr_fac = sphere_area(self.d)
        self._precompute()

    # -- inelastic_collisions.py:54-55
    def _pre_fac(self, r):
        return self.sphr_fac * r ** (self.gamma + self.d - 1)

    # -- inelastic_collisions.py:57-122
    def _precompute(self):
        d, N, L = self.d, self.N, self.L
        idx = spectral_index(N)
        idx_norm = index_norm(N, d)
        self.idx_norm = idx_norm
        self.lapl = -math.pi ** 2 / L ** 2 * idx_norm ** 2
        r, wr = self.r, self.wr
        quad_slice = (...,) + (None,) * d
        self.e_r = e_profile(r, self.R, self.e_in, self.e0)
        self.exp_fac = 0.25 * math.pi * (1 + self.e_r) / L
        idx_mesh = 0
        for i in range(d):
            idx_mesh = (
                np.expand_dims(idx_mesh, axis=-1)
                + idx * self.sigma[(slice(None), slice(i, i + 1)) + (None,) * i]
            )
        r_gain = r[quad_slice + (None,)]
        wr_gain = wr[quad_slice + (None,)]
        exp_fac = self.exp_fac[quad_slice + (None,)]
        idx_mesh = idx_mesh * r_gain
        gain = (
            self._pre_fac(r_gain)
            * np.exp(1j * exp_fac * idx_mesh)
            * sp_func(d, exp_fac * r_gain * idx_norm)
        )
        gain *= wr_gain
        gain *= self.wsigma[(slice(None),) + (None,) * d]
        self.gain = gain  # (Nr, M, N^d) complex128
        self.exp = np.exp(-1j * math.pi * idx_mesh / L)  # (Nr, M, N^d)
        r_loss = r[quad_slice]
        wr_loss = wr[quad_slice]
        loss = self._pre_fac(r_loss)
        loss *= wr_loss
        self.loss = loss  # (Nr, 1^d)
        self.sp = sp_func(d, math.pi * r_loss * idx_norm / L)  # (Nr, N^d)

    def _mirror(self, x):
        """x(-k) on the FFT grid of the last d axes."""
        d = self.d
        out = x
        for ax in range(-d, 0):
            out = np.roll(np.flip(out, axis=ax), 1, axis=ax)
        return out

    def symmetrized(self):
        """Operator with E and G replaced by their Hermitian parts (X + conj X(-k))/2.

        Not the reference algorithm: this is the operator the real-data FP32 form
        computes.  It equals the reference's C2C operator exactly when fhat
        vanishes on the Nyquist planes, and differs by O(|fhat on the Nyquist
        planes|) otherwise.
        """
        import copy
        op = copy.copy(self)
        op.exp = 0.5 * (self.exp + np.conj(self._mirror(self.exp)))
        op.gain = 0.5 * (self.gain + np.conj(self._mirror(self.gain)))
        return op

    def nyquist_term(self, f):
        """Gain spectrum G_A * fft(f * ifft(E_A fhat)) summed over (j, m), for d = 2 or 3.

        E_A, G_A are the anti-Hermitian parts of the translate and gain kernels; they
        vanish off the three Nyquist planes.  ``symmetrized()`` plus this term is
        the reference's C2C gain (``Re ifftn(term) / |S|`` added to the symmetrised
        gain).  Restated through the factorisation the device kernel uses:
        ifft(E_A fhat) = i beta with beta a sum of three plane waves (-1)^x P1(y,z),
        (-1)^y P2(x,z), (-1)^z P3(x,y) (2D inverse transforms of the phase-weighted
        Nyquist planes of fhat) and a corner term; fft(f b) on the plane kx = N/2 is
        i fft2(sum_x (-1)^x f beta), and likewise on ky = N/2, kz = N/2.
        Returns the (N, N, N) complex spectrum (zero off the planes), in the units of
        ``self.gain`` (|S| included).  For d = 2 the planes are the lines kx = N/2 and
        ky = N/2: ifft(E_A fhat) = i beta with beta = s_x (-1)^x P1(y) + s_y (-1)^y P2(x)
        (1D inverse transforms of the phase-weighted Nyquist lines; no corner term), and
        fft(f b) on kx = N/2 is i fft(sum_x (-1)^x f beta), on ky = N/2 i fft(sum_y ...).
        """
        if self.d == 2:
            return self._nyquist_term_2d(f)
        if self.d != 3:
            raise ValueError("the Nyquist term is defined here for d = 2 and 3")
        N, L = self.N, self.L
        p = N // 2
        f = np.asarray(f, dtype=np.float64)
        F = np.fft.fftn(f) / f.size
        kap = spectral_index(N).astype(np.float64)
        sgn = (-1.0) ** np.arange(N)
        sx3 = sgn[:, None, None] * sgn[None, :, None] * sgn[None, None, :]
        out = np.zeros((N, N, N), dtype=np.complex128)
        notp = np.arange(N) != p
        for j in range(self.nr):
            rj, aj = self.r[j], self.exp_fac[j]
            acc0 = np.zeros((N, N), dtype=np.complex128)
            acc1 = np.zeros((N, N), dtype=np.complex128)
            acc2 = np.zeros((N, N), dtype=np.complex128)
            for m in range(self.M):
                sg = self.sigma[m]
                e = [np.exp(-1j * math.pi * rj * sg[a] * kap / L) for a in range(3)]
                g = [np.exp(1j * aj * rj * sg[a] * kap) for a in range(3)]
                h = [x.copy() for x in e]
                for a in range(3):
                    h[a][p] = e[a][p].real
                s = [e[a][p].imag for a in range(3)]
                P1 = (np.fft.ifft2(h[1][:, None] * h[2][None, :] * F[p, :, :]) * N * N).real
                P2 = (np.fft.ifft2(h[0][:, None] * h[2][None, :] * F[:, p, :]) * N * N).real
                P3 = (np.fft.ifft2(h[0][:, None] * h[1][None, :] * F[:, :, p]) * N * N).real
                beta = (s[0] * sgn[:, None, None] * P1[None, :, :]
                        + s[1] * sgn[None, :, None] * P2[:, None, :]
                        + s[2] * sgn[None, None, :] * P3[:, :, None]
                        - s[0] * s[1] * s[2] * F[p, p, p].real * sx3)
                u = f * beta
                V1 = np.fft.fft2(np.tensordot(sgn, u, axes=(0, 0)))       # (ky, kz) on kx = p
                V2 = np.fft.fft2(np.einsum("y,xyz->xz", sgn, u))         # (kx, kz) on ky = p
                V3 = np.fft.fft2(np.tensordot(u, sgn, axes=(2, 0)))       # (kx, ky) on kz = p
                # weight w_sigma Im(prod over Nyquist axes of g(p)) * prod over the others of g(k)
                nyq = [np.where(notp, 1.0, g[a]) for a in range(3)]
                rest = [np.where(notp, g[a], 1.0) for a in range(3)]
                w = self.wsigma[m]
                acc0 += w * (g[0][p] * nyq[1][:, None] * nyq[2][None, :]).imag \
                    * rest[1][:, None] * rest[2][None, :] * V1
                acc1 += w * (g[1][p] * nyq[2][None, :] * np.ones((N, 1))).imag \
                    * rest[0][:, None] * rest[2][None, :] * V2
                acc2 += w * g[2][p].imag * rest[0][:, None] * rest[1][None, :] * V3
            crad = self._pre_fac(rj) * self.wr[j]
            sp = sp_func(3, aj * rj * self.idx_norm)
            out[p, :, :] += -crad * sp[p, :, :] * acc0
            t1 = -crad * sp[:, p, :] * acc1
            t1[p, :] = 0.0                       # kx = p points belong to the first plane
            out[:, p, :] += t1
            t2 = -crad * sp[:, :, p] * acc2
            t2[p, :] = 0.0
            t2[:, p] = 0.0
            out[:, :, p] += t2
        return out

    def _nyquist_term_2d(self, f):
        """d = 2 case of ``nyquist_term`` (the factorisation k_f2_nyq evaluates)."""
        N, L = self.N, self.L
        p = N // 2
        f = np.asarray(f, dtype=np.float64)
        F = np.fft.fft2(f) / f.size
        kap = spectral_index(N).astype(np.float64)
        sgn = (-1.0) ** np.arange(N)
        out = np.zeros((N, N), dtype=np.complex128)
        notp = np.arange(N) != p
        for j in range(self.nr):
            rj, aj = self.r[j], self.exp_fac[j]
            acc0 = np.zeros(N, dtype=np.complex128)
            acc1 = np.zeros(N, dtype=np.complex128)
            for m in range(self.M):
                sg = self.sigma[m]
                e = [np.exp(-1j * math.pi * rj * sg[a] * kap / L) for a in range(2)]
                g = [np.exp(1j * aj * rj * sg[a] * kap) for a in range(2)]
                h = [x.copy() for x in e]
                for a in range(2):
                    h[a][p] = e[a][p].real
                s = [e[a][p].imag for a in range(2)]
                P1 = (np.fft.ifft(h[1] * F[p, :]) * N).real
                P2 = (np.fft.ifft(h[0] * F[:, p]) * N).real
                beta = s[0] * sgn[:, None] * P1[None, :] + s[1] * sgn[None, :] * P2[:, None]
                u = f * beta
                V1 = np.fft.fft(sgn @ u)   # (ky) on kx = p
                V2 = np.fft.fft(u @ sgn)   # (kx) on ky = p
                w = self.wsigma[m]
                nyq1 = np.where(notp, 1.0, g[1])
                rest1 = np.where(notp, g[1], 1.0)
                acc0 += w * (g[0][p] * nyq1).imag * rest1 * V1
                acc1 += w * g[1][p].imag * np.where(notp, g[0], 1.0) * V2
            crad = self._pre_fac(rj) * self.wr[j]
            sp = sp_func(2, aj * rj * self.idx_norm)
            out[p, :] += -crad * sp[p, :] * acc0
            t1 = -crad * sp[:, p] * acc1
            t1[p] = 0.0                          # the kx = p point belongs to the first line
            out[:, p] += t1
        return out

    # -- inelastic_collisions.py:208-227 (looped over r for bounded memory)
    def collide_parts(self, f, workers=1):
        """Return (gain_sum_hat_ifft, loss_sum) before the final scaling.

        gain = ifftn(sum_{j,m} G * fftn(ifftn(E * fhat) * f)); loss = sum_j
        ifftn(S_j fhat) * l_j * f.  Both complex, shape f.shape.
        """
        d = self.d
        f = np.asarray(f, dtype=np.float64)
        nb = f.ndim - d
        bslice = (slice(None),) + (None,) * nb
        f_hat = _fftn(f, d, workers)
        gain_hat = np.zeros(f.shape, dtype=np.complex128)
        loss = np.zeros(f.shape, dtype=np.complex128)
        for j in range(self.nr):
            E = self.exp[j][bslice]  # (M, 1^nb, N^d)
            G = self.gain[j][bslice]
            t = _fftn(_ifftn(E * f_hat, d, workers) * f, d, workers)
            t *= G
            gain_hat += t.sum(axis=0)
            lj = _ifftn(self.sp[j] * f_hat, d, workers)
            lj *= self.loss[j] * f
            loss += lj
        gain = _ifftn(gain_hat, d, workers)
        return gain, loss

    def collide(self, f, with_loss=False, workers=1):
        gain, loss = self.collide_parts(f, workers)
        if with_loss:
            return (gain / self.sphr_fac - loss).real
        return (gain / self.sphr_fac).real

    def gain_loss(self, f, workers=1):
        """(Q+, Q-) real parts, for diagnostics and FP32 error reports."""
        gain, loss = self.collide_parts(f, workers)
        return (gain / self.sphr_fac).real, loss.real

    # -- inelastic_collisions.py:229-230
    def laplacian(self, f, workers=1):
        d = self.d
        return _ifftn(self.lapl * _fftn(f, d, workers), d, workers).real

    # -- base.py:32-42 (device-independent part)
    def __call__(self, f, heat_bath=None, workers=1):
        out = self.collide(f, workers=workers)
        if heat_bath:
            out += heat_bath * self.laplacian(f, workers)
        return out

    # ---- velocity grid and moments (state.py:31-52 generalised to any d)
    @property
    def delta(self):
        return 2 * self.L / self.N

    def v_center(self):
        return np.array([-self.L + (i + 0.5) * self.delta for i in range(self.N)])

    def v_grids(self):
        """List of d arrays of shape N^d: v_a on the grid, axis a <-> v_a."""
        v = self.v_center()
        return [v[np.indices((self.N,) * self.d)[a]] for a in range(self.d)]


def moments(f, v_grids, dv):
    """(rho, momentum (d,), energy) per cell, FP64; f shape (*B, N^d).

    rho = sum f dv^d, m_a = sum f v_a dv^d, E = 1/2 sum f |v|^2 dv^d.
    """
    d = len(v_grids)
    axes = tuple(range(-d, 0))
    w = dv ** d
    rho = f.sum(axis=axes) * w
    mom = np.stack([(f * v).sum(axis=axes) * w for v in v_grids], axis=-1)
    v2 = sum(v * v for v in v_grids)
    en = 0.5 * (f * v2).sum(axis=axes) * w
    return rho, mom, en


def temperature(f, v_grids, dv):
    """Kinetic temperature T = (2E/rho - |u|^2)/d (state.py:43-52 for d=2)."""
    d = len(v_grids)
    rho, mom, en = moments(f, v_grids, dv)
    u = mom / rho[..., None]
    return (2 * en / rho - (u * u).sum(axis=-1)) / d


def project_conservative(q, v_grids):
    """L2 projection of Q onto {sum Q = 0, sum Q v_a = 0} (new, FP64).

    On the symmetric velocity grid the Gram matrix of the constraint rows
    {1, v_1..v_d} is diagonal, so Q_c = Q - mean(Q) - sum_a <Q,v_a>/<v_a,v_a> v_a.
    """
    d = len(v_grids)
    axes = tuple(range(-d, 0))
    n = v_grids[0].size
    bshape = q.shape[: q.ndim - d] + (1,) * d
    out = q - (q.sum(axis=axes) / n).reshape(bshape)
    for v in v_grids:
        c = (q * v).sum(axis=axes) / (v * v).sum()
        out = out - c.reshape(bshape) * v
    return out


def operator_from_config(config):
    """Build an OracleOperator from a reference-style ini mapping.

    Sections/keys as in reference ``vhs-2d.ini``/``vhs-3d.ini`` and
    ``spectral_mesh.py``; 3D rules: ``ssrule = lebedev`` (new) or a
    tabulated rule passed as text under ``spherical-design-rule/rule``.
    """
    cm = config["collision-model"]
    vm = config["velocity-mesh"]
    d = int(cm["dim"])
    N = int(vm["nv"])
    nr = int(vm.get("nr", int(N / 2)))
    s = float(vm["s"])
    if d == 2:
        sigma, ws = quadrature.circle_rule(int(config["circle-quad-rule"]["nphi"]))
    elif d == 3:
        sect = config["spherical-design-rule"]
        npts = int(sect["nsphr"])
        if sect.get("ssrule", "lebedev") == "lebedev":
            sigma, ws = quadrature.lebedev_rule(npts)
        else:
            sigma, _ = quadrature.parse_tabulated_rule(sect["rule"])
            ws = np.full(sigma.shape[0], 4 * math.pi / npts)
    else:
        raise ValueError("Dimension must be 2 or 3.")
    return OracleOperator(
        d, N, s, nr, sigma, ws,
        e=float(cm.get("e", 1.0)), e0=float(cm.get("e0", 1.0)),
        gamma=float(cm.get("gamma", 0.0)),
    )
