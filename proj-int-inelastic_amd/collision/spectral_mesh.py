"""Velocity box and the radial / angular quadrature of the spectral method.

Public surface of reference ``src/collision/spectral_mesh.py:8-132`` (constructor
``SpectralMesh(config)``; ``v_center``, ``v_centers``, ``ncirc_or_nsphr``,
``circ_or_sphr_quad()``, ``ndim``, ``nv``, ``nv_s``, ``L``, ``R``, ``delta``,
``nr``, ``rquad()``) with its numbers:

* box: R = 2 s, L = (3 + sqrt 2) s / 2, cell width 2 L / nv, centres
  -L + (i + 1/2) 2 L / nv (``:46-48``, ``:68-74``, ``:123-125``);
* radial rule: Gauss-Legendre with ``nr`` nodes (default ``int(nv / 2)``,
  ``:43``) mapped affinely from [-1, 1] onto [0, R] (``:28-34``);
* circle (d = 2): phi = ``numpy.arange(0, 2 pi, 2 pi / nphi)`` with the weight
  2 pi / nphi (``:51-56``).  ``arange`` with a rounded step can return one node
  more than ``nphi`` (nphi = 61, 122); the rule keeps the nodes ``arange`` gives
  and ``ncirc_or_nsphr`` reports ``nphi``, as the reference does;
* sphere (d = 3): a named rule of ``nsphr`` points with the equal weight
  4 pi / nsphr (``:58-66``).  New: a rule that carries its own weights
  (Lebedev, built in) exposes them, normalised to 4 pi, by ``sphere_weights()``.

The reference prints the cell count and the box on construction; this class
does not print.
"""
import math

import numpy as np

from .spherical_design import get_sphrquadrule


def _gauss_legendre(n):
    return np.polynomial.legendre.leggauss(n)


# radial rules on [-1, 1] by config name ("quad-rule" of the velocity-mesh section)
RADIAL_RULES = {"legendre": _gauss_legendre}


def box_extent(s):
    """(R, L) of the velocity box for the support parameter s."""
    s = float(s)
    return 2.0 * s, 0.5 * (3.0 + math.sqrt(2)) * s


def radial_rule(name, nr, R):
    """Nodes and weights of the named rule on [0, R]."""
    try:
        make = RADIAL_RULES[name]
    except KeyError:
        raise KeyError("unknown radial quadrature rule %r (known: %s)"
                       % (name, ", ".join(sorted(RADIAL_RULES))))
    x, w = make(int(nr))
    return 0.5 * (x + 1) * R, 0.5 * R * w


def circle_rule(nphi):
    """Unit vectors (K, 2) at phi = arange(0, 2 pi, 2 pi / nphi) and the scalar weight."""
    w = 2 * math.pi / int(nphi)
    phi = np.arange(0, 2 * math.pi, w)
    return np.column_stack((np.cos(phi), np.sin(phi))), w


class SpectralMesh(object):
    """Velocity grid of nv^d cells on [-L, L]^d with the quadratures of Q(f,f)."""

    def __init__(self, config, *args, **kwargs):
        self.config = config
        vm = config["velocity-mesh"]
        self._nv = int(vm["nv"])
        self._nr = int(vm.get("nr", int(self._nv / 2)))
        self._R, self._L = box_extent(vm["s"])
        self._r, self._wr = radial_rule(vm.get("quad-rule", "legendre"), self._nr, self._R)
        self._ndim = int(config["collision-model"]["dim"])
        if self._ndim == 2:
            sect = config["circle-quad-rule"]
            self._nang = int(sect["nphi"])
            self._sigma, self._wsig = circle_rule(self._nang)
            self._own_weights = None
        elif self._ndim == 3:
            sect = config["spherical-design-rule"]
            kind = sect.get("ssrule", "lebedev")
            self._nang = int(sect["nsphr"])
            rule = get_sphrquadrule("symmetric", rule=sect.get("rule", kind), npts=self._nang,
                                    rule_dir=sect.get("rule-dir"))
            self._sigma = rule.pts
            self._wsig = 4 * math.pi / self._nang
            # equal-weight designs use 4 pi / M; a weighted rule keeps its own weights
            self._own_weights = (4 * math.pi * rule.wts / rule.wts.sum()) if kind == "lebedev" else None
        else:
            raise ValueError("Dimension must be 2 or 3.")
        self._centres = None
        self._grids = None

    # --- box -------------------------------------------------------------------------------
    @property
    def ndim(self):
        return self._ndim

    @property
    def nv(self):
        return self._nv

    @property
    def nv_s(self):
        return [self._nv] * self._ndim

    @property
    def L(self):
        return self._L

    @property
    def R(self):
        return self._R

    @property
    def delta(self):
        return 2 * self._L / self._nv

    @property
    def v_center(self):
        """Cell centres along one axis, (nv,)."""
        if self._centres is None:
            self._centres = (np.arange(self._nv) + 0.5) * self.delta - self._L
        return self._centres

    @property
    def v_centers(self):
        """The d coordinate arrays of the grid, each of shape nv_s (axis a varies v_a)."""
        if self._grids is None:
            self._grids = list(np.meshgrid(*([self.v_center] * self._ndim), indexing="ij"))
        return self._grids

    # --- quadratures -----------------------------------------------------------------------
    @property
    def nr(self):
        return self._nr

    def rquad(self):
        return self._r, self._wr

    @property
    def ncirc_or_nsphr(self):
        return self._nang

    def circ_or_sphr_quad(self):
        """(sigma (M, d), scalar weight): 2 pi / nphi on the circle, 4 pi / nsphr on the sphere."""
        return self._sigma, self._wsig

    def sphere_weights(self):
        """Per-node weights (M,): the scalar rule weight, or the rule's own weights."""
        if self._own_weights is not None:
            return self._own_weights
        return np.full(self._sigma.shape[0], self._wsig)
