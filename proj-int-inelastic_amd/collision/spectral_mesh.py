"""Velocity box and quadrature of the spectral method.

Same constructor, properties and quirks as reference
``src/collision/spectral_mesh.py:8-132``: L = (3 + sqrt 2) s / 2, R = 2 s,
nr defaults to nv / 2, Gauss-Legendre radial rule on [0, R], circle nodes
``arange(0, 2 pi, 2 pi / nphi)`` with weight 2 pi / nphi, sphere weight 4 pi / M
for equal-weight designs.  New: a rule with its own weights (Lebedev) keeps
them (``sphere_weights``), normalised to 4 pi.
"""
import math

import numpy as np

from .spherical_design import get_sphrquadrule


class SpectralMesh(object):
    def __init__(self, config, *args, **kwargs):
        self.config = config
        self._construct_velocity_mesh()
        self._load_radial_quadrature()
        self._ndim = int(self.config["collision-model"]["dim"])
        if self._ndim == 2:
            self._construct_polar_mesh()
        elif self._ndim == 3:
            self._construct_spherical_mesh()
        else:
            raise ValueError("Dimension must be 2 or 3.")
        self._v_center = None
        self._v_centers = None

    def _load_radial_quadrature(self):
        vm_sect = self.config["velocity-mesh"]
        quad_rule = vm_sect.get("quad-rule", "legendre")
        v_quads = {"legendre": np.polynomial.legendre.leggauss}
        r, wr = v_quads[quad_rule](self._nr)
        self._r = 0.5 * (r + 1) * self._R
        self._wr = 0.5 * self._R * wr

    def _construct_velocity_mesh(self):
        vm_sect = self.config["velocity-mesh"]
        self._nv = int(vm_sect["nv"])
        self._nr = int(vm_sect.get("nr", int(self._nv / 2)))
        self._S = float(vm_sect["s"])
        self._R = 2 * self._S
        self._L = 0.5 * (3.0 + math.sqrt(2)) * self._S

    def _construct_polar_mesh(self):
        circ_quad = self.config["circle-quad-rule"]
        self._nphi = int(circ_quad["nphi"])
        self._wphi = 2 * math.pi / self._nphi
        self._phi = np.arange(0, 2 * math.pi, self._wphi)

    def _construct_spherical_mesh(self):
        sphr_sect = self.config["spherical-design-rule"]
        self._ssrule = sphr_sect.get("ssrule", "lebedev")
        self._nsphr = int(sphr_sect["nsphr"])
        rule = sphr_sect.get("rule", self._ssrule)
        srule = get_sphrquadrule("symmetric", rule=rule, npts=self._nsphr,
                                 rule_dir=sphr_sect.get("rule-dir"))
        self._spts = srule.pts
        self._wspts = 4 * math.pi / self._nsphr
        if self._ssrule == "lebedev":
            self._wspts_arr = 4 * math.pi * srule.wts / srule.wts.sum()
        else:
            self._wspts_arr = None

    @property
    def v_center(self):
        if self._v_center is None:
            self._v_center = np.empty(self.nv)
            for i in range(self.nv):
                self._v_center[i] = -self.L + (i + 0.5) * self.delta
        return self._v_center

    @property
    def v_centers(self):
        if self._v_centers is None:
            index = np.indices(self.nv_s)
            self._v_centers = [self.v_center[index[i, ...]] for i in range(self.ndim)]
        return self._v_centers

    @property
    def ncirc_or_nsphr(self):
        if self.ndim == 2:
            return len(self._phi)
        elif self.ndim == 3:
            return self._nsphr
        raise ValueError("Dimension must be 2 or 3.")

    def circ_or_sphr_quad(self):
        if self.ndim == 2:
            sigma = np.stack((np.cos(self._phi), np.sin(self._phi)), axis=-1)
            return sigma, self._wphi
        elif self.ndim == 3:
            return self._spts, self._wspts
        raise ValueError("Dimension must be 2 or 3.")

    def sphere_weights(self):
        """Per-node weights (M,): the scalar rule weight, or the rule's own weights."""
        sigma, w = self.circ_or_sphr_quad()
        if self.ndim == 3 and self._wspts_arr is not None:
            return self._wspts_arr
        return np.full(sigma.shape[0], w)

    @property
    def ndim(self):
        return self._ndim

    @property
    def nv(self):
        return self._nv

    @property
    def nv_s(self):
        return [self.nv] * self.ndim

    @property
    def L(self):
        return self._L

    @property
    def R(self):
        return self._R

    @property
    def delta(self):
        return 2 * self.L / self.nv

    @property
    def nr(self):
        return self._nr

    def rquad(self):
        return self._r, self._wr
