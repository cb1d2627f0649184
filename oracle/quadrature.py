"""Quadrature rules of the spectral collision operator (FP64, numpy).

Follows reference ``src/collision/spectral_mesh.py``:
  * radial rule  :28-34  ``leggauss(nr)`` mapped to [0, R]
  * velocity box :36-49  L = (3 + sqrt 2) s / 2, R = 2 s, nr defaults to nv/2
  * circle rule  :51-56  phi = arange(0, 2 pi, 2 pi / M), w = 2 pi / M
  * sphere rule  :58-66  points from a stored symmetric design, w = 4 pi / M
and ``src/collision/spherical_design/__init__.py:7-40`` for the text format of
tabulated sphere rules (one point per line, optional trailing weight).

The Lebedev rules are new (the reference has none); their nodes and weights are
the closed forms published by Lebedev & Laikov (1999).

Pins: Gauss-Legendre is numpy's library routine; circle nodes are closed form;
every sphere rule is checked for polynomial exactness (tests/test_oracle_quadrature.py).
"""
import math

import numpy as np


def velocity_box(s):
    """(L, R) for the reference's box parameter s (spectral_mesh.py:46-48)."""
    s = float(s)
    return 0.5 * (3.0 + math.sqrt(2)) * s, 2.0 * s


def radial_rule(nr, R):
    """Gauss-Legendre nodes/weights on [0, R] (spectral_mesh.py:28-34)."""
    x, w = np.polynomial.legendre.leggauss(int(nr))
    return 0.5 * (x + 1) * R, 0.5 * R * w


def circle_rule(nphi):
    """Uniform circle rule (spectral_mesh.py:51-56, :94-97).

    Returns sigma (M, 2) and a per-node weight array (M,) equal to 2 pi / M.
    ``np.arange`` is used exactly as in the reference, so M values whose step
    rounds badly yield M + 1 nodes, as there.
    """
    wphi = 2 * math.pi / int(nphi)
    phi = np.arange(0, 2 * math.pi, wphi)
    sigma = np.stack((np.cos(phi), np.sin(phi)), axis=-1)
    return sigma, np.full(sigma.shape[0], wphi)


def _perm_sign(vec):
    """All distinct sign changes and permutations of a 3-vector."""
    out = set()
    import itertools
    for p in itertools.permutations(vec):
        for sx in (1, -1):
            for sy in (1, -1):
                for sz in (1, -1):
                    out.add((sx * p[0] + 0.0, sy * p[1] + 0.0, sz * p[2] + 0.0))
    return sorted(out)


def lebedev_rule(npts):
    """Lebedev-Laikov octahedral rules with M in {6, 14, 26, 38}.

    Returns points (M, 3) and weights (M,) normalised to sum to 4 pi.
    Degrees of exactness: 6 -> 3, 14 -> 5, 26 -> 7, 38 -> 9.
    """
    npts = int(npts)
    a1 = _perm_sign((1.0, 0.0, 0.0))
    a2 = _perm_sign((1 / math.sqrt(2), 1 / math.sqrt(2), 0.0))
    a3 = _perm_sign((1 / math.sqrt(3),) * 3)
    if npts == 6:
        groups = [(a1, 1.0 / 6.0)]
    elif npts == 14:
        groups = [(a1, 1.0 / 15.0), (a3, 3.0 / 40.0)]
    elif npts == 26:
        groups = [(a1, 1.0 / 21.0), (a2, 4.0 / 105.0), (a3, 9.0 / 280.0)]
    elif npts == 38:
        p = 0.4597008433809831
        q = math.sqrt(1.0 - p * p)
        c1 = _perm_sign((p, q, 0.0))
        groups = [(a1, 1.0 / 105.0), (a3, 9.0 / 280.0), (c1, 1.0 / 35.0)]
    else:
        raise ValueError("No suitable spherical quadrature rule found")
    pts, wts = [], []
    for g, w in groups:
        pts.extend(g)
        wts.extend([w] * len(g))
    pts = np.array(pts, dtype=np.float64)
    wts = np.array(wts, dtype=np.float64) * 4 * math.pi
    assert pts.shape[0] == npts
    return pts, wts


def parse_tabulated_rule(text, ndim=3):
    """Parse a tabulated sphere rule (spherical_design/__init__.py:7-40).

    Each non-empty line holds ``ndim`` coordinates and optionally a weight.
    Returns (pts (M, ndim), wts (M,) or empty).
    """
    import re
    rule = re.sub(r"(?<=\))\s*,?\s*(?!$)", "\n", text)
    rule = re.sub(r"\(|\)|,", "", rule).strip()
    rule = rule[1:-1] if rule.startswith("[") else rule
    pts, wts = [], []
    for line in rule.splitlines():
        if not line:
            continue
        args = [float(v) for v in line.split()]
        if len(args) == ndim:
            pts.append(args)
        elif len(args) == ndim + 1:
            pts.append(args[:-1])
            wts.append(args[-1])
        else:
            raise ValueError("Invalid points in quadrature rule")
    if len(wts) and len(wts) != len(pts):
        raise ValueError("Invalid number of weights")
    return np.array(pts, dtype=np.float64), np.array(wts, dtype=np.float64)
