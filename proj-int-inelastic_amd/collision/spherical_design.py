"""Sphere quadrature rules for the 3D operator.

Mirrors reference ``src/collision/spherical_design/__init__.py:7-116``: a rule
is selected by name and number of points; tabulated rules are text files of
points with an optional trailing weight.  The reference ships Womersley
symmetric designs as data files; this package does not redistribute them, so
a tabulated rule is read from ``rule_dir`` (config key ``rule-dir`` or the
environment variable ``INELASTIC_SPHERE_DIR``, same ``<name>-ss<deg>-m<npts>.txt``
naming) or from inline text (config key ``rule``).  The Lebedev-Laikov
octahedral rules (M = 6, 14, 26, 38; degrees 3, 5, 7, 9) are built in
(``ssrule = lebedev``) with their published closed-form nodes and weights.
"""
import itertools
import math
import os
import re

import numpy as np


class SphereRule:
    def __init__(self, pts, wts):
        self.pts = np.asarray(pts, dtype=np.float64)
        self.wts = np.asarray(wts, dtype=np.float64)


def _orbit(vec):
    out = set()
    for p in itertools.permutations(vec):
        for sg in itertools.product((1, -1), repeat=3):
            out.add(tuple(s * x + 0.0 for s, x in zip(sg, p)))
    return sorted(out)


def lebedev(npts):
    """Nodes (M, 3) and weights summing to 1 (Lebedev & Laikov 1999)."""
    a1 = _orbit((1.0, 0.0, 0.0))
    a2 = _orbit((math.sqrt(0.5), math.sqrt(0.5), 0.0))
    a3 = _orbit((1 / math.sqrt(3),) * 3)
    if npts == 6:
        groups = [(a1, 1 / 6)]
    elif npts == 14:
        groups = [(a1, 1 / 15), (a3, 3 / 40)]
    elif npts == 26:
        groups = [(a1, 1 / 21), (a2, 4 / 105), (a3, 9 / 280)]
    elif npts == 38:
        p = 0.4597008433809831
        groups = [(a1, 1 / 105), (a3, 9 / 280), (_orbit((p, math.sqrt(1 - p * p), 0.0)), 1 / 35)]
    else:
        raise ValueError("No suitable spherical quadrature rule found")
    pts = [q for g, _ in groups for q in g]
    wts = [w for g, w in groups for _ in g]
    return SphereRule(pts, wts)


def parse_rule(text, ndim=3):
    """Text format of spherical_design/__init__.py:7-40."""
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
    return SphereRule(pts, wts)


def get_sphrquadrule(sstype, rule=None, npts=None, qdeg=None, rule_dir=None):
    """Select a rule as spherical_design/__init__.py:91-116 does."""
    if sstype != "symmetric":
        raise ValueError("Unknown rule family %r" % sstype)
    if rule == "lebedev":
        return lebedev(int(npts))
    if rule and not re.match(r"[a-zA-z0-9\-~+]+$", rule):
        r = parse_rule(rule)
        if npts and npts != len(r.pts):
            raise ValueError("Invalid number of points in provided rule")
        return r
    if not npts and not qdeg:
        raise ValueError("Must specify either npts or qdeg")
    rule_dir = rule_dir or os.environ.get("INELASTIC_SPHERE_DIR")
    best = None
    if rule_dir and os.path.isdir(rule_dir):
        for path in sorted(os.listdir(rule_dir)):
            m = re.match(r"([a-zA-Z0-9\-~+]+)-ss(\d+)(?:-m(\d+))?\.txt$", path)
            if not m:
                continue
            rname, rqdeg, rnpts = m.group(1), int(m.group(2)), int(m.group(3) or -1)
            if (not rule or rule == rname) and (not npts or npts == rnpts) and (not qdeg or qdeg <= rqdeg):
                if not best or (npts and rqdeg > best[2]) or (qdeg and rnpts < best[1]):
                    best = (path, rnpts, rqdeg)
    if not best:
        raise ValueError("No suitable spherical quadrature rule found")
    with open(os.path.join(rule_dir, best[0])) as fh:
        return parse_rule(fh.read())
