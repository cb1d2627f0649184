"""Exactness pins of the quadrature rules (oracle/quadrature.py) — CPU only."""
import itertools
import math

import numpy as np
import pytest

from oracle import quadrature as Qd


def _sphere_monomial(a, b, c):
    """Exact integral of x^a y^b z^c over the unit sphere (closed form)."""
    if a % 2 or b % 2 or c % 2:
        return 0.0
    g = math.gamma
    return 2 * g((a + 1) / 2) * g((b + 1) / 2) * g((c + 1) / 2) / g((a + b + c + 3) / 2)


@pytest.mark.parametrize("npts,deg", [(6, 3), (14, 5), (26, 7), (38, 9)])
def test_lebedev_exactness(npts, deg):
    pts, w = Qd.lebedev_rule(npts)
    assert np.allclose((pts ** 2).sum(1), 1.0, atol=1e-15)
    assert abs(w.sum() - 4 * math.pi) < 1e-13
    worst = 0.0
    for a, b, c in itertools.product(range(deg + 1), repeat=3):
        if a + b + c > deg:
            continue
        q = (w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c).sum()
        worst = max(worst, abs(q - _sphere_monomial(a, b, c)))
    assert worst < 1e-14
    # not exact one even degree higher
    a = deg + 1
    q = (w * pts[:, 0] ** a).sum()
    assert abs(q - _sphere_monomial(a, 0, 0)) > 1e-6


def test_gauss_legendre_radial():
    R = 5.0
    r, w = Qd.radial_rule(8, R)
    for p in range(16):
        assert abs((w * r ** p).sum() - R ** (p + 1) / (p + 1)) < 1e-12 * R ** (p + 1)


def test_circle_rule_and_arange_quirk():
    s, w = Qd.circle_rule(8)
    assert s.shape == (8, 2) and np.allclose(w, 2 * math.pi / 8)
    assert np.allclose(s[2], [0, 1], atol=1e-15)
    # np.arange yields M + 1 nodes for M = 61 (reference quirk, kept)
    assert Qd.circle_rule(61)[0].shape[0] == 62


def test_tabulated_parser():
    txt = "  1.0 0.0 0.0\n -1.0 0.0 0.0\n 0 1 0\n 0 -1 0\n 0 0 1\n 0 0 -1\n"
    pts, wts = Qd.parse_tabulated_rule(txt)
    assert pts.shape == (6, 3) and wts.size == 0
    pts, wts = Qd.parse_tabulated_rule("(1, 0, 0, 0.5), (0, 1, 0, 0.5)")
    assert pts.shape == (2, 3) and np.allclose(wts, 0.5)
    with pytest.raises(ValueError):
        Qd.parse_tabulated_rule("1 2\n")


def test_velocity_box():
    L, R = Qd.velocity_box(3.624654714576)
    assert abs(L - 8.0) < 1e-11 and abs(R - 7.249309429152) < 1e-12
