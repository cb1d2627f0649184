"""CPU checks of the C ABI library: it loads, exports every declared symbol,
and its host-callable special functions match scipy (no GPU needed)."""
import math
import pathlib
import re

import numpy as np
import pytest

ROOT = pathlib.Path(__file__).resolve().parent.parent


def _declared():
    hdr = (ROOT / "include" / "inelastic.h").read_text()
    return sorted(set(re.findall(r"\b(ib_[a-z0-9_]+)\s*\(", hdr)))


def test_header_matches_binding(pkg):
    from proj_int_inelastic_amd import _lib
    names = _declared()
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_library_exports_every_symbol(pkg):
    from proj_int_inelastic_amd import _lib
    lib = _lib.load()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.ib_version().decode().startswith("inelastic-mi355x")


def test_device_special_functions_match_scipy(pkg):
    from scipy import special
    from proj_int_inelastic_amd import _lib
    lib = _lib.load()
    xs = np.concatenate([np.linspace(0, 150, 3001), [1e-8, 2.404825557695773, 64.3]])
    j0 = np.array([lib.ib_host_sp(2, float(x)) for x in xs])
    assert np.abs(j0 - special.jv(0, xs)).max() < 5e-15
    sc = np.array([lib.ib_host_sp(3, float(x)) for x in xs])
    assert np.abs(sc - np.sinc(xs / np.pi)).max() < 1e-15
