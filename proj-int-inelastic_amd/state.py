"""Solution state with the distribution function resident on the HIP device.

Mirrors reference ``src/state.py``: ``f`` (here a float64 torch CUDA tensor;
assigning a numpy array copies it to the device), ``rho``/``u``/``T``
(:31-52, 2D velocity as there; computed with the FP64 moment kernel),
``sum_f`` (:91-92), ``problem_data`` and ``t``.
"""
import numpy as np

from . import _lib
from . import device as dev
from .geometry import Domain, Grid


class State(object):
    def __getattr__(self, key):
        if key in ("num_dim", "num_dim_x", "c_centers", "num_cells", "lower", "upper", "delta",
                   "centers"):
            return getattr(self.grid, key)
        raise AttributeError("'State' object has no attribute '" + key + "'")

    def __init__(self, geom):
        if isinstance(geom, Grid):
            self.grid = geom
        elif isinstance(geom, Domain):
            self.grid = geom.grids[0]
        else:
            raise Exception("Must be initialzed with a Grid object.")
        self.problem_data = {}
        self.t = 0.0
        self._f = None

    @property
    def f(self):
        if self._f is None:
            t = dev.require_gpu()
            self._f = t.empty(tuple(self.grid.num_cells), dtype=t.float64, device="cuda")
        return self._f

    @f.setter
    def f(self, value):
        if dev.is_device_tensor(value) and value.dtype == dev.torch().float64 and value.is_contiguous():
            self._f = value
        else:
            self._f = dev.to_device(value)

    def get_f_global(self):
        return self.f.detach().cpu().numpy()

    def _ghost_dims(self, num_ghost, fbc):
        nd = self.grid.num_dim_x or 0
        if nd not in (1, 2):
            raise Exception("Assumption (1 <= num_dim <= 2) violated.")
        f = self.f
        n0 = int(f.shape[0])
        n1 = int(f.shape[1]) if nd == 2 else 1
        inner = int(np.prod(f.shape[nd:], dtype=np.int64))
        want = (n0 + 2 * num_ghost,) + ((n1 + 2 * num_ghost,) if nd == 2 else ()) + tuple(f.shape[nd:])
        if tuple(fbc.shape) != want or not dev.is_device_tensor(fbc) or not fbc.is_contiguous():
            raise ValueError("fbc must be a contiguous float64 device array of shape %r" % (want,))
        return nd, n0, n1, inner

    # -- state.py:109-122
    def get_fbc_from_f(self, num_ghost, fbc):
        """Copy f into the interior of the ghost-extended array fbc (a device tensor) and return it."""
        nd, n0, n1, inner = self._ghost_dims(num_ghost, fbc)
        _lib.check(_lib.load().ib_ghost_embed(self.f.data_ptr(), fbc.data_ptr(), nd, n0, n1, inner,
                                              int(num_ghost), 0, dev.stream_handle()), "ib_ghost_embed")
        return fbc

    # -- state.py:95-107
    def set_f_from_fbc(self, num_ghost, fbc):
        """Set f from the interior of the ghost-extended array fbc."""
        nd, n0, n1, inner = self._ghost_dims(num_ghost, fbc)
        _lib.check(_lib.load().ib_ghost_embed(self.f.data_ptr(), fbc.data_ptr(), nd, n0, n1, inner,
                                              int(num_ghost), 1, dev.stream_handle()), "ib_ghost_embed")

    def _vel_moments(self):
        nx = self.grid.num_dim_x or 0
        dv_dims = self.grid.dimensions[nx:]
        d = len(dv_dims)
        n = dv_dims[0].num_cells
        L = dv_dims[0].upper
        f = self.f
        nb = int(np.prod(f.shape[:-d], dtype=np.int64)) if f.dim() > d else 1
        t = dev.torch()
        out = t.empty((nb, d + 2), dtype=t.float64, device="cuda")
        _lib.check(_lib.load().ib_moments(f.data_ptr(), out.data_ptr(), nb, d, n, L,
                                          dev.stream_handle()), "ib_moments")
        return out.reshape(tuple(f.shape[:-d]) + (d + 2,)), d

    def sum_f(self, f):
        return f.sum(dim=(-1, -2)) * self.delta[-1] * self.delta[-2]

    @property
    def rho(self):
        m, _ = self._vel_moments()
        return m[..., 0]

    @property
    def u(self):
        m, d = self._vel_moments()
        return [m[..., 1 + a] / m[..., 0] for a in range(d)]

    @property
    def T(self):
        """E / rho - |u|^2 / 2 (reference state.py:43-52, kinetic T for 2D velocity)."""
        m, d = self._vel_moments()
        rho = m[..., 0]
        us = 0.0
        for a in range(d):
            ua = m[..., 1 + a] / rho
            us = us + ua ** 2
        return m[..., d + 1] / rho - 0.5 * us
