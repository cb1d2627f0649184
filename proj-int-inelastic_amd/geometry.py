"""Cartesian phase-space grid (subset of reference ``src/geometry.py`` used by
Solution/State: Dimension.centers :188-194, Grid.c_centers :70-73/:133-144,
Domain :4-42)."""
import numpy as np


class Dimension(object):
    def __init__(self, lower, upper, num_cells, name="x"):
        self.lower = float(lower)
        self.upper = float(upper)
        self.num_cells = int(num_cells)
        self.name = name
        self._centers = None

    @property
    def delta(self):
        return (self.upper - self.lower) / float(self.num_cells)

    @property
    def centers(self):
        if self._centers is None:
            self._centers = np.empty(self.num_cells)
            for i in range(self.num_cells):
                self._centers[i] = self.lower + (i + 0.5) * self.delta
        return self._centers

    @property
    def nodes(self):
        return np.array([self.lower + i * self.delta for i in range(self.num_cells + 1)])


class Grid(object):
    def __init__(self, dimensions):
        self._dimensions = list(dimensions)
        self._num_dim_x = None
        self._c_centers = None

    def get_dim_attribute(self, attr):
        return [getattr(d, attr) for d in self._dimensions]

    @property
    def dimensions(self):
        return self._dimensions

    @property
    def num_dim(self):
        return len(self._dimensions)

    @property
    def num_dim_x(self):
        return self._num_dim_x

    @num_dim_x.setter
    def num_dim_x(self, v):
        self._num_dim_x = v

    @property
    def num_cells(self):
        return self.get_dim_attribute("num_cells")

    @property
    def lower(self):
        return self.get_dim_attribute("lower")

    @property
    def upper(self):
        return self.get_dim_attribute("upper")

    @property
    def delta(self):
        return self.get_dim_attribute("delta")

    @property
    def centers(self):
        return self.get_dim_attribute("centers")

    @property
    def c_centers(self):
        if self._c_centers is None:
            index = np.indices(self.num_cells)
            self._c_centers = [c[index[i, ...]] for i, c in enumerate(self.centers)]
        return self._c_centers


class Domain(object):
    def __init__(self, lower, upper, n):
        names = ["x", "vx", "vy", "vz"]
        dims = [Dimension(lo, up, nn, name=nm)
                for lo, up, nn, nm in zip(lower, upper, n, names[: len(n) + 1])]
        self.grids = [Grid(dims)]

    @property
    def grid(self):
        return self.grids[0]

    @property
    def num_dim(self):
        return self.grid.num_dim
