"""Solver base (reference ``src/solver.py``): BC enum, fixed-dt
``evolve_to_time`` (:107-143), and the ghost-cell conventions (:51-105).  The
transport step's stencil kernel applies the ghost cells on the fly and never
builds a ghost array; ``_allocate_bc_arrays`` / ``_apply_bcs`` / ``_bc_lower`` /
``_bc_upper`` build ``fbc`` on the device for callers that use it (HIP kernels
ib_ghost_embed and ib_ghost_fill)."""
import numpy as np

from . import _lib
from . import device as dev


class BC:
    extrap = 1
    periodic = 2


class Solver(object):
    @property
    def all_bcs(self):
        return self.bc_lower, self.bc_upper

    @all_bcs.setter
    def all_bcs(self, all_bcs):
        for i in range(self.num_dim_x):
            self.bc_lower[i] = all_bcs
            self.bc_upper[i] = all_bcs

    def __init__(self, riemann_solver=None, collision_operator=None):
        self.dt_initial = 0.01
        self.dt_max = 1e99
        self.max_steps = 10000
        self.dt_variable = False
        self.fbc = None
        if riemann_solver is not None:
            self.rp = riemann_solver
        self.collision = collision_operator
        self.before_step = None
        self.dt = self.dt_initial
        self.bc_lower = [None] * (self.num_dim_x)
        self.bc_upper = [None] * (self.num_dim_x)
        self._is_set_up = False
        super(Solver, self).__init__()

    def setup(self, solution):
        self._is_set_up = True

    # -- solver.py:51-56
    def _allocate_bc_arrays(self, state):
        t = dev.require_gpu()
        nd = self.num_dim_x
        shape = [n + 2 * self.num_ghost for n in state.f.shape[:nd]] + list(state.f.shape[nd:])
        self.fbc = t.zeros(shape, dtype=t.float64, device="cuda")
        self._apply_bcs(state)

    # -- solver.py:57-65
    def _apply_bcs(self, state):
        self.fbc = state.get_fbc_from_f(self.num_ghost, self.fbc)
        for idim in range(self.num_dim_x):
            self._bc_lower(self.bc_lower[idim], self.fbc, axis=idim)
            self._bc_upper(self.bc_upper[idim], self.fbc, axis=idim)

    def _ghost_call(self, array, axis, lo, hi):
        g = self.num_ghost
        shape = tuple(array.shape)
        if not (dev.is_device_tensor(array) and array.is_contiguous()):
            raise ValueError("the ghost-extended array must be a contiguous device tensor")
        outer = int(np.prod(shape[:axis], dtype=np.int64))
        inner = int(np.prod(shape[axis + 1:], dtype=np.int64))
        _lib.check(_lib.load().ib_ghost_fill(array.data_ptr(), outer, shape[axis] - 2 * g, inner, g, lo, hi,
                                             dev.stream_handle()), "ib_ghost_fill")

    @staticmethod
    def _bc_code(bc_type, side):
        if bc_type in (BC.extrap, BC.periodic):
            return bc_type
        if bc_type is None:
            raise Exception("%s boundary condition not specified." % side)
        raise NotImplementedError("Boundary condition %s not implemented" % bc_type)

    # -- solver.py:67-85 (the axis the reference selects with np.rollaxis is an argument here)
    def _bc_lower(self, bc_type, array, axis=0):
        self._ghost_call(array, axis, self._bc_code(bc_type, "Lower"), 0)

    # -- solver.py:87-105
    def _bc_upper(self, bc_type, array, axis=0):
        self._ghost_call(array, axis, 0, self._bc_code(bc_type, "Upper"))

    def _bc_codes(self, idim=0):
        lo, hi = self.bc_lower[idim], self.bc_upper[idim]
        if lo is None:
            raise Exception("Lower boundary condition not specified.")
        if hi is None:
            raise Exception("Upper boundary condition not specified.")
        for b in (lo, hi):
            if b not in (BC.extrap, BC.periodic):
                raise NotImplementedError("Boundary condition %s not implemented" % b)
        return lo, hi

    def evolve_to_time(self, solution, tend=None):
        if not self._is_set_up:
            self.setup(solution)
        take_one_step = tend is None
        tstart = solution.t
        if not self.dt_variable:
            if take_one_step:
                self.max_steps = 1
            else:
                self.max_steps = int((tend - tstart + 1e-10) / self.dt)
                if abs(self.max_steps * self.dt - (tend - tstart)) > 1e-5 * (tend - tstart):
                    raise Exception("dt does not divide (tend-tstart) and dt is fixed!")
        for n in range(self.max_steps):
            if self.before_step is not None:
                self.before_step(self, solution.state)
            self.step(solution, take_one_step, tstart, tend)
            solution.t = tstart + (n + 1) * self.dt

    def step(self, solution, take_one_step, tstart, tend):
        raise NotImplementedError("No stepping routine has been defined!")
