"""Solver base (reference ``src/solver.py``): BC enum, fixed-dt
``evolve_to_time`` (:107-143), and the ghost-cell conventions (:51-105), which
the device stencil kernel applies on the fly (no ghost array is copied)."""


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
