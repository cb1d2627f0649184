"""Iterative steady-state solvers (reference ``src/iter_solver.py:1-62``).

``BoltzmannFastSweep.step`` evaluates the gain term Q+(f) with the collision
operator (its gain-only ``__call__``, as the reference), the density with the
FP64 moment kernel, then runs the Gauss-Seidel space sweep of every velocity
row in one HIP kernel (``ib_fast_sweep``: a thread carries one velocity point
through the recurrence over x) and returns max |f_new - f_old| as the
reference does.  The state's f stays on the device.
"""
import logging
import math

import numpy as np

from . import _lib
from . import device as dev


def set_logger(log_path):
    """Log INFO to ``log_path`` (reference ``src/utility.py:175-200``); None logs nowhere new."""
    logger = logging.getLogger()
    logger.setLevel(logging.INFO)
    if log_path and not logger.handlers:
        handler = logging.FileHandler(log_path)
        handler.setFormatter(logging.Formatter("%(asctime)s:%(levelname)s: %(message)s"))
        logger.addHandler(handler)


class IterativeBoltzmannSolver(object):
    def __init__(self, log_path, collision=None, kn=0.1, tol=1e-6):
        set_logger(log_path)
        self.collision = collision
        self.kn = kn
        self.tol = tol

    def solve(self, solution, n_iter=0, disp=False):
        """iter_solver.py:19-35: iterate until err <= tol (or n_iter steps); (err, steps)."""
        err = 1
        num_steps = 1
        while err > self.tol:
            if n_iter and num_steps >= n_iter:
                return err, num_steps
            err = self.step(solution.state)
            num_steps += 1
            if disp and (num_steps % 1 == 0):
                logging.info("Current error: {}; Number of iterations: {}".format(err, num_steps))
        return err, num_steps

    def step(self, solution):
        raise NotImplementedError("No stepping routine has been defined!")


class BoltzmannFastSweep(IterativeBoltzmannSolver):
    _vx_cache = None

    def _vx(self, state):
        v = np.ascontiguousarray(np.asarray(state.problem_data["v1"])[:, 0], dtype=np.float64)
        key = v.tobytes()
        if self._vx_cache is None or self._vx_cache[0] != key:
            self._vx_cache = (key, dev.to_device(v))
        return self._vx_cache[1]

    def step(self, state):
        """iter_solver.py:38-62 (one sweep; f updated in place on the device)."""
        t = dev.require_gpu()
        f = state.f
        if f.dim() != 3:
            raise ValueError("the fast sweep expects f of shape (nx, N, N)")
        nx, dx = int(state.num_cells[0]), float(state.delta[0])
        nu = 1 / np.sqrt(math.pi / 2) / self.kn
        rho = state.rho.contiguous()
        gain = self.collision(f)
        if not dev.is_device_tensor(gain):
            gain = dev.to_device(gain)
        gain = gain.contiguous()
        vx = self._vx(state)
        err = t.empty(1, dtype=t.float64, device="cuda")
        _lib.check(_lib.load().ib_fast_sweep(f.data_ptr(), gain.data_ptr(), rho.data_ptr(),
                                             vx.data_ptr(), nx, f.shape[1], f.shape[2], dx,
                                             float(nu), err.data_ptr(), dev.stream_handle()),
                   "ib_fast_sweep")
        return float(err.item())
