"""Solution (reference ``src/solution.py:27-60``): Domain + State, with
``problem_data["v1"]`` = v_x on the velocity grid (:57) and, new, ``"v2"`` = v_y
(the speed along the second space axis of a 2D-space grid)."""
from .geometry import Domain
from .state import State


class Solution(object):
    def __getattr__(self, key):
        if key in ("t", "f", "rho", "u", "T", "num_dim", "num_dim_x", "c_centers", "num_cells",
                   "lower", "upper", "delta", "centers", "grid"):
            return getattr(self.state, key)
        raise AttributeError("'Solution' object has no attribute '" + key + "'")

    def __setattr__(self, key, value):
        if key == "t":
            setattr(self.state, key, value)
        else:
            self.__dict__[key] = value

    def __init__(self, config):
        dim_x, dim_v = config.dim_x, config.dim_v
        lower = [config.xmin] * dim_x + [-config.l] * dim_v
        upper = [config.xmax] * dim_x + [config.l] * dim_v
        nn = [config.nx] * dim_x + [config.nv] * dim_v
        self.domain = Domain(lower, upper, nn)
        self.state = State(self.domain)
        self.state.grid.num_dim_x = dim_x
        cc = self.state.grid.c_centers
        v1 = cc[dim_x]
        for _ in range(dim_x):
            v1 = v1[0]
        self.state.problem_data = {"v1": v1}
        if dim_v >= 2:  # new: v_y, the speed of the second space axis of a 2D-space grid
            v2 = cc[dim_x + 1]
            for _ in range(dim_x):
                v2 = v2[0]
            self.state.problem_data["v2"] = v2
