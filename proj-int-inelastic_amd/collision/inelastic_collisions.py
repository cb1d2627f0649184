"""Fast spectral inelastic VHS collision operator on the HIP device.

Drop-in for reference ``src/collision/inelastic_collisions.py``:
``FSInelasticVHSCollision(config, spectral_mesh)`` with ``collide(input_f,
with_loss=False)`` (:208-227) and ``laplacian(input_f)`` (:229-230).  The
parameters are read as in ``load_parameters`` (:30-52): e, e0 (default 1.0,
giving the tanh blend of :77-79), gamma.  ``perform_precomputation`` builds
the device plan (quadrature, per-restitution kernel tables) instead of the
reference's (Nr, M, n) complex tables.

Extensions (keyword-only, all defaulting to the reference behaviour):
  precision   constructor keyword: "fp64" (plain FP64 form), "fp32" (FP32
              transforms and angular sums, FP64 radial sums, final transform
              and loss term; fused kernels for 3D 32^3 and 16^3 and 2D 16^2,
              32^2 and 64^2) or "fp32k" (as "fp32" with Kahan-compensated angular
              sums in the fused 3D 32^3 kernel).
  with_loss   constructor keyword: make ``__call__`` return gain - loss.
  e, e0       per-cell restitution arrays for ``collide`` (batched problems
              with different e; e0 defaults to e, i.e. constant restitution).
  gamma       per-cell VHS exponent array for ``collide`` (Maxwell and hard
              spheres mixed in one batched launch).
  project     apply the FP64 conservative projection (mass, momentum) to Q.
  autotune    constructor keyword: time the engine's launch-shape candidates
              on the first batch of each size bucket and keep the fastest.
  static_scratch  ``collide`` keyword: run through ``ib_collide_static``, whose
              scratch is a static device region (no device allocation; batches
              whose scratch exceeds it raise ValueError).
"""
import ctypes
import math

import numpy as np

from .. import _lib
from .. import device as dev
from .base import BaseCollision


def _per_cell(x, nb):
    """A scalar, a (nb,) array or an array of the batch shape -> (nb,) float64."""
    a = np.asarray(x, dtype=np.float64)
    return a.reshape(nb) if a.size == nb else np.broadcast_to(a, (nb,))


class FSInelasticVHSCollision(BaseCollision):

    collision_model = "vhs"

    def __init__(self, config, spectral_mesh, precision="fp64", with_loss=False, autotune=False,
                 **kwargs):
        if precision not in ("fp64", "fp32", "fp32k"):
            raise ValueError("precision must be 'fp64', 'fp32' or 'fp32k'")
        self.precision = precision
        self.autotune = bool(autotune)
        self._tuned = set()
        self.call_with_loss = bool(with_loss)
        self._plan = None
        self._param_key = None
        super().__init__(config, spectral_mesh, **kwargs)

    # -- inelastic_collisions.py:30-52
    def load_parameters(self):
        cm = self.config["collision-model"]
        self._e = float(cm.get("e", 1.0))
        self._e0 = float(cm.get("e0", 1.0))
        self._gamma = float(cm.get("gamma", 0.0))
        if self.ndim not in (2, 3):
            raise ValueError("Dimension must be 2 or 3.")
        self._sphr_fac = 2 * math.pi ** (0.5 * self.ndim) / math.gamma(0.5 * self.ndim)

    def e_profile(self, e, e0):
        """e_j = (e0 - e)/2 tanh(r_j - R/2) + (e0 + e)/2 (inelastic_collisions.py:77-79)."""
        r, _ = self.sm.rquad()
        return 0.5 * (e0 - e) * np.tanh(r - 0.5 * self.sm.R) + 0.5 * (e0 + e)

    # -- replaces inelastic_collisions.py:57-137
    def perform_precomputation(self):
        lib = _lib.load()
        r, wr = self.sm.rquad()
        sigma, _ = self.sm.circ_or_sphr_quad()
        wsig = self.sm.sphere_weights()
        self._quad = [np.ascontiguousarray(a, dtype=np.float64) for a in (r, wr, sigma, wsig)]
        self.e_r = np.ascontiguousarray(self.e_profile(self._e, self._e0)[None, :])
        self._plan = self._create_plan(self.e_r)
        self._param_key = ("config",)
        self._pidx_dev = None

    def _create_plan(self, e_r):
        lib = _lib.load()
        r, wr, sigma, wsig = self._quad
        d = _lib.ib_plan_desc()
        d.dim = self.ndim
        d.n = self.sm.nv
        d.nr = r.size
        d.m = sigma.shape[0]
        d.L = self.sm.L
        d.gamma = self._gamma
        d.precision = {"fp64": _lib.PREC_FP64, "fp32": _lib.PREC_FP32,
                       "fp32k": _lib.PREC_FP32_KAHAN}[self.precision]
        d.n_param = e_r.shape[0]
        dp = ctypes.POINTER(ctypes.c_double)
        d.r = r.ctypes.data_as(dp)
        d.wr = wr.ctypes.data_as(dp)
        d.sigma = sigma.ctypes.data_as(dp)
        d.wsigma = wsig.ctypes.data_as(dp)
        d.e_r = e_r.ctypes.data_as(dp)
        h = ctypes.c_void_p()
        _lib.check(lib.ib_plan_create(ctypes.byref(d), ctypes.byref(h)), "ib_plan_create")
        return h

    def __del__(self):
        if getattr(self, "_plan", None) is not None:
            try:
                _lib.load().ib_plan_destroy(self._plan)
            except Exception:
                pass
            self._plan = None

    def set_tuning(self, chunk_pairs=0, fused_cells=0, scratch_budget=0):
        """Plan knobs (0 = automatic): pairs per staged chunk, cells per fused launch,
        scratch budget in bytes (staged sub-batches and fused launches are cut to fit it)."""
        _lib.check(_lib.load().ib_plan_set_tuning(self._plan, int(chunk_pairs), int(fused_cells),
                                                  int(scratch_budget)), "ib_plan_set_tuning")

    def set_tiles(self, angle_split=1, pair_rounds=0):
        """Tile shape of the fused engines: workgroups per (cell, radial node), each over a
        contiguous range of the angles (forms with one workgroup per (cell, j)), and rounds of
        pairs per workgroup of the 2D 16^2 / 32^2 form (0 = 16 at 16^2, 8 at 32^2)."""
        _lib.check(_lib.load().ib_plan_set_tiles(self._plan, int(angle_split), int(pair_rounds)),
                   "ib_plan_set_tiles")

    def tiles(self):
        """(angle_split, pair_rounds) in use."""
        a, r = ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().ib_plan_get_tiles(self._plan, ctypes.byref(a), ctypes.byref(r)),
                   "ib_plan_get_tiles")
        return a.value, r.value

    def scratch_need(self, cells, grouped=True):
        """(need, held): device bytes the plan's scratch will hold after a call of `cells`
        cells (a segment of collide_group when grouped) and the bytes it holds now."""
        need, held = ctypes.c_longlong(), ctypes.c_longlong()
        _lib.check(_lib.load().ib_plan_scratch_need(self._plan, int(cells), int(bool(grouped)),
                                                    ctypes.byref(need), ctypes.byref(held)),
                   "ib_plan_scratch_need")
        return need.value, held.value

    def set_timing(self, enable=True):
        """Record HIP events around the dominant kernel of every collide call."""
        _lib.check(_lib.load().ib_plan_set_timing(self._plan, int(bool(enable))), "ib_plan_set_timing")

    def get_timing(self):
        """(milliseconds, launches, units) summed over the recorded launches."""
        ms, nl, nu = ctypes.c_double(), ctypes.c_longlong(), ctypes.c_longlong()
        _lib.check(_lib.load().ib_plan_get_timing(self._plan, ctypes.byref(ms), ctypes.byref(nl),
                                                  ctypes.byref(nu)), "ib_plan_get_timing")
        return ms.value, nl.value, nu.value

    def set_param_sets(self, e_r, gamma=None):
        """Install parameter sets: restitution profiles (n_param, nr) and optional
        VHS exponents (n_param,); cells select one set by index."""
        e_r = np.ascontiguousarray(e_r, dtype=np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        g = None if gamma is None else np.ascontiguousarray(gamma, dtype=np.float64)
        _lib.check(_lib.load().ib_plan_set_params(
            self._plan, e_r.shape[0], e_r.ctypes.data_as(dp),
            None if g is None else g.ctypes.data_as(dp), dev.stream_handle()), "ib_plan_set_params")
        self.e_r = e_r

    def _param_index(self, nb, e, e0, gamma=None):
        """Per-cell (e, e0, gamma) -> device parameter-set index tensor (or None)."""
        t = dev.torch()
        if e is None and gamma is None:
            if self._param_key != ("config",):
                self.set_param_sets(self.e_profile(self._e, self._e0)[None, :])
                self._param_key = ("config",)
            return None
        if e is None:
            e, e0 = self._e, self._e0
        e = _per_cell(e, nb)
        e0 = e if e0 is None else _per_cell(e0, nb)
        g = _per_cell(self._gamma if gamma is None else gamma, nb)
        triples = np.stack([e, e0, g], axis=1)
        cache = getattr(self, "_pidx_cache", None)
        tkey = triples.tobytes()
        if cache is not None and cache[0] == tkey and self._param_key == cache[1]:
            return cache[2]  # no host-to-device copy (safe inside graph capture)
        uniq, inv = np.unique(triples, axis=0, return_inverse=True)
        key = ("cells", uniq.tobytes())
        if self._param_key != key:
            self.set_param_sets(np.stack([self.e_profile(a, b) for a, b, _ in uniq]), uniq[:, 2])
            self._param_key = key
        idx = t.from_numpy(inv.reshape(-1).astype(np.int32)).to("cuda")
        self._pidx_cache = (tkey, key, idx)
        return idx

    def _run(self, input_f, fn):
        is_dev = dev.is_device_tensor(input_f)
        f = dev.to_device(input_f)
        d = self.ndim
        if tuple(f.shape[-d:]) != (self.sm.nv,) * d:
            raise ValueError("input_f must end with %d velocity axes of size %d" % (d, self.sm.nv))
        nb = int(np.prod(f.shape[:-d], dtype=np.int64)) if f.dim() > d else 1
        out = dev.torch().empty_like(f)
        fn(f, out, nb)
        if is_dev:
            return out
        return out.cpu().numpy()

    def collide(self, input_f, with_loss=False, heat_bath=None, e=None, e0=None, project=False,
                gamma=None, static_scratch=False):
        lib = _lib.load()
        entry = lib.ib_collide_static if static_scratch else lib.ib_collide

        def go(f, out, nb):
            if self.autotune and not static_scratch:
                from .. import autotune as at
                b = at.bucket(nb)
                if b not in self._tuned:
                    self._tuned.add(b)  # before timing: tune() calls collide()
                    at.tune(self, f, with_loss=with_loss)
            pidx = self._param_index(nb, e, e0, gamma)
            _lib.check(entry(self._plan, f.data_ptr(), out.data_ptr(), nb, dev.ptr(pidx),
                             int(bool(with_loss)), float(heat_bath or 0.0), dev.stream_handle()),
                       "ib_collide_static" if static_scratch else "ib_collide")
            if project:
                _lib.check(lib.ib_project_conservative(out.data_ptr(), nb, self.ndim, self.sm.nv,
                                                       self.sm.L, dev.stream_handle()),
                           "ib_project_conservative")

        return self._run(input_f, go)

    def laplacian(self, input_f):
        lib = _lib.load()

        def go(f, out, nb):
            _lib.check(lib.ib_laplacian(self._plan, f.data_ptr(), out.data_ptr(), nb,
                                        dev.stream_handle()), "ib_laplacian")

        return self._run(input_f, go)

    def moments(self, input_f):
        """FP64 [rho, momentum (d), energy] per cell (times dv^d)."""
        lib = _lib.load()
        t = dev.torch()
        is_dev = dev.is_device_tensor(input_f)
        f = dev.to_device(input_f)
        d = self.ndim
        nb = int(np.prod(f.shape[:-d], dtype=np.int64)) if f.dim() > d else 1
        out = t.empty((nb, d + 2), dtype=t.float64, device="cuda")
        _lib.check(lib.ib_moments(f.data_ptr(), out.data_ptr(), nb, d, self.sm.nv, self.sm.L,
                                  dev.stream_handle()), "ib_moments")
        out = out.reshape(tuple(f.shape[:-d]) + (d + 2,))
        return out if is_dev else out.cpu().numpy()


def set_group_mode(fused_grouped=True):
    """collide_group's handling of FP32 operators with a fused form: one grouped launch of
    all their quadrature-pair loops (True, the default) or one fused launch per operator
    (False); both give bit-identical results.  Process-wide."""
    _lib.check(_lib.load().ib_set_group_mode(int(bool(fused_grouped))), "ib_set_group_mode")


def collide_group(colls, fs, with_loss=False, heat_bath=None, es=None, gammas=None):
    """Q(f, f) of an uneven batch in grouped launches (one launch per stage: the pair loops
    of all FP32 fused-form operators share one launch, the staged operators' passes another).

    colls: operators of one precision (any grid shapes); fs: device tensors
    (*B_i, N_i^d_i); es/gammas: optional per-cell arrays per operator.
    Returns the list of Q tensors.
    """
    lib = _lib.load()
    t = dev.torch()
    n = len(colls)
    if len({id(c) for c in colls}) != n:
        # one operator's parameter-set tables and scratch serve one segment of the launch
        raise ValueError("an operator appears twice in collide_group; concatenate its cells instead")
    outs, keep, pidx = [], [], []
    nb = (ctypes.c_longlong * n)()
    for i, (c, f) in enumerate(zip(colls, fs)):
        f = dev.to_device(f)
        keep.append(f)
        d = c.ndim
        nb[i] = int(np.prod(f.shape[:-d], dtype=np.int64)) if f.dim() > d else 1
        outs.append(t.empty_like(f))
        e = None if es is None else es[i]
        g = None if gammas is None else gammas[i]
        pi = c._param_index(nb[i], e, None, g)
        keep.append(pi)
        pidx.append(dev.ptr(pi))
    P = ctypes.c_void_p
    arr = lambda vals: (P * n)(*[P(v) for v in vals])
    _lib.check(lib.ib_collide_group(
        n, arr([c._plan.value for c in colls]), arr([f.data_ptr() for f in keep[0::2]]),
        arr([o.data_ptr() for o in outs]), nb, arr(pidx), int(bool(with_loss)),
        float(heat_bath or 0.0), dev.stream_handle()), "ib_collide_group")
    return outs
