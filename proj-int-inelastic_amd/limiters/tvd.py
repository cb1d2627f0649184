"""TVD wave limiters (reference ``src/limiters/tvd.py``).

``minmod`` (:30-40) and ``mc`` (new: monotonized central) select the limiter
of the fused HIP stencil kernel (``ib_tvd1d``), which applies ``limit``
(:6-27) inside the wave-propagation update.  Called directly on device
tensors they evaluate phi(theta) with device tensor ops.  ``limit`` itself
is the standalone HIP kernel ``ib_tvd_limit``.
"""
import numpy as np

from .. import _lib
from .. import device as dev

KERNEL_CODE = {}


def _register(code):
    def deco(fn):
        KERNEL_CODE[fn] = code
        fn.kernel_code = code
        return fn
    return deco


@_register(0)
def minmod(r, cfl):
    """phi = max(0, min(1, r)); cfl is unused, as in the reference."""
    return r.clamp(min=0.0, max=1.0) if hasattr(r, "clamp") else __import__("numpy").clip(r, 0.0, 1.0)


@_register(1)
def mc(r, cfl):
    """phi = max(0, min((1 + r)/2, 2, 2 r))."""
    import numpy as np
    if hasattr(r, "clamp"):
        return ((1.0 + r) * 0.5).clamp(max=2.0).minimum(2.0 * r).clamp(min=0.0)
    return np.maximum(0.0, np.minimum(np.minimum(0.5 * (1.0 + r), 2.0), 2.0 * r))


def limiter_code(fn):
    code = getattr(fn, "kernel_code", None)
    if code is None:
        name = getattr(fn, "__name__", "")
        code = {"minmod": 0, "mc": 1}.get(name)
    if code is None:
        raise NotImplementedError("limiter %r has no device kernel (minmod or mc)" % (fn,))
    return code


def limit(wave, s, limiter, dtdx):
    """Limit the waves of interfaces 1 .. n - 2 in place and return ``wave`` (tvd.py:6-27).

    wave: (n, ...) float64, a numpy array or a device tensor; s: the speeds, broadcastable to
    wave[1:-1] (one value, the trailing shape, or the full shape of wave[1:-1]); limiter:
    ``minmod`` or ``mc``.  theta = (s > 0 ? W[i-1] W[i] : W[i+1] W[i]) / W[i]^2, 0 where numpy's
    masked division masks; W = phi(theta) W where W != 0.  dtdx only enters the cfl argument,
    which neither limiter reads.  Computed by one HIP kernel on the current stream.
    """
    t = dev.require_gpu()
    code = limiter_code(limiter)
    on_dev = dev.is_device_tensor(wave)
    w = dev.to_device(wave)
    n = int(w.shape[0]) if w.dim() else 0
    inner = int(np.prod(w.shape[1:], dtype=np.int64)) if w.dim() > 1 else 1
    sh = np.asarray(s.cpu().numpy() if dev.is_device_tensor(s) else s, dtype=np.float64)
    if n > 2:
        full = (n - 2,) + tuple(w.shape[1:])
        # broadcasting fixes which speed each wave sees: a scalar, one per column, or one per wave
        if sh.ndim == len(full) and sh.shape[0] == 1:
            sh = sh[0]
        if sh.size == 1:
            sh = sh.reshape(1)
        elif sh.ndim <= len(full) - 1:
            sh = np.broadcast_to(sh, full[1:]).reshape(-1)
        else:
            sh = np.broadcast_to(sh, full).reshape(-1)
    else:
        sh = sh.reshape(-1)[:1] if sh.size else np.zeros(1)
    sd = dev.to_device(np.ascontiguousarray(sh))
    out = t.empty_like(w)
    _lib.check(_lib.load().ib_tvd_limit(dev.ptr(w), dev.ptr(out), n, inner, sd.data_ptr(), sd.numel(),
                                        code, dev.stream_handle()), "ib_tvd_limit")
    if on_dev:
        wave.copy_(out.view(wave.shape))
        return wave
    wave[...] = out.cpu().numpy().reshape(wave.shape)
    return wave
