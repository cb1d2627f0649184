"""TVD wave limiters (reference ``src/limiters/tvd.py``).

``minmod`` (:30-40) and ``mc`` (new: monotonized central) select the limiter
of the fused HIP stencil kernel (``ib_tvd1d``), which applies ``limit``
(:6-27) inside the wave-propagation update.  Called directly on device
tensors they evaluate phi(theta) with device tensor ops.
"""
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
