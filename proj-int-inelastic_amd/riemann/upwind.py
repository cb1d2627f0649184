"""Flux-form second-order upwind scheme with the van Leer limiter
(reference ``src/riemann/upwind.py:1-50``; no solver of the reference calls it).

Same function names and argument meaning.  ``F_p_2``, ``F_m_2``, ``grad_vf2``
and ``grad_vf1`` run one HIP stencil kernel (``ib_upwind_flux``) along axis 0
of f, with numpy's evaluation order, so results equal the reference's bit for
bit.  ``v`` (and ``vp``/``vm``) must have f's number of dimensions, with the
reference's row count along axis 0 and size-1 axes where they broadcast.
numpy inputs are copied to the device and the result copied back; torch
device tensors stay resident.  The limiter argument accepts only
``van_leer_limiter`` (the one the kernel implements); any other callable raises
``NotImplementedError``.
"""
import numpy as np

from .. import _lib
from .. import device as dev

EPS = 1e-8


def diff_x_up(psi):
    """upwind.py:6-7."""
    return psi[1:-1] - psi[:-2]


def diff_x_down(psi):
    """upwind.py:10-11."""
    return psi[2:] - psi[1:-1]


def van_leer_limiter(r):
    """upwind.py:16-17 (elementwise; numpy or torch)."""
    return (r + abs(r)) / (1. + abs(r))


def _run(f, v, rows_v, mode, dx=0.0, dt=0.0):
    was_numpy = not dev.is_device_tensor(f)
    t = dev.require_gpu()
    fd = dev.to_device(f)
    vd = dev.to_device(v)
    if fd.dim() < 1 or vd.dim() != fd.dim():
        raise ValueError("v must have the dimensions of f")
    n = fd.shape[0]
    if vd.shape[0] != rows_v(n):
        raise ValueError("v has %d rows along axis 0, expected %d" % (vd.shape[0], rows_v(n)))
    vd = vd.expand((vd.shape[0],) + tuple(fd.shape[1:])).contiguous()
    inner = int(np.prod(fd.shape[1:], dtype=np.int64)) if fd.dim() > 1 else 1
    rows = n - 2 if mode == 0 else n - 4
    out = t.empty((max(rows, 0),) + tuple(fd.shape[1:]), dtype=t.float64, device="cuda")
    _lib.check(_lib.load().ib_upwind_flux(fd.data_ptr(), vd.data_ptr(), out.data_ptr(), n, inner,
                                          float(dx), float(dt), mode, dev.stream_handle()),
               "ib_upwind_flux")
    return out.cpu().numpy() if was_numpy else out


def _check_limiter(limiter):
    if limiter is not van_leer_limiter:
        raise NotImplementedError("only van_leer_limiter is implemented by the device kernel")


def F_p_2(f, vp, dx, dt, limiter=van_leer_limiter):
    """upwind.py:20-26: F[2:-2] - F[1:-3] with the upwind flux; vp aligned with f[1:-1]."""
    _check_limiter(limiter)
    return _run(f, vp, lambda n: n - 2, 2, dx, dt)


def F_m_2(f, vm, dx, dt, limiter=van_leer_limiter):
    """upwind.py:30-37: F[3:-1] - F[2:-2] with the downwind flux; vm aligned with f[1:-1]."""
    _check_limiter(limiter)
    return _run(f, vm, lambda n: n - 2, 3, dx, dt)


def grad_vf2(f, v, dx, dt):
    """upwind.py:40-44: second-order v grad f on rows 2..n-3."""
    return _run(f, v, lambda n: n, 1, dx, dt)


def grad_vf1(f, v):
    """upwind.py:47-50: first-order v grad f on rows 1..n-2."""
    return _run(f, v, lambda n: n, 0)
