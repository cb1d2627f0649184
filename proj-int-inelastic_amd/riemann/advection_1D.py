"""Upwind Riemann solver for v_x d/dx f (reference ``src/riemann/advection_1D.py:4-21``).

The solver's ``df_advection`` recognises this function and runs the fused
HIP stencil (ghost cells + waves + limiter + second-order correction).
Called directly on device tensors it returns (wave, s, amdq, apdq).
"""


def advection_1D(q_l, q_r, problem_data):
    from .. import device as dev
    t = dev.torch()
    wave = q_r - q_l
    s = dev.to_device(problem_data["v1"]) if not dev.is_device_tensor(q_l) else \
        t.as_tensor(problem_data["v1"], dtype=q_l.dtype, device=q_l.device)
    s = s.expand(q_l.shape[1:]).contiguous()
    apdq = s.clamp(min=0.0) * wave
    amdq = s.clamp(max=0.0) * wave
    return wave, s, amdq, apdq


advection_1D.kernel_rp = "advection_1D"
