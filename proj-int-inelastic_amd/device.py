"""Device-buffer plumbing: torch-ROCm tensors as HIP buffers, current stream."""
import numpy as np

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


def require_gpu():
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("a HIP device is required (no CPU fallback)")
    return t


def stream_handle():
    return torch().cuda.current_stream().cuda_stream


def is_device_tensor(x):
    t = torch()
    return isinstance(x, t.Tensor) and x.is_cuda


def to_device(x, dtype=None):
    """numpy / torch -> contiguous float64 (or dtype) CUDA tensor."""
    t = require_gpu()
    dtype = dtype or t.float64
    if isinstance(x, t.Tensor):
        return x.to(device="cuda", dtype=dtype).contiguous()
    return t.from_numpy(np.ascontiguousarray(x)).to(device="cuda", dtype=dtype).contiguous()


def ptr(x):
    return None if x is None else x.data_ptr()
