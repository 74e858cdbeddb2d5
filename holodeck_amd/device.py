"""Device buffers and streams (PyTorch-ROCm tensors used only as HBM buffers)."""
import numpy as np

try:
    import torch
except ImportError:   # pragma: no cover
    torch = None

from . import _lib


def require_device():
    """Return the HIP device, or raise: this package has no CPU path."""
    if torch is None or not torch.cuda.is_available():
        raise _lib.HoloHipError("holodeck_amd needs a HIP (ROCm) GPU; none is available")
    _lib.load()
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr():
    return torch.cuda.current_stream().cuda_stream


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return t.data_ptr()


def to_dev(a, dtype=None):
    """numpy array / scalar sequence / tensor -> contiguous device tensor (float64 default)."""
    dev = require_device()
    if isinstance(a, torch.Tensor):
        t = a.to(dev)
        if dtype is not None:
            t = t.to(dtype)
        return t.contiguous()
    arr = np.ascontiguousarray(a)
    if dtype is None:
        dtype = torch.int64 if np.issubdtype(arr.dtype, np.integer) else torch.float64
    return torch.as_tensor(arr).to(device=dev, dtype=dtype).contiguous()


def empty(shape, dtype=None):
    dev = require_device()
    return torch.empty(shape, dtype=dtype or torch.float64, device=dev)


def workspace(nbytes):
    dev = require_device()
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=dev)


def to_host(t):
    return t.detach().cpu().numpy()
