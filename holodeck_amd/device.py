"""Device buffers and streams (PyTorch-ROCm tensors used only as HBM buffers)."""
import collections

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
    """numpy array / scalar sequence / tensor -> contiguous device tensor (float64 default).

    Host arrays up to STAGE_MAX_BYTES are copied into a pinned staging ring and uploaded
    asynchronously on the current stream, so the host does not wait for the stream's earlier
    work; larger arrays are copied synchronously (host writes into pinned memory run well below
    the runtime's own pageable-copy path at that size).  A tensor returned here is ready for work
    on the current stream; another stream must wait for it (an event).
    """
    dev = require_device()
    if isinstance(a, torch.Tensor):
        t = a.to(dev)
        if dtype is not None:
            t = t.to(dtype)
        return t.contiguous()
    arr = np.ascontiguousarray(a)
    if dtype is None:
        dtype = torch.int64 if np.issubdtype(arr.dtype, np.integer) else torch.float64
    host = torch.as_tensor(arr).to(dtype)
    if host.numel() == 0:
        return torch.empty(host.shape, dtype=dtype, device=dev)
    staged = _staging().put(host) if host.numel() * host.element_size() <= STAGE_MAX_BYTES else None
    if staged is None:                          # large: a synchronous copy (the runtime's own staging)
        return host.to(dev)
    out = torch.empty(host.shape, dtype=dtype, device=dev)
    out.copy_(staged, non_blocking=True)
    _staging().done()
    return out


STAGE_MAX_BYTES = 256 << 10


class _Staging:
    """A pinned host ring (allocated once) for asynchronous uploads.  Each upload copies its
    data into the next free region and records an event after the device copy on the current
    stream; a region is rewritten only after the event of the copy that last used it."""

    def __init__(self, nbytes=64 << 20):
        self.buf = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
        self.n = nbytes
        self.head = 0
        self.lap = 0
        self.pending = collections.deque()      # (lap, start, end, event) in allocation order
        self._last = None

    def put(self, host):
        nb = host.numel() * host.element_size()
        size = (nb + 255) & ~255
        if size > self.n // 4:
            return None
        if self.head + size > self.n:
            self.head = 0
            self.lap += 1
        start, end = self.head, self.head + size
        while self.pending and (self.pending[0][0] < self.lap - 1 or
                                (self.pending[0][0] < self.lap and self.pending[0][1] < end)):
            self.pending.popleft()[3].synchronize()
        view = self.buf[start:start + nb].view(host.dtype).view(host.shape)
        view.copy_(host)
        self.head = end
        self._last = (self.lap, start, end)
        return view

    def done(self):
        ev = torch.cuda.Event()
        ev.record()
        self.pending.append(self._last + (ev,))


_STAGING = {}


def _staging():
    dev = torch.cuda.current_device()
    st = _STAGING.get(dev)
    if st is None:
        st = _STAGING[dev] = _Staging()
    return st


def empty(shape, dtype=None):
    dev = require_device()
    return torch.empty(shape, dtype=dtype or torch.float64, device=dev)


def workspace(nbytes):
    dev = require_device()
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=dev)


def to_host(t):
    return t.detach().cpu().numpy()
