"""Device arrays for the C-ABI, held in torch HIP tensors.

torch is used here only as the HBM allocator and for host<->device copies;
all compute goes through libvpichip.so.
"""
import numpy as np
import torch

from . import types as T


class DeviceArray:
    """A host-shaped numpy array's bytes in HBM (16-byte aligned)."""

    def __init__(self, shape, dtype, device="cuda"):
        self.dtype = np.dtype(dtype)
        self.shape = tuple(np.atleast_1d(shape).tolist()) if not isinstance(
            shape, tuple) else shape
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.t = torch.empty(max(self.nbytes, 16), dtype=torch.uint8,
                             device=device)
        assert self.t.data_ptr() % 16 == 0

    @classmethod
    def from_host(cls, a, device="cuda"):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, device)
        d.upload(a)
        return d

    @property
    def ptr(self):
        return self.t.data_ptr()

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes, (a.nbytes, self.nbytes)
        if self.nbytes:
            src = torch.from_numpy(a.reshape(-1).view(np.uint8))
            self.t[:self.nbytes].copy_(src)

    def download(self):
        torch.cuda.synchronize()
        raw = self.t[:self.nbytes].cpu().numpy()
        return raw.view(self.dtype).reshape(self.shape)

    def zero(self):
        self.t.zero_()


def particles(n, device="cuda"):
    return DeviceArray((n,), T.PARTICLE, device)


def movers(n, device="cuda"):
    return DeviceArray((max(n, 1),), T.MOVER, device)


def interpolators(g, device="cuda"):
    return DeviceArray((g.nv,), T.INTERPOLATOR, device)


def accumulators(g, device="cuda"):
    return DeviceArray((g.nv, 12), np.float32, device)


def fields(g, device="cuda"):
    return DeviceArray((g.nv,), T.FIELD, device)
