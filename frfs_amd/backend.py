"""Device matrices in the solver's AoSoA layout (plumbing over torch device
memory; the layout follows frfs/backends/base/types.py:9-102 and
frfs/backends/cuda/types.py:17-53).

A 3D matrix of ioshape (nupts, nvar, neles) is stored as nrow = nupts rows
of leaddim doubles; element e, variable v, point u lives at
u*leaddim + ((e//soasz)*nvar + v)*soasz + e%soasz.
"""
import numpy as np


class HIPBackend:
    """The two layout constants the operator needs (frfs/backends/cuda/base.py:48-52)."""
    soasz = 32
    alignb = 128

    def __init__(self, device=None):
        import torch
        if not torch.cuda.is_available():
            raise RuntimeError('HIPBackend needs a ROCm GPU')
        self.torch = torch
        self.device = torch.device('cuda', torch.cuda.current_device() if device is None else device)

    def matrix(self, ioshape, initval=None, dtype=np.float64):
        return HIPMatrix(self, ioshape, initval, dtype)

    def stream_handle(self):
        return self.torch.cuda.current_stream(self.device).cuda_stream


def aosoa_geometry(ioshape, soasz=32, alignb=128, itemsize=8):
    nupts, nvar, neles = ioshape
    nparr = neles - neles % -soasz
    ncol = nvar * nparr
    ldmod = alignb // itemsize
    leaddim = ncol - (ncol % -ldmod)
    return nupts, ncol, leaddim, nparr


def pack(ary, soasz=32, alignb=128):
    """ioshape array -> (nrow, leaddim) AoSoA buffer (host)."""
    nupts, nvar, neles = ary.shape
    nrow, ncol, ld, nparr = aosoa_geometry(ary.shape, soasz, alignb)
    buf = np.zeros((nrow, ld), dtype=ary.dtype)
    a = np.zeros((nupts, nvar, nparr), dtype=ary.dtype)
    a[..., :neles] = ary
    a = a.reshape(nupts, nvar, nparr // soasz, soasz).transpose(0, 2, 1, 3)
    buf[:, :ncol] = a.reshape(nupts, ncol)
    return buf


def unpack(buf, ioshape, soasz=32):
    nupts, nvar, neles = ioshape
    _, ncol, _, nparr = aosoa_geometry(ioshape, soasz)
    a = buf[:, :ncol].reshape(nupts, nparr // soasz, nvar, soasz).transpose(0, 2, 1, 3)
    return np.ascontiguousarray(a.reshape(nupts, nvar, nparr)[..., :neles])


class HIPMatrix:
    def __init__(self, backend, ioshape, initval=None, dtype=np.float64):
        torch = backend.torch
        self.backend = backend
        self.ioshape = tuple(ioshape)
        self.dtype = np.dtype(dtype)
        self.nrow, self.ncol, self.leaddim, _ = aosoa_geometry(
            self.ioshape, backend.soasz, backend.alignb, self.dtype.itemsize)
        self.traits = (self.nrow, self.leaddim, self.dtype)
        self.data = torch.zeros(self.nrow * self.leaddim, dtype=torch.float64,
                                device=backend.device)
        if initval is not None:
            self.set(initval)

    @property
    def _as_parameter_(self):
        return self.data.data_ptr()

    def set(self, ary):
        ary = np.asarray(ary, dtype=self.dtype)
        if ary.shape != self.ioshape:
            raise ValueError('Invalid initial value')
        buf = pack(ary, self.backend.soasz, self.backend.alignb)
        self.data.copy_(self.backend.torch.from_numpy(buf.reshape(-1)))

    def get(self):
        buf = self.data.cpu().numpy().reshape(self.nrow, self.leaddim)
        return unpack(buf, self.ioshape, self.backend.soasz)
