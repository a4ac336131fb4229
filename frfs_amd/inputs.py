"""Seeded synthetic inputs: batches of two-Maxwellian mixtures, generated on
the device in the solver's AoSoA layout.

Per DOF and component k = 1, 2 (drawn in this order from
``numpy.random.default_rng(seed)``): n_k ~ U(dens), u_k ~ U(-1, 1)^3,
T_k ~ U(0.5, 2); f = sum_k n_k (m/(pi T_k))^{3/2} exp(-m |c - u_k|^2 / T_k),
the non-dimensional Maxwellian of frfs/solvers/dgfs/initcond.py:56-58
(species form frfs/solvers/dgfsbi/initcond.py:71-73).
"""
import numpy as np


def mixture_params(ndof, seed, dens_range=(0.5, 1.5)):
    rng = np.random.default_rng(seed)
    out = np.empty((ndof, 2, 5))
    for d in range(ndof):
        for k in range(2):
            nk = rng.uniform(*dens_range)
            uk = rng.uniform(-1, 1, 3)
            Tk = rng.uniform(0.5, 2)
            out[d, k] = (nk, uk[0], uk[1], uk[2], Tk)
    return out


def mixture_device(cv, params, mass=1.0, device='cuda', chunk=512):
    """(ndof, N^3) float64 tensor on ``device``."""
    import torch
    c = torch.as_tensor(np.ascontiguousarray(cv), dtype=torch.float64, device=device)   # (3, n)
    P = torch.as_tensor(params, dtype=torch.float64, device=device)
    ndof = P.shape[0]
    out = torch.empty((ndof, c.shape[1]), dtype=torch.float64, device=device)
    for d0 in range(0, ndof, chunk):
        p = P[d0:d0 + chunk]                                   # (b, 2, 5)
        acc = torch.zeros((p.shape[0], c.shape[1]), dtype=torch.float64, device=device)
        for k in range(2):
            nk, u, Tk = p[:, k, 0:1], p[:, k, 1:4], p[:, k, 4:5]
            r2 = ((c.unsqueeze(0) - u.unsqueeze(2)) ** 2).sum(1)   # (b, n)
            acc += nk * (mass / (np.pi * Tk)) ** 1.5 * torch.exp(-mass * r2 / Tk)
        out[d0:d0 + chunk] = acc
    return out


def to_aosoa(f, matrix):
    """Scatter (ndof, n) device rows, DOF order elem*nupts + upt, into an
    AoSoA matrix of ioshape (nupts, n, neles)."""
    nupts, n, neles = matrix.ioshape
    soasz = matrix.backend.soasz
    nparr = -(-neles // soasz) * soasz
    a = f.reshape(neles, nupts, n).permute(1, 2, 0)            # (nupts, n, neles)
    buf = matrix.data.view(nupts, matrix.leaddim)
    pad = a.new_zeros((nupts, n, nparr))
    pad[..., :neles] = a
    buf[:, :n * nparr] = pad.reshape(nupts, n, nparr // soasz, soasz).permute(0, 2, 1, 3).reshape(nupts, -1)


def from_aosoa(matrix):
    """Inverse of :func:`to_aosoa` -> (ndof, n) device tensor."""
    nupts, n, neles = matrix.ioshape
    soasz = matrix.backend.soasz
    nparr = -(-neles // soasz) * soasz
    buf = matrix.data.view(nupts, matrix.leaddim)[:, :n * nparr]
    a = buf.reshape(nupts, nparr // soasz, n, soasz).permute(0, 2, 1, 3).reshape(nupts, n, nparr)
    return a[..., :neles].permute(2, 0, 1).reshape(neles * nupts, n)


def shear_mixture_params(ndof, seed, max_comp=3, shear=0.25, dens_range=(0.5, 1.5)):
    """Per DOF (drawn in this order from ``numpy.random.default_rng(seed)``):
    the number of components K ~ U{1..max_comp}, then per component
    n_k ~ U(dens), u_k ~ U(-1, 1)^3, T_k ~ U(0.5, 2) and the three
    off-diagonal entries (xy, xz, yz) of a symmetric shear S_k ~ U(-shear,
    shear).  Unused component slots have n_k = 0.  Returns (ndof, max_comp, 8)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((ndof, max_comp, 8))
    out[..., 4] = 1.0
    for d in range(ndof):
        nc = int(rng.integers(1, max_comp + 1))
        for k in range(nc):
            nk = rng.uniform(*dens_range)
            uk = rng.uniform(-1, 1, 3)
            Tk = rng.uniform(0.5, 2)
            sk = rng.uniform(-shear, shear, 3)
            out[d, k] = (nk, uk[0], uk[1], uk[2], Tk, sk[0], sk[1], sk[2])
    return out


def shear_mixture_device(cv, params, device='cuda', chunk=512):
    """(ndof, N^3) float64 tensor: f = sum_k n_k sqrt(det A_k) / pi^{3/2}
    exp(-(c - u_k)^T A_k (c - u_k)) with A_k = (I + S_k) / T_k, a Maxwellian
    of density n_k sheared by S_k (|S_k| entries <= 0.25 keep A_k positive
    definite)."""
    import torch
    c = torch.as_tensor(np.ascontiguousarray(cv), dtype=torch.float64, device=device)   # (3, n)
    P = torch.as_tensor(params, dtype=torch.float64, device=device)
    ndof, ncomp = P.shape[:2]
    out = torch.empty((ndof, c.shape[1]), dtype=torch.float64, device=device)
    for d0 in range(0, ndof, chunk):
        p = P[d0:d0 + chunk]
        acc = torch.zeros((p.shape[0], c.shape[1]), dtype=torch.float64, device=device)
        for k in range(ncomp):
            nk, u, Tk = p[:, k, 0:1], p[:, k, 1:4], p[:, k, 4:5]
            sxy, sxz, syz = p[:, k, 5:6], p[:, k, 6:7], p[:, k, 7:8]
            d = c.unsqueeze(0) - u.unsqueeze(2)                      # (b, 3, n)
            dx, dy, dz = d[:, 0], d[:, 1], d[:, 2]
            q = dx * dx + dy * dy + dz * dz + 2 * (sxy * dx * dy + sxz * dx * dz + syz * dy * dz)
            det = 1 + 2 * sxy * sxz * syz - sxy ** 2 - sxz ** 2 - syz ** 2
            acc += nk * torch.sqrt(det) * (np.pi * Tk) ** -1.5 * torch.exp(-q / Tk)
        out[d0:d0 + chunk] = acc
    return out
