"""Test infrastructure only: the CPU oracle of the Fourier-layer surrogate.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module; the product path (frfs_amd.surrogate) runs the HIP
kernels of libfrfs_fq.so and never calls into it.

The surrogate is not part of the reference (frfs has no learned operator);
its definition is the one the HIP kernels implement, restated here in
float64 NumPy for any grid N, width W, modes m (even, m <= N) and depth L:

    v_0(c, x)   = P_c * g(x) + p0_c                              (lift)
    X_l(c, k)   = (1/n) sum_x v_l(c, x) exp(-2 pi i k.x / N)     (k in the mode set)
    Y_l(o, k)   = sum_i X_l(i, k) R_l(i, o, k)
    s_l(o, x)   = Re sum_k Y_l(o, k) exp(+2 pi i k.x / N)
    z_l(o, x)   = sum_i Wm_l(o, i) v_l(i, x) + b_l(o) + s_l(o, x)
    v_{l+1}     = gelu(z_l) for l < L-1, z_l for the last layer
    out(x)      = sum_c Qw_c v_L(c, x) + q0                      (projection)

with n = N^3, x = (I, J, K) flattened as (I*N + J)*N + K, the mode set per
axis {0..m/2-1} u {N-m/2..N-1} (wavenumbers -m/2..m/2-1 in FFT order) and
gelu(z) = z * Phi(z) (the erf form).  The parameter vector theta is laid out
as [P (W), p0 (W), per layer: R (W*W*m^3 complex, (i, o, k) order, real and
imaginary parts interleaved), Wm (W*W, (o, i) order), b (W)], Qw (W), q0.

Parity pins (tests/test_surrogate_cpu.py): the truncated transforms against
numpy.fft.fftn / ifftn restricted to the mode set; the backward pass against
central finite differences of the float64 forward; the layer with R = 0 and
Wm = identity against a closed form.
"""
from __future__ import annotations

import math

import numpy as np
from scipy.special import erf


def mode_index(N, m):
    """Grid wavenumber indices of the mode set along one axis (FFT order)."""
    h = m // 2
    return np.array(list(range(h)) + list(range(N - h, N)))


def layout(W, m, L):
    """Offsets (in floats) of the parameter blocks of theta."""
    m3 = m ** 3
    off = {'P': 0, 'p0': W}
    base = 2 * W
    per = 2 * W * W * m3 + W * W + W
    for l in range(L):
        off[f'R{l}'] = base + l * per
        off[f'W{l}'] = base + l * per + 2 * W * W * m3
        off[f'b{l}'] = base + l * per + 2 * W * W * m3 + W * W
    off['Qw'] = base + L * per
    off['q0'] = base + L * per + W
    off['total'] = base + L * per + W + 1
    return off


def unpack(theta, W, m, L):
    """theta -> dict of arrays (views into a float64 copy)."""
    th = np.asarray(theta, dtype=np.float64)
    off = layout(W, m, L)
    m3 = m ** 3
    p = {'P': th[off['P']:off['P'] + W], 'p0': th[off['p0']:off['p0'] + W],
         'Qw': th[off['Qw']:off['Qw'] + W], 'q0': th[off['q0']]}
    for l in range(L):
        r = th[off[f'R{l}']:off[f'R{l}'] + 2 * W * W * m3].reshape(W, W, m3, 2)
        p[f'R{l}'] = r[..., 0] + 1j * r[..., 1]
        p[f'W{l}'] = th[off[f'W{l}']:off[f'W{l}'] + W * W].reshape(W, W)
        p[f'b{l}'] = th[off[f'b{l}']:off[f'b{l}'] + W]
    return p


def pack_grad(g, W, m, L):
    """dict of gradients -> flat vector in theta's layout.  Complex
    gradients are stored as (d/d Re, d/d Im)."""
    off = layout(W, m, L)
    out = np.zeros(off['total'])
    out[off['P']:off['P'] + W] = g['P']
    out[off['p0']:off['p0'] + W] = g['p0']
    out[off['Qw']:off['Qw'] + W] = g['Qw']
    out[off['q0']] = g['q0']
    m3 = m ** 3
    for l in range(L):
        r = np.stack([g[f'R{l}'].real, g[f'R{l}'].imag], axis=-1).reshape(-1)
        out[off[f'R{l}']:off[f'R{l}'] + 2 * W * W * m3] = r
        out[off[f'W{l}']:off[f'W{l}'] + W * W] = g[f'W{l}'].reshape(-1)
        out[off[f'b{l}']:off[f'b{l}'] + W] = g[f'b{l}']
    return out


def init_theta(W, m, L, seed):
    """Seeded initial parameters (the host code draws the same way)."""
    rng = np.random.default_rng(seed)
    off = layout(W, m, L)
    th = np.zeros(off['total'])
    th[off['P']:off['P'] + W] = rng.uniform(-1, 1, W)
    th[off['p0']:off['p0'] + W] = rng.uniform(-0.1, 0.1, W)
    sc = 1.0 / (W * W)
    m3 = m ** 3
    for l in range(L):
        th[off[f'R{l}']:off[f'R{l}'] + 2 * W * W * m3] = sc * rng.uniform(-1, 1, 2 * W * W * m3)
        lim = 1.0 / math.sqrt(W)
        th[off[f'W{l}']:off[f'W{l}'] + W * W] = rng.uniform(-lim, lim, W * W)
        th[off[f'b{l}']:off[f'b{l}'] + W] = rng.uniform(-lim, lim, W)
    lim = 1.0 / math.sqrt(W)
    th[off['Qw']:off['Qw'] + W] = rng.uniform(-lim, lim, W)
    return th


def _phases(N, m):
    """E[k, x] = exp(-2 pi i k x / N) along one axis for the mode set."""
    k = mode_index(N, m)
    return np.exp(-2j * np.pi * np.outer(k, np.arange(N)) / N)


def dft_modes(v, N, m):
    """(..., N^3) real -> (..., m^3) complex: (1/n) sum_x v exp(-i k.x)."""
    E = _phases(N, m)
    a = v.reshape(v.shape[:-1] + (N, N, N))
    a = np.einsum('...ijk,ck->...ijc', a, E)
    a = np.einsum('...ijc,bj->...ibc', a, E)
    a = np.einsum('...ibc,ai->...abc', a, E)
    return a.reshape(v.shape[:-1] + (m ** 3,)) / N ** 3


def idft_modes_re(Y, N, m):
    """(..., m^3) complex -> (..., N^3) real: Re sum_k Y_k exp(+i k.x)."""
    E = np.conj(_phases(N, m))
    a = Y.reshape(Y.shape[:-1] + (m, m, m))
    a = np.einsum('...abc,ai->...ibc', a, E)
    a = np.einsum('...ibc,bj->...ijc', a, E)
    a = np.einsum('...ijc,ck->...ijk', a, E)
    return a.real.reshape(Y.shape[:-1] + (N ** 3,))


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / math.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2 * math.pi)


def forward(theta, g, N, W, m, L, keep=False):
    """g: (B, N^3) normalised input -> (B, N^3) output (and the saved
    activations when keep is set)."""
    p = unpack(theta, W, m, L)
    g = np.asarray(g, dtype=np.float64)
    v = p['P'][None, :, None] * g[:, None, :] + p['p0'][None, :, None]
    acts = {'v0': v}
    for l in range(L):
        X = dft_modes(v, N, m)                                   # (B, W, m3)
        Y = np.einsum('bik,iok->bok', X, p[f'R{l}'])
        s = idft_modes_re(Y, N, m)
        z = np.einsum('oi,bix->box', p[f'W{l}'], v) + p[f'b{l}'][None, :, None] + s
        v = gelu(z) if l < L - 1 else z
        acts[f'X{l}'], acts[f'z{l}'], acts[f'v{l + 1}'] = X, z, v
    out = np.einsum('c,bcx->bx', p['Qw'], v) + p['q0']
    return (out, acts) if keep else out


def backward(theta, g, gout, N, W, m, L):
    """Gradient of sum(gout * out) with respect to theta (flat layout)."""
    p = unpack(theta, W, m, L)
    _, acts = forward(theta, g, N, W, m, L, keep=True)
    n = N ** 3
    gr = {}
    vL = acts[f'v{L}']
    gr['Qw'] = np.einsum('bx,bcx->c', gout, vL)
    gr['q0'] = gout.sum()
    gv = p['Qw'][None, :, None] * gout[:, None, :]
    for l in reversed(range(L)):
        z, v = acts[f'z{l}'], acts[f'v{l}']
        gz = gv * gelu_grad(z) if l < L - 1 else gv
        gr[f'W{l}'] = np.einsum('box,bix->oi', gz, v)
        gr[f'b{l}'] = gz.sum(axis=(0, 2))
        gvin = np.einsum('oi,box->bix', p[f'W{l}'], gz)
        # s = Re IDFT(Y): dL/dY = sum_x gz exp(-i k.x) (unscaled forward DFT)
        gY = dft_modes(gz, N, m) * n
        X = acts[f'X{l}']
        gr[f'R{l}'] = np.einsum('bik,bok->iok', np.conj(X), gY)
        gX = np.einsum('bok,iok->bik', gY, np.conj(p[f'R{l}']))
        # X = (1/n) DFT(v): dL/dv = (1/n) Re IDFT(gX)
        gvin = gvin + idft_modes_re(gX, N, m) / n
        gv = gvin
    gr['P'] = np.einsum('bcx,bx->c', gv, g)
    gr['p0'] = gv.sum(axis=(0, 2))
    return pack_grad(gr, W, m, L)


def invariants(cv):
    """The collision invariants 1, c_x, c_y, c_z, |c|^2 on the grid."""
    return np.stack([np.ones(cv.shape[1]), cv[0], cv[1], cv[2], (cv ** 2).sum(axis=0)])


def defect_and_project(q, cv, cw):
    """Conservation defect of q (largest cosine between q and a collision
    invariant in the cw-weighted inner product) and q with its component in
    the span of the invariants removed (least-squares projection)."""
    phi = invariants(cv)
    G = phi @ phi.T * cw
    mom = phi @ q * cw
    nq = math.sqrt(float(q @ q) * cw)
    dfct = 0.0 if nq == 0 else float(np.max(np.abs(mom) / np.sqrt(np.diag(G)) / nq))
    lam = np.linalg.solve(G, mom)
    return dfct, q - lam @ phi
