"""CPU checks of the learned stand-in: the oracle's transforms against
numpy.fft, its backward pass against central finite differences, a closed
form of the network with the Fourier and pointwise paths switched off, the
conservation projection, the host parameter layout and the sheared-mixture
generator."""
import math

import numpy as np
import pytest

from oracle import surrogate as osur


@pytest.mark.parametrize('N,m', [(4, 2), (8, 4), (16, 8)])
def test_truncated_transforms_match_fft(N, m):
    rng = np.random.default_rng(N)
    v = rng.standard_normal((3, N ** 3))
    k = osur.mode_index(N, m)
    full = np.fft.fftn(v.reshape(3, N, N, N), axes=(1, 2, 3)) / N ** 3
    want = full[:, k][:, :, k][:, :, :, k].reshape(3, -1)
    assert np.max(np.abs(osur.dft_modes(v, N, m) - want)) < 1e-15
    Y = rng.standard_normal((3, m ** 3)) + 1j * rng.standard_normal((3, m ** 3))
    spec = np.zeros((3, N, N, N), complex)
    spec[np.ix_(range(3), k, k, k)] = Y.reshape(3, m, m, m)
    want = (np.fft.ifftn(spec, axes=(1, 2, 3)) * N ** 3).real.reshape(3, -1)
    assert np.max(np.abs(osur.idft_modes_re(Y, N, m) - want)) < 1e-12


@pytest.mark.parametrize('N,W,m,L', [(4, 3, 2, 2), (8, 4, 4, 3)])
def test_backward_matches_finite_differences(N, W, m, L):
    rng = np.random.default_rng(W)
    th = osur.init_theta(W, m, L, seed=L)
    # larger Fourier weights so that the spectral path matters in the check
    off = osur.layout(W, m, L)
    for l in range(L):
        a = off[f'R{l}']
        th[a:a + 2 * W * W * m ** 3] *= W
    g = rng.standard_normal((2, N ** 3))
    go = rng.standard_normal((2, N ** 3))
    grad = osur.backward(th, g, go, N, W, m, L)
    idx = np.arange(th.size) if th.size < 800 else np.unique(rng.integers(0, th.size, 600))
    # every block is sampled
    idx = np.union1d(idx, [off['P'], off['p0'], off['Qw'], off['q0']] +
                     [off[f'{b}{l}'] for b in 'RWb' for l in range(L)])
    h = 1e-6
    for j in idx:
        tp, tm = th.copy(), th.copy()
        tp[j] += h
        tm[j] -= h
        fd = (np.sum(go * osur.forward(tp, g, N, W, m, L)) - np.sum(go * osur.forward(tm, g, N, W, m, L))) / (2 * h)
        assert abs(grad[j] - fd) <= 1e-6 * max(1.0, abs(fd)), (j, grad[j], fd)


def test_network_closed_form_without_fourier_path():
    """R = 0, Wm = 0, b = 0 in the last layer: out = q0 (and with a single
    layer whose Wm = I: out = Qw . (P g + p0) + q0)."""
    N, W, m, L = 8, 4, 4, 1
    rng = np.random.default_rng(3)
    th = osur.init_theta(W, m, L, seed=1)
    off = osur.layout(W, m, L)
    th[off['R0']:off['R0'] + 2 * W * W * m ** 3] = 0
    th[off['W0']:off['W0'] + W * W] = np.eye(W).reshape(-1)
    th[off['b0']:off['b0'] + W] = 0
    th[off['q0']] = 0.3
    g = rng.standard_normal((2, N ** 3))
    p = osur.unpack(th, W, m, L)
    want = (p['Qw'] * p['P']).sum() * g + (p['Qw'] * p['p0']).sum() + 0.3
    assert np.allclose(osur.forward(th, g, N, W, m, L), want, atol=1e-14)


def test_projection_removes_invariants():
    from frfs_amd.surrogate import gram
    N, L = 16, 5.0
    c0 = np.linspace(-L + L / N, L - L / N, N)
    idx = np.arange(N ** 3)
    cv = np.stack([c0[idx // (N * N)], c0[(idx // N) % N], c0[idx % N]])
    cw = (2 * L / N) ** 3
    rng = np.random.default_rng(0)
    q = rng.standard_normal(N ** 3)
    d, qp = osur.defect_and_project(q, cv, cw)
    phi = osur.invariants(cv)
    assert np.max(np.abs(phi @ qp * cw)) < 1e-10 * np.abs(phi @ q * cw).max()
    assert 0 < d < 1
    # the host's Gram inverse and diagonal agree with the oracle's
    gi, gd = gram(cv, cw)
    G = phi @ phi.T * cw
    assert np.allclose(gi @ G, np.eye(5), atol=1e-10)
    assert np.allclose(gd, np.diag(G))
    # q orthogonal to the invariants has defect 0, q = phi_0 has defect 1
    assert osur.defect_and_project(qp, cv, cw)[0] < 1e-12
    assert abs(osur.defect_and_project(phi[0], cv, cw)[0] - 1) < 1e-12


def test_host_parameter_layout_matches_oracle():
    from frfs_amd.surrogate import init_params
    W, m, L = 32, 8, 4
    th = init_params(L, seed=5)
    assert th.size == osur.layout(W, m, L)['total']
    p = osur.unpack(th, W, m, L)
    assert np.abs(p['R0'].real).max() <= 1 / W ** 2 and p['q0'] == 0
    assert np.abs(p['W3']).max() <= 1 / math.sqrt(W)


def test_sheared_mixture_density_and_positivity():
    import torch
    from frfs_amd.inputs import shear_mixture_device, shear_mixture_params
    N, L = 16, 5.0
    c0 = np.linspace(-L + L / N, L - L / N, N)
    idx = np.arange(N ** 3)
    cv = np.stack([c0[idx // (N * N)], c0[(idx // N) % N], c0[idx % N]])
    cw = (2 * L / N) ** 3
    prm = shear_mixture_params(64, seed=4)
    f = shear_mixture_device(cv, prm, device='cpu').numpy()
    assert np.all(f > 0)
    rho = f.sum(axis=1) * cw
    want = prm[:, :, 0].sum(axis=1)
    # grid quadrature of Gaussians with T >= 0.5 on [-5, 5]^3: spectrally exact
    # up to the tails truncated at the box edge
    assert np.max(np.abs(rho / want - 1)) < 2e-3
    # a single unsheared component is the Maxwellian of the two-Maxwellian generator
    one = np.zeros((1, 3, 8))
    one[0, 0] = (1.2, 0.1, -0.2, 0.3, 0.8, 0, 0, 0)
    one[0, 1:, 4] = 1
    g = shear_mixture_device(cv, one, device='cpu').numpy()[0]
    r2 = (cv[0] - 0.1) ** 2 + (cv[1] + 0.2) ** 2 + (cv[2] - 0.3) ** 2
    assert np.allclose(g, 1.2 / (np.pi * 0.8) ** 1.5 * np.exp(-r2 / 0.8), rtol=1e-13)
    del torch


def test_sur_abi_rejects_bad_shapes_and_arguments():
    """The C ABI validates before touching the device: wrong ABI version,
    unsupported shapes and null buffers return error codes with a message."""
    import ctypes as C
    from frfs_amd import _lib
    lib = _lib.load()
    n = C.c_int64()
    good = _lib.fq_sur_shape(_lib.SUR_ABI_VERSION, 16, 32, 8, 4)
    assert lib.fq_sur_sizes(C.byref(good), 64, 1, C.byref(n), None, None) == _lib.SUCCESS
    assert n.value == osur.layout(32, 8, 4)['total']
    act, work = C.c_int64(), C.c_int64()
    assert lib.fq_sur_sizes(C.byref(good), 2, 1, None, C.byref(act), C.byref(work)) == _lib.SUCCESS
    vol, md = 2 * 32 * 4096, 2 * 2 * 32 * 512
    assert act.value == 5 * vol + 4 * vol + 4 * md + vol + md
    assert work.value == 2 * vol + 2 * md
    # the 32^3 grid: the same parameters, activations sized by its 32768 points
    good32 = _lib.fq_sur_shape(_lib.SUR_ABI_VERSION, 32, 32, 8, 4)
    n32 = C.c_int64()
    assert lib.fq_sur_sizes(C.byref(good32), 2, 1, C.byref(n32), C.byref(act), C.byref(work)) == _lib.SUCCESS
    vol32 = 2 * 32 * 32768
    assert n32.value == n.value
    assert act.value == 5 * vol32 + 4 * vol32 + 4 * md + vol32 + md
    assert work.value == 2 * vol32 + 2 * md
    for bad, code in ((_lib.fq_sur_shape(99, 16, 32, 8, 4), _lib.ERR_INVALID),
                      (_lib.fq_sur_shape(_lib.SUR_ABI_VERSION, 24, 32, 8, 4), _lib.ERR_UNSUPPORTED),
                      (_lib.fq_sur_shape(_lib.SUR_ABI_VERSION, 16, 16, 8, 4), _lib.ERR_UNSUPPORTED),
                      (_lib.fq_sur_shape(_lib.SUR_ABI_VERSION, 16, 32, 8, 9), _lib.ERR_UNSUPPORTED)):
        assert lib.fq_sur_sizes(C.byref(bad), 1, 0, C.byref(n), None, None) == code
        assert lib.fq_last_error()
    assert lib.fq_sur_forward(C.byref(good), None, None, 1, 0, 0, None, None, None) == _lib.ERR_INVALID
    assert lib.fq_sur_forward(C.byref(good), 8, 8, _lib.SUR_MAX_BATCH + 1, 0, 0, 8, 8, None) == _lib.ERR_INVALID
    assert lib.fq_sur_forward(C.byref(good), 8, 8, 1, 1, 5, 8, 8, None) == _lib.ERR_INVALID
    assert lib.fq_sur_backward(C.byref(good), 8, 8, 1, 0, 8, None, 8, 8, None) == _lib.ERR_INVALID
    assert lib.fq_sur_loss(-1, 8, 8, 8, 8, None) == _lib.ERR_INVALID
    assert lib.fq_sur_adam(10, 8, 8, 8, 8, C.c_float(1e-3), C.c_float(0.9), C.c_float(0.999),
                           C.c_float(1e-8), 0, None) == _lib.ERR_INVALID
    assert lib.fq_sur_gather(16, 0.1, 1.0, 1, 8, 1, 10, 4, 3, 2, None, None, 8, None) == _lib.ERR_INVALID
    assert lib.fq_sur_finish(16, 5.0, 0.1, 1.0, 1e-4, 8, 8, 8, 8, 8, 1, 10, 4, 0, 2, 8, 8,
                             None) == _lib.ERR_INVALID
    # folded inference needs two or more layers and its buffers
    one = _lib.fq_sur_shape(_lib.SUR_ABI_VERSION, 16, 32, 8, 1)
    assert lib.fq_sur_fuse(C.byref(one), 8, 8, None) == _lib.ERR_UNSUPPORTED
    assert lib.fq_sur_fuse(C.byref(good), None, 8, None) == _lib.ERR_INVALID
    assert lib.fq_sur_forward_fused(C.byref(one), 8, 8, 8, 1, 0, 8, 8, None) == _lib.ERR_UNSUPPORTED
    assert lib.fq_sur_forward_fused(C.byref(good), 8, None, 8, 1, 0, 8, 8, None) == _lib.ERR_INVALID
    assert lib.fq_sur_forward_fused(C.byref(good), 8, 8, 8, 1, 3, 8, 8, None) == _lib.ERR_INVALID
    # zero-sized work launches nothing and succeeds
    assert lib.fq_sur_forward(C.byref(good), 8, 8, 0, 0, 2, 8, 8, None) == _lib.SUCCESS
    assert lib.fq_sur_forward_fused(C.byref(good), 8, 8, 8, 0, 1, 8, 8, None) == _lib.SUCCESS


def test_stand_in_plugin_is_registered():
    from frfs_amd.scattering import DGFSScatteringModel, subclass_where
    from frfs_amd.surrogate import DGFSVHSGLLFNOScatteringModel
    assert subclass_where(DGFSScatteringModel, scattering_model='vhs-gll-fno') is DGFSVHSGLLFNOScatteringModel
