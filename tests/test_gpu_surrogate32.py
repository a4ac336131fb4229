"""The learned stand-in on the 32^3 grid, through the C ABI of
include/frfs_sur.h: forward (unfolded and folded) and backward kernels
against the float64 oracle (pinned by numpy.fft and finite differences on
the CPU), 16-bit activation storage, and the vss-gll-fno plugin around them
(tolerance 0 gives the spectral operator bit for bit; a short training run
gains over the untrained network).  Tolerances as on the 16^3 grid: FP32
arithmetic with sums of at most 512 terms per transform stage."""
import numpy as np
import pytest

from oracle import surrogate as osur
from tests.cases import rel_l2, vss_cfg

pytestmark = pytest.mark.gpu

N, W, M = 32, 32, 8
FWD_TOL = 1e-5
GRAD_TOL = 1e-4


def _grid(L=6.0):
    c0 = np.linspace(-L + L / N, L - L / N, N)
    idx = np.arange(N ** 3)
    return np.stack([c0[idx // (N * N)], c0[(idx // N) % N], c0[idx % N]]), (2 * L / N) ** 3


def _theta(L, seed=3, r_scale=8.0):
    from frfs_amd.surrogate import init_params
    th = init_params(L, seed)
    off = osur.layout(W, M, L)
    for l in range(L):
        th[off[f'R{l}']:off[f'R{l}'] + 2 * W * W * M ** 3] *= r_scale
    return th.astype(np.float32).astype(np.float64)


def _inputs(B, seed):
    from frfs_amd.inputs import shear_mixture_device, shear_mixture_params
    from frfs_amd.surrogate import F_REF
    cv, cw = _grid()
    f = shear_mixture_device(cv, shear_mixture_params(B, seed), device='cpu').numpy()
    rho = f.sum(axis=1) * cw
    return (f / (rho[:, None] * F_REF)).astype(np.float32).astype(np.float64)


def _dev(a, dtype=None):
    import torch
    return torch.tensor(a, dtype=dtype or torch.float32, device='cuda')


def test_forward32_matches_oracle(backend):
    from frfs_amd.surrogate import FourierSurrogate
    L = 4
    th = _theta(L)
    g = _inputs(2, seed=21)
    net = FourierSurrogate(backend, layers=L, theta=th, nv=N)
    print('[surrogate32 forward L=4 B=2]', flush=True)
    out = net.predict(_dev(g), fused=False).cpu().numpy()
    want = osur.forward(th, g, N, W, M, L)
    for b in range(2):
        assert rel_l2(out[b], want[b]) <= FWD_TOL, b
    of = net.predict(_dev(g)).cpu().numpy()
    for b in range(2):
        assert rel_l2(of[b], want[b]) <= FWD_TOL, b
    for acts, tol in (('fp16', 2e-3), ('bf16', 3e-2)):
        o16 = net.predict(_dev(g), acts=acts).cpu().numpy()
        for b in range(2):
            assert 0 < rel_l2(o16[b], of[b]) <= tol, (acts, b)


def test_backward32_matches_oracle(backend):
    import torch
    from frfs_amd.surrogate import FourierSurrogate
    L = 2
    th = _theta(L, seed=7)
    g = _inputs(2, seed=31)
    rng = np.random.default_rng(5)
    go = rng.standard_normal(g.shape).astype(np.float32).astype(np.float64)
    net = FourierSurrogate(backend, layers=L, theta=th, nv=N)
    print('[surrogate32 backward L=2 B=2]', flush=True)
    gd = _dev(g)
    out = torch.empty_like(gd)
    acts_buf = net.forward(gd, out, training=True)
    grad = torch.zeros(net.nparams, dtype=torch.float32, device='cuda')
    net.backward(gd, acts_buf, _dev(go), grad)
    got = grad.cpu().numpy().astype(np.float64)
    want = osur.backward(th, g, go, N, W, M, L)
    off = osur.layout(W, M, L)
    blocks = [('P', W), ('p0', W), ('Qw', W), ('q0', 1)]
    blocks += [(f'R{l}', 2 * W * W * M ** 3) for l in range(L)]
    blocks += [(f'W{l}', W * W) for l in range(L)] + [(f'b{l}', W) for l in range(L)]
    for name, size in blocks:
        a = off[name]
        assert rel_l2(got[a:a + size], want[a:a + size]) <= GRAD_TOL, name


def _plugin(backend, tol, weights=''):
    from frfs_amd.surrogate import DGFSBiVSSGLLFNOScatteringModel
    from frfs_amd.velocitymesh import DGFSBiVelocityMesh
    cfg = vss_cfg(32, 2, 12, 6.0, project=1)
    cfg.set('scattering-model', 'type', 'vss-gll-fno')
    cfg.set('scattering-model', 'surrogate-tol', str(tol))
    if weights:
        cfg.set('scattering-model', 'surrogate-weights', weights)
    vm = DGFSBiVelocityMesh(backend, cfg, 1)
    return vm, DGFSBiVSSGLLFNOScatteringModel(backend, cfg, vm)


def test_vss_plugin_tolerance_zero_is_spectral(backend):
    """With tolerance 0 every DOF falls back: the plugin's output is the
    spectral operator's, bit for bit."""
    import torch
    from frfs_amd.inputs import mixture_device, mixture_params, to_aosoa
    vm, sm = _plugin(backend, 0.0)
    n = vm.vsize()
    fin = backend.matrix((1, n, 5))
    to_aosoa(mixture_device(vm.cv(), mixture_params(5, 9)), fin)
    qa, qb = backend.matrix(fin.ioshape), backend.matrix(fin.ioshape)
    print('[vss-gll-fno tolerance 0]', flush=True)
    st = sm.fs_batch(fin, qa)
    sm.fs_spectral(fin, qb)
    torch.cuda.synchronize()
    assert st['fallback'] == 5
    assert np.array_equal(qa.get(), qb.get())


def test_vss_training_gain(backend):
    """What the trained network gains over the method without it: 512 pairs
    of the projected FP64 vss-gll operator (32^3, M = 12, Nrho = 2), 64 held
    out, 40 epochs (as on the 16^3 grid) take the held-out median rel-L2 well
    below the untrained network's and below Q = 0's (which scores 1)."""
    from frfs_amd import train_surrogate as ts
    print('[vss-gll-fno training gain]', flush=True)
    sm, rep = ts.train(backend, grid='vss32', pairs=512, heldout=64, epochs=40, batch=32, lr=2e-3,
                       seed=1, nrho=2, log=None)
    print(f"[vss-gll-fno training gain] held-out median rel-L2 {rep['heldout_rel_l2_untrained_median']:.3e} "
          f"-> {rep['heldout_rel_l2_median']:.3e}", flush=True)
    assert rep['heldout_rel_l2_untrained_median'] > 0.5
    assert rep['heldout_rel_l2_median'] < 0.5 and \
        rep['heldout_rel_l2_median'] < 0.5 * rep['heldout_rel_l2_untrained_median']
