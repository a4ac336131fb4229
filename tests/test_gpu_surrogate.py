"""Learned stand-in on the GPU, through the C ABI of include/frfs_sur.h:
forward and backward kernels against the float64 oracle (which is pinned by
numpy.fft and finite differences on the CPU), the loss and Adam kernels, the
AoSoA gather and the conservation check / scatter, the fallback to the
spectral operator, and what training gains over the untrained network.

Tolerances: the kernels compute in FP32 with FP32 sums of at most 512 terms
per stage, so the network output agrees with the oracle to rel-L2 <= 1e-5
per DOF and the gradient to rel-L2 <= 1e-4 per parameter block.
"""
import numpy as np
import pytest

from oracle import surrogate as osur
from tests.cases import rel_l2, vhs_cfg

pytestmark = pytest.mark.gpu

N, W, M = 16, 32, 8
FWD_TOL = 1e-5
GRAD_TOL = 1e-4


def _grid(L=5.0):
    c0 = np.linspace(-L + L / N, L - L / N, N)
    idx = np.arange(N ** 3)
    return np.stack([c0[idx // (N * N)], c0[(idx // N) % N], c0[idx % N]]), (2 * L / N) ** 3


def _theta(L, seed=3, r_scale=8.0):
    """Seeded parameters, Fourier weights scaled up so that the spectral path
    carries a share of the output comparable to the pointwise one; rounded
    to float32 so the oracle sees the parameters the kernels see."""
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


def test_forward_matches_oracle(backend):
    from frfs_amd.surrogate import FourierSurrogate
    L = 4
    th = _theta(L)
    g = _inputs(3, seed=21)
    net = FourierSurrogate(backend, layers=L, theta=th)
    print('[surrogate forward L=4 B=3]', flush=True)
    out = net.predict(_dev(g), fused=False).cpu().numpy()
    want = osur.forward(th, g, N, W, M, L)
    for b in range(3):
        assert rel_l2(out[b], want[b]) <= FWD_TOL, b
    # the training forward (which keeps activations) gives the same bits
    import torch
    o2 = torch.empty_like(_dev(g))
    net.forward(_dev(g), o2, training=True)
    assert np.array_equal(o2.cpu().numpy(), out)
    # inference with the lift and projection folded into the first and last
    # layers: the same network up to FP32 rounding, against the oracle too
    of = net.predict(_dev(g)).cpu().numpy()
    for b in range(3):
        assert rel_l2(of[b], out[b]) <= 2e-6, b
        assert rel_l2(of[b], want[b]) <= FWD_TOL, b


@pytest.mark.parametrize('L', [2, 3])
def test_fused_inference_shallow(backend, L):
    """Two and three layers: the folded first and last layers are adjacent
    (L = 2) or around one ordinary layer (L = 3); the folded weights follow
    theta after a training step."""
    import torch
    from frfs_amd.surrogate import FourierSurrogate
    th = _theta(L, seed=40 + L)
    g = _inputs(2, seed=50 + L)
    net = FourierSurrogate(backend, layers=L, theta=th)
    print(f'[surrogate fused L={L}]', flush=True)
    of = net.predict(_dev(g)).cpu().numpy()
    want = osur.forward(th, g, N, W, M, L)
    for b in range(2):
        assert rel_l2(of[b], want[b]) <= FWD_TOL, b
    net.train_step(_dev(g), torch.zeros_like(_dev(g)), 1e-3)
    th2 = net.theta.double().cpu().numpy()
    of2 = net.predict(_dev(g)).cpu().numpy()
    want2 = osur.forward(th2, g, N, W, M, L)
    assert not np.array_equal(of2, of)
    for b in range(2):
        assert rel_l2(of2[b], want2[b]) <= FWD_TOL, b


@pytest.mark.parametrize('L,acts,tol', [(1, 'fp32', GRAD_TOL), (4, 'fp32', GRAD_TOL), (4, 'bf16', 5e-2)])
def test_backward_matches_oracle(backend, L, acts, tol):
    """FP32 activation storage: every parameter block within 1e-4 of the
    oracle.  BF16 storage (training in BF16 with FP32 accumulation): within
    5e-2, the rounding of 8 significant bits carried through four layers."""
    import torch
    from frfs_amd.surrogate import FourierSurrogate
    th = _theta(L, seed=L)
    g = _inputs(2, seed=30 + L)
    rng = np.random.default_rng(L)
    go = rng.standard_normal(g.shape).astype(np.float32).astype(np.float64)
    net = FourierSurrogate(backend, layers=L, theta=th)
    net.train_acts = acts
    print(f'[surrogate backward L={L} {acts}]', flush=True)
    gd = _dev(g)
    out = torch.empty_like(gd)
    acts_buf = net.forward(gd, out, training=True)
    grad = torch.zeros(net.nparams, dtype=torch.float32, device='cuda')
    net.backward(gd, acts_buf, _dev(go), grad)
    got = grad.cpu().numpy().astype(np.float64)
    want = osur.backward(th, g, go, N, W, M, L)
    if acts != 'fp32':
        fw = osur.forward(th, g, N, W, M, L)
        for b in range(2):
            assert rel_l2(out.cpu().numpy()[b], fw[b]) <= tol, b
    off = osur.layout(W, M, L)
    blocks = [('P', W), ('p0', W), ('Qw', W), ('q0', 1)]
    blocks += [(f'R{l}', 2 * W * W * M ** 3) for l in range(L)]
    blocks += [(f'W{l}', W * W) for l in range(L)] + [(f'b{l}', W) for l in range(L)]
    for name, size in blocks:
        a = off[name]
        assert rel_l2(got[a:a + size], want[a:a + size]) <= tol, name
    # the gradient accumulates into a caller-zeroed vector
    net.backward(gd, acts_buf, _dev(go), grad)
    assert rel_l2(grad.cpu().numpy(), 2 * got) <= 1e-6


def test_loss_and_adam_kernels(backend):
    import ctypes as C
    import torch
    from frfs_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(5)
    out = rng.standard_normal(5000).astype(np.float32)
    tgt = rng.standard_normal(5000).astype(np.float32)
    go = torch.empty(5000, dtype=torch.float32, device='cuda')
    sums = torch.zeros(2, dtype=torch.float64, device='cuda')
    print('[surrogate loss/adam]', flush=True)
    # device copies are held in names: a temporary's memory may be reused before the kernel runs
    od, td = _dev(out), _dev(tgt)
    _lib.check(lib.fq_sur_loss(5000, od.data_ptr(), td.data_ptr(), go.data_ptr(), sums.data_ptr(), None))
    d = (out - tgt).astype(np.float64)        # the kernel forms the difference in FP32
    assert np.allclose(go.cpu().numpy(), 2 * d / 5000, rtol=1e-6, atol=0)
    s = sums.cpu().numpy()
    assert abs(s[0] / np.sum(d * d) - 1) < 1e-12 and abs(s[1] / np.sum(tgt.astype(np.float64) ** 2) - 1) < 1e-12
    th = rng.standard_normal(1000).astype(np.float32)
    gr = rng.standard_normal(1000).astype(np.float32)
    m = np.zeros(1000, np.float32)
    v = np.zeros(1000, np.float32)
    tt, mt, vt, gt = _dev(th), _dev(m), _dev(v), _dev(gr)
    for step in (1, 2):
        _lib.check(lib.fq_sur_adam(1000, tt.data_ptr(), gt.data_ptr(), mt.data_ptr(), vt.data_ptr(),
                                   C.c_float(1e-2), C.c_float(0.9), C.c_float(0.999), C.c_float(1e-8), step, None))
    # numpy Adam, two steps with the same gradient
    th64, m64, v64 = th.astype(np.float64), np.zeros(1000), np.zeros(1000)
    for step in (1, 2):
        m64 = 0.9 * m64 + 0.1 * gr
        v64 = 0.999 * v64 + 0.001 * gr.astype(np.float64) ** 2
        th64 -= 1e-2 * (m64 / (1 - 0.9 ** step)) / (np.sqrt(v64 / (1 - 0.999 ** step)) + 1e-8)
    assert np.max(np.abs(tt.cpu().numpy() - th64)) < 1e-6


def test_gather_finish_against_oracle(backend):
    """Gather from an AoSoA bank (2 points x 3 elements), then the
    conservation check and scatter with a tolerance that flags some DOFs."""
    import torch
    from frfs_amd import _lib
    from frfs_amd.inputs import shear_mixture_device, shear_mixture_params, to_aosoa
    from frfs_amd.surrogate import F_REF, gram
    lib = _lib.load()
    cv, cw = _grid()
    nupts, neles = 2, 3
    nd = nupts * neles
    f = shear_mixture_device(cv, shear_mixture_params(nd, 8), device='cuda')
    fin = backend.matrix((nupts, N ** 3, neles))
    to_aosoa(f, fin)
    g = torch.empty((nd, N ** 3), dtype=torch.float32, device='cuda')
    rho = torch.empty(nd, dtype=torch.float64, device='cuda')
    print('[surrogate gather/finish]', flush=True)
    _lib.check(lib.fq_sur_gather(N, cw, F_REF, 1, fin._as_parameter_, nupts, fin.leaddim, neles, 0, nd, None,
                                 rho.data_ptr(), g.data_ptr(), None))
    fh = f.cpu().numpy()
    rh = fh.sum(axis=1) * cw
    assert np.allclose(rho.cpu().numpy(), rh, rtol=1e-13)
    assert np.allclose(g.cpu().numpy(), fh / (rh[:, None] * F_REF), rtol=1e-6)
    # network-like outputs with a controlled share along the invariants
    rng = np.random.default_rng(2)
    phi = osur.invariants(cv)
    q = np.array([osur.defect_and_project(x, cv, cw)[1] for x in rng.standard_normal((nd, N ** 3))])
    for b in range(nd):
        q[b] += 0.1 ** (b + 1) * np.linalg.norm(q[b]) * phi[b % 5] / np.linalg.norm(phi[b % 5])
    q = q.astype(np.float32)
    tol = 3e-3
    gi, gd = gram(cv, cw)
    dst = backend.matrix((nupts, N ** 3, neles), np.full((nupts, N ** 3, neles), 7.0))
    dfc = torch.empty(nd, dtype=torch.float64, device='cuda')
    flag = torch.empty(nd, dtype=torch.int32, device='cuda')
    qd, gid, gdd = _dev(q), _dev(gi.reshape(-1), torch.float64), _dev(gd, torch.float64)
    _lib.check(lib.fq_sur_finish(N, 5.0, cw, 2.5, tol, qd.data_ptr(), rho.data_ptr(), gid.data_ptr(), gdd.data_ptr(),
                                 dst._as_parameter_, nupts, dst.leaddim, neles, 0, nd, dfc.data_ptr(),
                                 flag.data_ptr(), None))
    out = dst.get()
    dd, ff = dfc.cpu().numpy(), flag.cpu().numpy()
    assert 0 < ff.sum() < nd
    for b in range(nd):
        e, u = divmod(b, nupts)
        d, qp = osur.defect_and_project(q[b].astype(np.float64), cv, cw)
        assert abs(dd[b] - d) <= 1e-12 + 1e-9 * d, b
        assert ff[b] == int(d > tol)
        if ff[b]:
            assert np.all(out[u, :, e] == 7.0)
        else:
            assert rel_l2(out[u, :, e], rh[b] ** 2 * 2.5 * qp) <= 1e-12


def _fno_model(backend, tol):
    from frfs_amd.surrogate import DGFSVHSGLLFNOScatteringModel
    from frfs_amd.velocitymesh import DGFSVelocityMesh
    cfg = vhs_cfg(16, 8, 6, 5.0, model='vhs-gll-fno')
    cfg.set('scattering-model', 'surrogate-tol', tol)
    vm = DGFSVelocityMesh(backend, cfg)
    return DGFSVHSGLLFNOScatteringModel(backend, cfg, vm), vm


def test_fallback_and_conservation(backend):
    """tol = 0: every DOF falls back and the result is the spectral
    operator's, bit for bit; tol = 2 (above any cosine): no DOF falls back and
    the result is the conservation-projected network output."""
    import torch
    from frfs_amd import _lib
    from frfs_amd.inputs import shear_mixture_device, shear_mixture_params, to_aosoa
    from frfs_amd.surrogate import F_REF
    sm, vm = _fno_model(backend, 0.0)
    nupts, neles = 3, 5
    nd = nupts * neles
    f = shear_mixture_device(vm.cv(), shear_mixture_params(nd, 12), device='cuda')
    fin = backend.matrix((nupts, vm.vsize(), neles))
    to_aosoa(f, fin)
    spec = backend.matrix(fin.ioshape)
    sm.fs_spectral(fin, spec)
    q0 = backend.matrix(fin.ioshape)
    print('[surrogate fallback]', flush=True)
    st = sm.fs_batch(fin, q0)
    torch.cuda.synchronize()
    assert st['fallback'] == nd and np.array_equal(q0.get(), spec.get())
    sm.tol = 2.0
    sm.check_tol = 1e30              # the sampled spectral check always passes
    q1 = backend.matrix(fin.ioshape)
    st = sm.fs_batch(fin, q1)
    assert st['fallback'] == 0 and len(st['check_rel_l2_median']) == 1
    st_defect = st['defect'].copy()
    cv, cw = vm.cv(), vm.cw()
    # the network's own input (the gather kernel is checked against NumPy above)
    g = torch.empty((nd, vm.vsize()), dtype=torch.float32, device='cuda')
    rd = torch.empty(nd, dtype=torch.float64, device='cuda')
    _lib.check(_lib.load().fq_sur_gather(N, cw, F_REF, 1, fin._as_parameter_, nupts, fin.leaddim, neles, 0, nd,
                                         None, rd.data_ptr(), g.data_ptr(), None))
    rho = rd.cpu().numpy()
    qn = sm.net.predict(g).cpu().numpy().astype(np.float64)
    got = q1.get()
    phi = osur.invariants(cv)
    for b in range(nd):
        e, u = divmod(b, nupts)
        d, qp = osur.defect_and_project(qn[b], cv, cw)
        assert abs(st['defect'][b] - d) <= 1e-12 + 1e-9 * d
        assert rel_l2(got[u, :, e], rho[b] ** 2 * sm.net.q_ref * qp) <= 1e-12
        mom = phi @ got[u, :, e] * cw
        assert np.max(np.abs(mom)) <= 1e-10 * np.abs(got[u, :, e]).sum() * cw * 25
    # the sampled spectral check fails (tolerance 0): the whole batch is the
    # spectral operator's, bit for bit, with every DOF counted as a fallback
    sm.check_tol = 0.0
    sm.check_every = 4
    q3 = backend.matrix(fin.ioshape)
    st = sm.fs_batch(fin, q3)
    assert st['fallback'] == nd and st['check_failed_chunks'] == 1 and st['check_rel_l2_median'][0] > 0
    assert np.array_equal(q3.get(), spec.get())
    sm.check_tol = 1e30
    # a tolerance between the DOFs' defects flags a scattered subset: the
    # flagged DOFs, gathered into one compact spectral batch, get the spectral
    # operator's result bit for bit and the others keep the network's
    sm.check_every = 64
    sm.tol = float(np.median(st_defect))
    sm.inplace_runs = 0
    q4 = backend.matrix(fin.ioshape)
    st = sm.fs_batch(fin, q4)
    bad = st_defect > sm.tol
    assert 0 < bad.sum() < nd and st['fallback'] == bad.sum()
    a4, a1, asp = q4.get(), q1.get(), spec.get()
    for b in range(nd):
        e, u = divmod(b, nupts)
        assert np.array_equal(a4[u, :, e], (asp if bad[b] else a1)[u, :, e]), b
    sm.tol, sm.inplace_runs = 2.0, 4
    # a DOF sub-range writes only its DOFs
    q2 = backend.matrix(fin.ioshape, np.full(fin.ioshape, -1.0))
    sm.fs_batch(fin, q2, first_dof=4, ndof=6)
    a, bq = q2.get(), q1.get()
    for b in range(nd):
        e, u = divmod(b, nupts)
        if 4 <= b < 10:
            assert np.array_equal(a[u, :, e], bq[u, :, e])
        else:
            assert np.all(a[u, :, e] == -1.0)


@pytest.mark.parametrize('acts', ['fp32', 'bf16'])
def test_training_gain(backend, acts):
    """What the trained network gains over the method without it: 512
    training pairs from the 16^3 vhs-gll operator, 64 held out.  Without the
    network (Q = 0) the held-out rel-L2 is 1, the untrained network's is
    larger; a short training run, with FP32 or BF16 activation storage, must
    bring it well below both."""
    import torch
    from frfs_amd.inputs import shear_mixture_device, shear_mixture_params
    from frfs_amd.surrogate import training_pairs
    sm, vm = _fno_model(backend, 1e-4)
    sm.net.train_acts = acts
    f = shear_mixture_device(vm.cv(), shear_mixture_params(576, 40), device='cuda')
    g, t, _ = training_pairs(sm, f)
    qref = float(t[:512].pow(2).mean().sqrt())
    t /= qref
    sm.net.q_ref = qref

    def held_out():
        p = sm.net.predict(g[512:])
        return float(((p - t[512:]).pow(2).sum(1) / t[512:].pow(2).sum(1)).sqrt().median())
    before = held_out()
    print(f'[surrogate training {acts}] untrained held-out median rel-L2 {before:.3e}', flush=True)
    hist = sm.net.fit(g[:512], t[:512], epochs=40, batch=32, lr=2e-3, seed=1)
    after = held_out()
    print(f'[surrogate training {acts}] trained held-out median rel-L2 {after:.3e} '
          f'(train rel. sq. error {hist[0]:.3e} -> {hist[-1]:.3e})', flush=True)
    assert hist[-1] < 0.1 * hist[0]
    assert after < 0.5 and after < 0.5 * before
    torch.cuda.synchronize()


@pytest.mark.parametrize('acts,tol', [('fp16', 2e-3), ('bf16', 3e-2)])
def test_16bit_activation_storage(backend, acts, tol):
    """Folded inference with the activations stored in 16 bits between
    kernels: within tol (rel-L2 per DOF) of the FP32 storage and 1.5 tol of
    the oracle (FP16 keeps 11 significant bits, BF16 8)."""
    from frfs_amd.surrogate import FourierSurrogate
    L = 4
    th = _theta(L, seed=17)
    g = _inputs(3, seed=27)
    net = FourierSurrogate(backend, layers=L, theta=th)
    print(f'[surrogate {acts} activations]', flush=True)
    o32 = net.predict(_dev(g), acts='fp32').cpu().numpy()
    o16 = net.predict(_dev(g), acts=acts).cpu().numpy()
    want = osur.forward(th, g, N, W, M, L)
    for b in range(3):
        assert 0 < rel_l2(o16[b], o32[b]) <= tol, b
        assert rel_l2(o16[b], want[b]) <= 1.5 * tol, b


@pytest.mark.parametrize('NG', [16, 32])
def test_conservation_penalty_gradient(backend, NG):
    """fq_sur_cons adds 2 lambda u / count to the loss gradient, u the
    orthogonal projection of each output row onto the collision invariants
    {1, cx, cy, cz, |c|^2} of the grid (computed here in float64 NumPy by least
    squares), and sum |u|^2 to sums[2]."""
    import ctypes as C
    import torch
    from frfs_amd import _lib
    from frfs_amd.surrogate import gram
    L0, B, lam = 5.0, 3, 7.5
    c0 = np.linspace(-L0 + L0 / NG, L0 - L0 / NG, NG)
    idx = np.arange(NG ** 3)
    cv = np.stack([c0[idx // (NG * NG)], c0[(idx // NG) % NG], c0[idx % NG]])
    cw = (2 * L0 / NG) ** 3
    gi, _ = gram(cv, cw)
    rng = np.random.default_rng(NG)
    out = rng.standard_normal((B, NG ** 3)).astype(np.float32)
    # gout holds an earlier gradient: a small one, so that the FP32 sum keeps
    # the increment's digits (the increment is ~1e-4 of a unit-size row)
    base = (1e-4 * rng.standard_normal((B, NG ** 3))).astype(np.float32)
    phi = np.stack([np.ones(NG ** 3), cv[0], cv[1], cv[2], (cv ** 2).sum(axis=0)])
    coef, *_ = np.linalg.lstsq(phi.T, out.astype(np.float64).T, rcond=None)
    u = (phi.T @ coef).T                                     # (B, n) projections
    od, gd = _dev(out), _dev(base)
    sums = torch.zeros(3, dtype=torch.float64, device='cuda')
    ginv = torch.tensor(gi.reshape(-1), dtype=torch.float64, device='cuda')
    print(f'[conservation penalty {NG}^3]', flush=True)
    _lib.check(_lib.load().fq_sur_cons(NG, L0, cw, lam, B, od.data_ptr(), ginv.data_ptr(), gd.data_ptr(),
                                       sums.data_ptr(), None))
    torch.cuda.synchronize()
    want = base + 2 * lam / (B * NG ** 3) * u
    got = gd.cpu().numpy()
    assert rel_l2(got - base, want - base) <= 1e-5
    assert abs(float(sums[2]) - float((u ** 2).sum())) <= 1e-6 * float((u ** 2).sum())
