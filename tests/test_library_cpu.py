"""CPU tests of the C-ABI library (no GPU calls) and of the host side.

* the shared library loads and exports every function include/frfs_fq.h
  declares;
* the host-only plan: work counts, and a NumPy execution of the exported
  plan (separable phases, deduplicated transforms, shared products) that
  must reproduce the oracle -- this checks the restructured algorithm the
  kernels implement, independently of the GPU;
* argument validation and error reporting;
* velocity mesh / quadrature / parameters / AoSoA layout against the oracle.
"""
import os
import re

import numpy as np
import pytest

from frfs_amd import _lib
from oracle import fastspectral as fsp
from tests.cases import ARGON, oracle_vhs, oracle_vss, rel_l2, vhs_cfg, vss_cfg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = ''.join(open(os.path.join(ROOT, 'include', h)).read() for h in ('frfs_fq.h', 'frfs_sur.h'))
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(fq_\w+)\s*\(', txt)))


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    names = _header_functions()
    assert len(names) >= 22
    for nm in names:
        assert hasattr(lib, nm), nm
        assert nm in _lib.SIGNATURES, nm


def _vhs_params(cfg, model=_lib.MODEL_VHS, precision=_lib.PREC_FP64):
    from frfs_amd.scattering import DGFSVHSGLLScatteringModel
    from frfs_amd.velocitymesh import DGFSVelocityMesh

    class Host(DGFSVHSGLLScatteringModel):
        def perform_precomputation(self):
            self.params = self._vhs_params(model)

    vm = DGFSVelocityMesh(None, cfg)
    return Host(None, cfg, vm).params, vm


def _vss_params(cfg, ns):
    from frfs_amd.scattering import DGFSBiVSSGLLScatteringModel
    from frfs_amd.velocitymesh import DGFSBiVelocityMesh

    class Host(DGFSBiVSSGLLScatteringModel):
        def perform_precomputation(self):
            vm = self.vm
            self.params = _lib.Params(
                model=_lib.MODEL_VSS, Nv=vm.Nv(), Nrho=vm.Nrho(), M=vm.M(), L=vm.L(),
                sw=vm.sw(), cw=vm.cw(), qz=vm.qz(), qw=vm.qw(), sz=vm.sz(),
                masses=np.asarray(vm.masses(), float), eta=self._eta, gammapq=self._gamma,
                prefacpq=self._prefactor, szpre=self._szpre, swpre=self._swpre)

    vm = DGFSBiVelocityMesh(None, cfg, ns)
    return Host(None, cfg, vm).params, vm


def test_plan_counts():
    p, _ = _vhs_params(vhs_cfg(16, 8, 6, 5.0))
    i = _lib.Plan(p).info()
    assert (i['n_items'], i['n_products'], i['n_groups'], i['n_fft_ref'], i['n_fft']) == (96, 48, 8, 147, 107)
    p, _ = _vss_params(vss_cfg(32, 16, 12, 6.0), 1)
    i = _lib.Plan(p).info()
    # antipodal directions share inverse transforms and products
    assert (i['n_items'], i['n_groups'], i['n_fft_ref'], i['n_fft']) == (192, 96, 580, 291)
    m0 = ARGON['molarMass0']
    p, _ = _vss_params(vss_cfg(32, 16, 12, 6.0, masses=(m0, 2 * m0)), 2)
    i = _lib.Plan(p).info()
    assert i['n_sources'] == 2 and i['n_fft_ref'] == 4 * 580 and i['n_fft'] == 1158


def _packed_pair(T1, T2, th1, th2, N):
    """Both item transforms of a product the way the pair-packed product
    kernel forms them: one DFT+ of T1 + i T2, and the Nyquist-plane
    corrections of k_nyq (three 2D DFT+ of the packed anti-Hermitian part,
    on the planes K = N/2; J = N/2, K != N/2; I = N/2, J, K != N/2)."""
    h = N // 2
    dft2 = lambda x: np.fft.ifft2(x) * x.size
    E = np.fft.ifftn(T1 + 1j * T2) * T1.size
    A = 0
    for t, (T, th) in enumerate(((T1, th1), (T2, th2))):
        w = np.ones((N, N, N), complex)
        w[h, :, :] *= np.exp(1j * N * th[0])
        w[:, h, :] *= np.exp(1j * N * th[1])
        w[:, :, h] *= np.exp(1j * N * th[2])
        A = A + (1j ** t) * T * (1 - w) / 2
    xk = dft2(A[:, :, h])                                      # (vx, vy)
    aj = A[:, h, :].copy(); aj[:, h] = 0
    xj = dft2(aj)                                              # (vx, vz)
    ai = A[h, :, :].copy(); ai[h, :] = 0; ai[:, h] = 0
    xi = dft2(ai)                                              # (vy, vz)
    sg = (-1.0) ** np.arange(N)
    U = (sg[None, None, :] * xk[:, :, None] + sg[None, :, None] * xj[:, None, :]
         + sg[:, None, None] * xi[None, :, :])
    R = E - U
    return R.real + 1j * U.imag, R.imag - 1j * U.real


def _emulate(plan, F, weight_of_contrib, loss_tables, f_src, N, pack=False):
    """NumPy execution of an exported plan (complex128); with pack=True the
    item transforms of every product come from _packed_pair."""
    n = N ** 3
    l0 = np.concatenate((np.arange(0, N // 2), np.arange(-N // 2, 0))).astype(float)
    T, E = [], []
    for s, th in zip(plan['item_src'], plan['item_theta']):
        ph = np.exp(1j * (th[0] * l0[:, None, None] + th[1] * l0[None, :, None] + th[2] * l0[None, None, :]))
        T.append(ph * F[s].reshape(N, N, N))
        E.append(np.fft.ifftn(T[-1]) * n)
    ngroups = plan['prod_group'].max() + 1
    G = [np.zeros((N, N, N), complex) for _ in range(ngroups)]
    for i1, i2, cf, g in zip(plan['prod_i1'], plan['prod_i2'], plan['prod_coef'], plan['prod_group']):
        if pack:
            e1, e2 = _packed_pair(T[i1], T[i2], plan['item_theta'][i1], plan['item_theta'][i2], N)
            assert np.allclose(e1, E[i1], rtol=0, atol=1e-12) and np.allclose(e2, E[i2], rtol=0, atol=1e-12)
            G[g] += cf * e1 * e2
        else:
            G[g] += cf * E[i1] * E[i2]
    Gh = [np.fft.fftn(g).ravel() for g in G]
    ntgt = plan['contrib_target'].max() + 1
    H = [np.zeros(n, complex) for _ in range(ntgt)]
    Lt = [np.zeros(n, complex) for _ in range(ntgt)]
    for c, (g, t) in enumerate(zip(plan['contrib_group'], plan['contrib_target'])):
        H[t] += weight_of_contrib(c) * Gh[g]
    for j, (s, t) in enumerate(zip(plan['loss_src'], plan['loss_target'])):
        Lt[t] += loss_tables[j] * F[s]
    out = []
    for t in range(ntgt):
        gain = np.real(np.fft.ifftn(H[t].reshape(N, N, N)) * n).ravel()
        loss = np.real(np.fft.ifftn(Lt[t].reshape(N, N, N)) * n).ravel()
        out.append(gain - f_src[t] * loss)
    return out


@pytest.mark.parametrize('pack', [False, True])
def test_plan_execution_matches_oracle_vhs(pack):
    cfg = vhs_cfg(8, 4, 6, 4.0)
    p, _ = _vhs_params(cfg)
    plan = _lib.Plan(p).export()
    vm, prm, tab = oracle_vhs(cfg)
    N, n = 8, 512
    f = fsp.two_maxwellian_batch(vm['cv'], 1, seed=8)[0]
    F = [np.fft.fftn(f.reshape(N, N, N)).ravel() / n]
    w = lambda c: prm['prefac'] * (2 * vm['qw'][c] * vm['sw'] / n) * tab['bb1'][c]
    q = _emulate(plan, F, w, [prm['prefac'] * tab['bb2']], [f], N, pack)[0]
    assert rel_l2(q, fsp.vhs_fs(f, vm, tab, prm['prefac'])) < 1e-13


@pytest.mark.parametrize('pack', [False, True])
@pytest.mark.parametrize('ns', [1, 2])
def test_plan_execution_matches_oracle_vss(ns, pack):
    m0 = ARGON['molarMass0']
    cfg = vss_cfg(8, 3, 12, 4.0, masses=(m0, 2 * m0)[:ns], Mpre=32)
    p, _ = _vss_params(cfg, ns)
    plan = _lib.Plan(p).export()
    vm, prm, tab = oracle_vss(cfg, ns)
    N, n = 8, 512
    fs = [fsp.two_maxwellian_batch(vm['cv'], 1, seed=20 + s, mass=vm['masses'][s])[0] for s in range(ns)]
    F = [np.fft.fftn(f.reshape(N, N, N)).ravel() / n for f in fs]
    # weight of each contribution: recover its (pair, x, y) terms by locating the
    # two inverse transforms of every reference product in the exported items
    key = {(int(s), *map(float, th)): i for i, (s, th) in enumerate(zip(plan['item_src'], plan['item_theta']))}
    prod_of = {}
    for g, (i1, i2) in enumerate(zip(plan['prod_i1'], plan['prod_i2'])):
        prod_of[(min(i1, i2), max(i1, i2))] = plan['prod_group'][g]
    contrib_of = {(int(g), int(t)): c for c, (g, t) in enumerate(zip(plan['contrib_group'], plan['contrib_target']))}
    W = {}
    ms = p.c.masses
    for pp in range(ns):
        for qq in range(ns):
            mp, mq = ms[pp], ms[qq]
            kq, kp = mq / (mp + mq), mp / (mp + mq)
            for x in range(vm['Nrho']):
                c = np.pi / vm['L'] * vm['qz'][x]
                for y in range(vm['M']):
                    s = [c * vm['sz'][y, d] for d in range(3)]
                    i1 = key[(pp, s[0] * kq, s[1] * kq, s[2] * kq)]
                    i2 = key[(qq, -(s[0] * kp), -(s[1] * kp), -(s[2] * kp))]
                    cidx = contrib_of[(prod_of[(min(i1, i2), max(i1, i2))], pp)]
                    pq = f'{pp}{qq}'
                    W[cidx] = W.get(cidx, 0) + prm['prefac'][pq] * (vm['qw'][x] * vm['sw'] / n) * tab['b'][pq][x, y]
    Lt = [prm['prefac'][f'{t}{s}'] * tab['c'][f'{t}{s}'] for s, t in zip(plan['loss_src'], plan['loss_target'])]
    q = _emulate(plan, F, lambda c: W[c], Lt, fs, N, pack)
    ref = fsp.vss_fs(fs, vm, tab, prm)
    for t in range(ns):
        assert rel_l2(q[t], ref[t]) < 1e-13


def _emulate_group_pairs(plan, F, weight_of_contrib, loss_tables, f_src, N):
    """NumPy execution of the group-pair path of the 32^3 FP32 kernels at any
    N (complex128): two groups' real parts as one volume, the Hermitian part of
    the gain accumulated on rows I' <= N/2 from each pencil and its conjugate
    partner, the imaginary parts entering through their Nyquist-plane sums
    (k_nyqb), rows I' > N/2 as conjugates, a Hermitian loss and one packed
    inverse transform of gain + i loss (k_out1)."""
    n, h = N ** 3, N // 2
    l0 = np.concatenate((np.arange(0, h), np.arange(-h, 0))).astype(float)
    E = []
    for s, th in zip(plan['item_src'], plan['item_theta']):
        ph = np.exp(1j * (th[0] * l0[:, None, None] + th[1] * l0[None, :, None] + th[2] * l0[None, None, :]))
        E.append(np.fft.ifftn(ph * F[s].reshape(N, N, N)) * n)
    ngroups = plan['prod_group'].max() + 1
    G = [np.zeros((N, N, N), complex) for _ in range(ngroups)]
    for i1, i2, cf, g in zip(plan['prod_i1'], plan['prod_i2'], plan['prod_coef'], plan['prod_group']):
        G[g] += cf * E[i1] * E[i2]
    nvol = (ngroups + 1) // 2
    Z = [np.fft.fftn(G[2 * u].real + 1j * (G[2 * u + 1].real if 2 * u + 1 < ngroups else 0)) for u in range(nvol)]
    mi = lambda a: np.roll(a[::-1, ::-1, ::-1], 1, axis=(0, 1, 2))      # a(-xi)
    sg = (-1.0) ** np.arange(N)
    # B_g on the three Nyquist planes from the sums of b = Im G (k_prod1 / k_nyqb)
    BK, BJ, BI = {}, {}, {}
    for g in range(ngroups):
        b = G[g].imag                                                    # [vx][vy][vz]
        BK[g] = np.fft.fft2(np.einsum('xyz,z->xy', b, sg))               # (I', J')
        BJ[g] = np.fft.fft2(np.einsum('xyz,y->xz', b, sg))               # (I', K')
        BI[g] = np.fft.fft2(np.einsum('xyz,x->yz', b, sg))               # (J', K')
    ntgt = plan['contrib_target'].max() + 1
    out = []
    for t in range(ntgt):
        H = np.zeros((N, N, N), complex)
        for c, (g, tt) in enumerate(zip(plan['contrib_group'], plan['contrib_target'])):
            if tt != t:
                continue
            W = weight_of_contrib(c).reshape(N, N, N)
            U, V = (W + np.conj(mi(W))) / 2, 1j * (W - np.conj(mi(W))) / 2
            u, m = divmod(int(g), 2)
            s_, d_ = Z[u] + np.conj(mi(Z[u])), Z[u] - np.conj(mi(Z[u]))
            H += U / 2 * s_ if m == 0 else U / 2j * d_
            # V on the Nyquist set, each point once (k_vtab)
            H[:, :, h] += V[:, :, h] * BK[g]
            mj = np.ones(N, bool); mj[h] = False
            H[:, h, mj] += (V[:, h, :] * BJ[g])[:, mj]
            H[h][np.ix_(mj, mj)] += (V[h] * BI[g])[np.ix_(mj, mj)]
        # rows I' <= N/2 evaluated, the others their conjugates
        H[h + 1:] = 0
        H[h + 1:] = np.conj(mi(H))[h + 1:]
        Lt = np.zeros(n, complex)
        for j, (s, tt) in enumerate(zip(plan['loss_src'], plan['loss_target'])):
            if tt == t:
                Lt += loss_tables[j] * F[s]
        Lt = Lt.reshape(N, N, N)
        Hl = (Lt + np.conj(mi(Lt))) / 2
        y = np.fft.ifftn(H + 1j * Hl) * n
        out.append(y.real.ravel() - f_src[t] * y.imag.ravel())
    return out


@pytest.mark.parametrize('ns', [1, 2])
def test_group_pair_execution_matches_oracle_vss(ns):
    """The group-pair restatement (used by the 32^3 FP32 kernels) against the
    oracle to 1e-13 on inputs with multiplicative noise, where the
    Nyquist-plane terms are a visible part of the gain."""
    m0 = ARGON['molarMass0']
    cfg = vss_cfg(8, 3, 12, 4.0, masses=(m0, 2 * m0)[:ns], Mpre=32)
    p, _ = _vss_params(cfg, ns)
    plan = _lib.Plan(p).export()
    vm, prm, tab = oracle_vss(cfg, ns)
    N, n = 8, 512
    rng = np.random.default_rng(40)
    fs = [fsp.two_maxwellian_batch(vm['cv'], 1, seed=30 + s, mass=vm['masses'][s])[0] for s in range(ns)]
    fs = [f * (1 + 0.3 * rng.standard_normal(f.shape)) for f in fs]
    F = [np.fft.fftn(f.reshape(N, N, N)).ravel() / n for f in fs]
    key = {(int(s), *map(float, th)): i for i, (s, th) in enumerate(zip(plan['item_src'], plan['item_theta']))}
    prod_of = {}
    for g, (i1, i2) in enumerate(zip(plan['prod_i1'], plan['prod_i2'])):
        prod_of[(min(i1, i2), max(i1, i2))] = plan['prod_group'][g]
    contrib_of = {(int(g), int(t)): c for c, (g, t) in enumerate(zip(plan['contrib_group'], plan['contrib_target']))}
    W = {}
    ms = p.c.masses
    for pp in range(ns):
        for qq in range(ns):
            kq, kp = ms[qq] / (ms[pp] + ms[qq]), ms[pp] / (ms[pp] + ms[qq])
            for x in range(vm['Nrho']):
                c = np.pi / vm['L'] * vm['qz'][x]
                for y in range(vm['M']):
                    s = [c * vm['sz'][y, d] for d in range(3)]
                    i1 = key[(pp, s[0] * kq, s[1] * kq, s[2] * kq)]
                    i2 = key[(qq, -(s[0] * kp), -(s[1] * kp), -(s[2] * kp))]
                    cidx = contrib_of[(prod_of[(min(i1, i2), max(i1, i2))], pp)]
                    pq = f'{pp}{qq}'
                    W[cidx] = W.get(cidx, 0) + prm['prefac'][pq] * (vm['qw'][x] * vm['sw'] / n) * tab['b'][pq][x, y]
    Lt = [prm['prefac'][f'{t}{s}'] * tab['c'][f'{t}{s}'] for s, t in zip(plan['loss_src'], plan['loss_target'])]
    q = _emulate_group_pairs(plan, F, lambda c: W[c], Lt, fs, N)
    ref = fsp.vss_fs(fs, vm, tab, prm)
    for t in range(ns):
        assert rel_l2(q[t], ref[t]) < 1e-13


def test_invalid_parameters_raise():
    p, _ = _vhs_params(vhs_cfg(16, 8, 6, 5.0))
    p.c.Nv = 15
    with pytest.raises(_lib.FQError, match='even'):
        _lib.Plan(p)
    p.c.Nv = 16
    p.c.abi_version = 99
    with pytest.raises(_lib.FQError, match='abi_version'):
        _lib.Plan(p)


def test_alpha_below_one_with_antipodal_predesign_raises():
    # the 12-point design contains exact antipodes of the 12 directions: with
    # alpha < 1 the reference's (1 + sigma.sigma')^eta is infinite
    p, _ = _vss_params(vss_cfg(8, 2, 12, 4.0, alpha={'00': 0.8}, Mpre=12), 1)
    with pytest.raises(_lib.FQError, match='Mpre'):
        _lib.Plan(p)


def test_create_without_gpu_reports_error():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip('GPU present')
    except ImportError:
        pass
    p, _ = _vhs_params(vhs_cfg(16, 8, 6, 5.0))
    with pytest.raises(_lib.FQError):
        _lib.Handle(p)


# ---------------------------------------------------------------- host side
def test_velocity_mesh_matches_oracle():
    from frfs_amd.velocitymesh import DGFSBiVelocityMesh, DGFSVelocityMesh
    cfg = vhs_cfg(16, 8, 6, 5.0)
    vm = DGFSVelocityMesh(None, cfg)
    ovm, _, _ = oracle_vhs(cfg)
    for k in ('L', 'S', 'R', 'cw', 'sw'):
        assert getattr(vm, k)() == pytest.approx(ovm[k], rel=1e-15), k
    for k in ('cv', 'qz', 'qw', 'sz'):
        assert np.allclose(getattr(vm, k)(), ovm[k], rtol=1e-14, atol=1e-14), k
    m0 = ARGON['molarMass0']
    cfg = vss_cfg(16, 8, 12, 5.0, masses=(m0, 2 * m0))
    vb = DGFSBiVelocityMesh(None, cfg, 2)
    ob = fsp.velocity_mesh(16, 8, 12, dev=5.0, masses=[m0, 2 * m0], **ARGON)
    for k in ('L', 'S', 'R', 'sw'):
        assert getattr(vb, k)() == pytest.approx(ob[k], rel=1e-15), k
    assert np.allclose(vb.masses(), ob['masses'])
    assert np.allclose(vb.sz(), ob['sz'])


@pytest.mark.parametrize('n', [3, 6, 11])
def test_quadratures_match_oracle_and_radau_exactness(n):
    from frfs_amd import quadratures as q
    for a, b in [(0.0, 0.0), (0.3, -0.4)]:
        z, w = q.zwgj(n, a, b)
        zo, wo = fsp.zwgj(n, a, b)
        assert np.allclose(z, zo, atol=1e-14) and np.allclose(w, wo, atol=1e-14)
        z, w = q.zwglj(n, a, b)
        zo, wo = fsp.zwglj(n, a, b)
        assert np.allclose(z, zo, atol=1e-14) and np.allclose(w, wo, atol=1e-14)
    for rule, end in ((q.zwgrjm, -1.0), (q.zwgrjp, 1.0)):
        z, w = rule(n, 0.0, 0.0)
        assert np.min(np.abs(z - end)) < 1e-13
        for k in range(2 * n - 1):      # Radau: exact to degree 2n-2
            exact = (1 - (-1) ** (k + 1)) / (k + 1)
            assert abs(np.dot(w, z ** k) - exact) < 1e-12


def test_scattering_parameters_match_oracle():
    p, vm = _vhs_params(vhs_cfg(16, 8, 6, 5.0, omega=0.74))
    ovm, oprm, _ = oracle_vhs(vhs_cfg(16, 8, 6, 5.0, omega=0.74))
    assert p.c.prefac == pytest.approx(oprm['prefac'], rel=1e-15)
    assert p.c.gamma == pytest.approx(oprm['gamma'], rel=1e-15)
    m0 = ARGON['molarMass0']
    cfg = vss_cfg(8, 3, 12, 4.0, masses=(m0, 2 * m0))
    p, _ = _vss_params(cfg, 2)
    ovm = fsp.velocity_mesh(8, 3, 12, dev=4.0, masses=[m0, 2 * m0], **ARGON)
    oprm = fsp.vss_parameters(ovm, {k: 1.4 for k in ('00', '01', '10', '11')},
                              {k: 0.74 for k in ('00', '01', '10', '11')},
                              {k: 4.17e-10 for k in ('00', '01', '10', '11')},
                              {k: 273.0 for k in ('00', '01', '10', '11')}, 'womersley', 32)
    pref = np.ctypeslib.as_array(p.c.prefacpq, shape=(4,))
    for i, pq in enumerate(('00', '01', '10', '11')):
        assert pref[i] == pytest.approx(oprm['prefac'][pq], rel=1e-15)


def test_plugin_registry():
    from frfs_amd.scattering import (DGFSBiScatteringModel, DGFSBiVSSGLLScatteringModel,
                                     DGFSScatteringModel, DGFSVHSGLLNodalScatteringModel,
                                     DGFSVHSGLLScatteringModel, subclass_where)
    assert subclass_where(DGFSScatteringModel, scattering_model='vhs-gll') is DGFSVHSGLLScatteringModel
    assert subclass_where(DGFSScatteringModel, scattering_model='vhs-gll-nodal') is DGFSVHSGLLNodalScatteringModel
    assert subclass_where(DGFSBiScatteringModel, scattering_model='vss-gll') is DGFSBiVSSGLLScatteringModel
    with pytest.raises(KeyError):
        subclass_where(DGFSScatteringModel, scattering_model='nope')


def test_aosoa_pack_matches_oracle():
    from frfs_amd.backend import pack, unpack
    rng = np.random.default_rng(1)
    a = rng.normal(size=(2, 7, 45))
    assert np.array_equal(pack(a), fsp.aosoa_pack(a))
    assert np.array_equal(unpack(pack(a), a.shape), a)


def test_input_generator_matches_oracle():
    import torch
    from frfs_amd.inputs import mixture_device, mixture_params
    vm = fsp.velocity_mesh(8, 2, 6, dev=4.0, **ARGON)
    f = fsp.two_maxwellian_batch(vm['cv'], 5, 77)
    g = mixture_device(vm['cv'], mixture_params(5, 77), device='cpu').numpy()
    assert np.max(np.abs(f - g)) < 1e-14 * np.max(np.abs(f))


def test_inifile_defaults():
    from frfs_amd.inifile import Inifile
    ini = Inifile('[a]\nx = 3\n')
    assert ini.getint('a', 'x') == 3
    assert ini.getfloat('a', 'y', 2.5) == 2.5
    with pytest.raises(KeyError):
        ini.get('a', 'z')


def test_grid_range_checked_before_the_device():
    """fq_create refuses grids outside the library path's range (odd, below 4,
    above 128) and FQ_FLAG_NO_SCRATCH off the native grids, before any HIP
    call; the native grids and the even grids in range get past the check."""
    for nv, msg in ((130, 'from 4 to 128'), (2, 'from 4 to 128')):
        p, _ = _vhs_params(vhs_cfg(16, 8, 6, 5.0))
        p.c.Nv = nv
        with pytest.raises(_lib.FQError, match=msg):
            _lib.Handle(p)
    p, _ = _vhs_params(vhs_cfg(16, 8, 6, 5.0))
    p.c.Nv = 20
    p.c.flags |= _lib.FLAG_NO_SCRATCH
    with pytest.raises(_lib.FQError, match='native kernels'):
        _lib.Handle(p)
    p, _ = _vhs_params(vhs_cfg(16, 8, 6, 5.0))
    p.c.flags |= _lib.FLAG_NO_SCRATCH | _lib.FLAG_LIBRARY_FFT
    with pytest.raises(_lib.FQError, match='native kernels'):
        _lib.Handle(p)
