"""GPU parity: the HIP operator (through the C ABI, via the plugin classes)
against the CPU oracle on identical seeded inputs.

Tolerances: FP64 mode rel-L2 <= 1e-12 per DOF; mixed mode rel-L2 <= 1e-5.
"""
import numpy as np
import pytest

from tests.cases import oracle_vhs, oracle_vss, rel_l2, vhs_cfg, vss_cfg

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-12
MIXED_TOL = 1e-5


def _io(f, nupts, neles):
    """(ndof, n) in DOF order elem*nupts + upt -> ioshape (nupts, n, neles)."""
    ndof, n = f.shape
    return f.reshape(neles, nupts, n).transpose(1, 2, 0).copy()


def _from_io(a):
    nupts, n, neles = a.shape
    return a.transpose(2, 0, 1).reshape(neles * nupts, n)


def _run_vhs(backend, cfg, f, nupts, neles, model_cls=None, per_dof=False):
    import torch
    from frfs_amd.scattering import DGFSVHSGLLScatteringModel
    from frfs_amd.velocitymesh import DGFSVelocityMesh
    vm = DGFSVelocityMesh(backend, cfg)
    sm = (model_cls or DGFSVHSGLLScatteringModel)(backend, cfg, vm)
    n = vm.vsize()
    fin = backend.matrix((nupts, n, neles), _io(f, nupts, neles))
    fout = backend.matrix((nupts, n, neles))
    if per_dof:
        for elem in range(neles):
            for upt in range(nupts):
                sm.fs(fin, fout, elem, upt)
    else:
        sm.fs_batch(fin, fout)
    torch.cuda.synchronize()
    return _from_io(fout.get()), sm


@pytest.mark.parametrize('fused', [1, 0])
@pytest.mark.parametrize('prec,tol', [('double', FP64_TOL), ('mixed', MIXED_TOL)])
def test_vhs_config1_shape(backend, prec, tol, fused):
    """VHS omega=0.5, 16^3 on [-5,5]^3, M=6 half sphere, Nrho=8."""
    from oracle import fastspectral as fsp
    cfg = vhs_cfg(16, 8, 6, 5.0, precision=prec, fused=fused)
    vm, prm, tab = oracle_vhs(cfg)
    nupts, neles = 2, 3
    f = fsp.two_maxwellian_batch(vm['cv'], nupts * neles, seed=11)
    print(f'[vhs-config1 {prec}] N=16 ndof={nupts * neles}', flush=True)
    Q, sm = _run_vhs(backend, cfg, f, nupts, neles)
    errs = [rel_l2(Q[d], fsp.vhs_fs(f[d], vm, tab, prm['prefac'])) for d in range(len(f))]
    print(prec, 'max rel-L2', max(errs), sm.plan_info())
    assert max(errs) <= tol


def test_vhs_per_dof_fs_matches_batch(backend):
    from oracle import fastspectral as fsp
    cfg = vhs_cfg(16, 8, 6, 5.0)
    vm, prm, tab = oracle_vhs(cfg)
    f = fsp.two_maxwellian_batch(vm['cv'], 4, seed=3)
    Qb, _ = _run_vhs(backend, cfg, f, 2, 2)
    Qd, _ = _run_vhs(backend, cfg, f, 2, 2, per_dof=True)
    assert np.array_equal(Qb, Qd)


@pytest.mark.parametrize('prec,tol', [('double', FP64_TOL), ('mixed', MIXED_TOL)])
@pytest.mark.parametrize('N,Nrho,dev', [(8, 4, 4.0), (24, 6, 5.0), (32, 4, 6.0)])
def test_vhs_grid_sizes(backend, N, Nrho, dev, prec, tol):
    from oracle import fastspectral as fsp
    cfg = vhs_cfg(N, Nrho, 6, dev, precision=prec)
    vm, prm, tab = oracle_vhs(cfg)
    f = fsp.two_maxwellian_batch(vm['cv'], 2, seed=N)
    print(f'[vhs N={N}]', flush=True)
    Q, _ = _run_vhs(backend, cfg, f, 1, 2)
    errs = [rel_l2(Q[d], fsp.vhs_fs(f[d], vm, tab, prm['prefac'])) for d in range(len(f))]
    print(N, prec, max(errs))
    assert max(errs) <= tol


def test_vhs_nodal_bilinear(backend):
    import torch
    from oracle import fastspectral as fsp
    from frfs_amd.scattering import DGFSVHSGLLNodalScatteringModel
    from frfs_amd.velocitymesh import DGFSVelocityMesh
    cfg = vhs_cfg(16, 6, 6, 5.0, model='vhs-gll-nodal')
    vm, prm, tab = oracle_vhs(cfg)
    f = fsp.two_maxwellian_batch(vm['cv'], 2, seed=5)
    g = fsp.two_maxwellian_batch(vm['cv'], 2, seed=6)
    hvm = DGFSVelocityMesh(backend, cfg)
    sm = DGFSVHSGLLNodalScatteringModel(backend, cfg, hvm)
    n = hvm.vsize()
    fi = backend.matrix((1, n, 2), _io(f, 1, 2))
    gi = backend.matrix((1, n, 2), _io(g, 1, 2))
    qo = backend.matrix((1, n, 2))
    sm.fs_batch(fi, gi, qo)
    torch.cuda.synchronize()
    Q = _from_io(qo.get())
    errs = [rel_l2(Q[d], fsp.vhs_nodal_fs(f[d], g[d], vm, tab, prm['prefac'])) for d in range(2)]
    assert max(errs) <= FP64_TOL


def _run_vss(backend, cfg, fs, nupts, neles):
    import torch
    from frfs_amd.scattering import DGFSBiVSSGLLScatteringModel
    from frfs_amd.velocitymesh import DGFSBiVelocityMesh
    ns = len(fs)
    vm = DGFSBiVelocityMesh(backend, cfg, ns)
    sm = DGFSBiVSSGLLScatteringModel(backend, cfg, vm)
    n = vm.vsize()
    fin = [backend.matrix((nupts, n, neles), _io(f, nupts, neles)) for f in fs]
    qout = [backend.matrix((nupts, n, neles)) for _ in fs]
    sm.fs_batch(*fin, *qout)
    torch.cuda.synchronize()
    return [_from_io(q.get()) for q in qout], sm


@pytest.mark.parametrize('fused', [1, 0])
@pytest.mark.parametrize('prec,tol', [('double', FP64_TOL), ('mixed', MIXED_TOL)])
def test_vss_single_species(backend, prec, tol, fused):
    """VSS omega=0.74 alpha=1.4, full sphere M=12, Mpre=32."""
    from oracle import fastspectral as fsp
    cfg = vss_cfg(16, 6, 12, 5.0, precision=prec, fused=fused)
    vm, prm, tab = oracle_vss(cfg, 1)
    f = fsp.two_maxwellian_batch(vm['cv'], 3, seed=21)
    print(f'[vss-1 {prec}]', flush=True)
    (Q,), sm = _run_vss(backend, cfg, [f], 1, 3)
    errs = [rel_l2(Q[d], fsp.vss_fs([f[d]], vm, tab, prm)[0]) for d in range(3)]
    print(prec, max(errs), sm.plan_info())
    assert max(errs) <= tol


@pytest.mark.parametrize('prec,tol,lanes', [('double', FP64_TOL, 1), ('mixed', MIXED_TOL, 1),
                                            ('mixed', MIXED_TOL, 2)])
def test_vss_single_species_32(backend, prec, tol, lanes):
    """32^3 grid, one product per group: the FP32 product kernel with one and
    with two lanes per pencil, and the FP64 kernel, against the oracle."""
    from oracle import fastspectral as fsp
    cfg = vss_cfg(32, 4, 12, 6.0, precision=prec, lanes=lanes)
    vm, prm, tab = oracle_vss(cfg, 1)
    f = fsp.two_maxwellian_batch(vm['cv'], 2, seed=32)
    print(f'[vss-1 32^3 {prec}]', flush=True)
    (Q,), sm = _run_vss(backend, cfg, [f], 1, 2)
    errs = [rel_l2(Q[d], fsp.vss_fs([f[d]], vm, tab, prm)[0]) for d in range(2)]
    print(prec, max(errs), sm.plan_info())
    assert max(errs) <= tol


def test_vss_two_species(backend):
    """Two species, mass ratio 1:2, four pairs, uneven omega/diameters."""
    from oracle import fastspectral as fsp
    m0 = 39.96e-3
    cfg = vss_cfg(16, 4, 12, 5.0, masses=(m0, 2 * m0),
                  omega={'00': 0.74, '01': 0.775, '10': 0.775, '11': 0.81},
                  dref={'00': 4.17e-10, '01': 4.17e-10 * 1.15, '10': 4.17e-10 * 1.15,
                        '11': 4.17e-10 * 1.3})
    vm, prm, tab = oracle_vss(cfg, 2)
    f0 = fsp.two_maxwellian_batch(vm['cv'], 2, seed=31, mass=vm['masses'][0])
    f1 = fsp.two_maxwellian_batch(vm['cv'], 2, seed=32, mass=vm['masses'][1], dens_range=(0.2, 0.8))
    print('[vss-2]', flush=True)
    (Q0, Q1), sm = _run_vss(backend, cfg, [f0, f1], 1, 2)
    for d in range(2):
        r0, r1 = fsp.vss_fs([f0[d], f1[d]], vm, tab, prm)
        assert rel_l2(Q0[d], r0) <= FP64_TOL
        assert rel_l2(Q1[d], r1) <= FP64_TOL
    print(sm.plan_info())


def _two_species_case(prec='double', project=0):
    m0 = 39.96e-3
    return vss_cfg(16, 4, 12, 5.0, masses=(m0, 2 * m0), precision=prec, project=project,
                   omega={'00': 0.74, '01': 0.775, '10': 0.775, '11': 0.81},
                   dref={'00': 4.17e-10, '01': 4.17e-10 * 1.15, '10': 4.17e-10 * 1.15,
                         '11': 4.17e-10 * 1.3})


def test_vss_two_species_mixed(backend):
    from oracle import fastspectral as fsp
    cfg = _two_species_case('mixed')
    vm, prm, tab = oracle_vss(cfg, 2)
    f0 = fsp.two_maxwellian_batch(vm['cv'], 2, seed=41, mass=vm['masses'][0])
    f1 = fsp.two_maxwellian_batch(vm['cv'], 2, seed=42, mass=vm['masses'][1], dens_range=(0.2, 0.8))
    print('[vss-2 mixed]', flush=True)
    (Q0, Q1), _ = _run_vss(backend, cfg, [f0, f1], 1, 2)
    for d in range(2):
        r0, r1 = fsp.vss_fs([f0[d], f1[d]], vm, tab, prm)
        assert rel_l2(Q0[d], r0) <= MIXED_TOL
        assert rel_l2(Q1[d], r1) <= MIXED_TOL


@pytest.mark.parametrize('nspecies', [1, 2])
def test_projection_zeroes_invariants(backend, nspecies):
    """The projection leaves every species' mass and the total mass-weighted
    momentum and energy at zero, and its correction lies in the span of the
    constraint functions {e_s, m_s v, m_s |v|^2}: the two properties that fix
    the least-squares correction uniquely.  Momentum and energy of a single
    species of a mixture are not constrained (they move between species)."""
    from oracle import fastspectral as fsp
    if nspecies == 1:
        cfg0, cfg1 = vss_cfg(16, 4, 12, 5.0), vss_cfg(16, 4, 12, 5.0, project=1)
    else:
        cfg0, cfg1 = _two_species_case(), _two_species_case(project=1)
    vm, prm, tab = oracle_vss(cfg0, nspecies)
    masses = vm['masses'][:nspecies]
    fs = [fsp.two_maxwellian_batch(vm['cv'], 3, seed=50 + s, mass=masses[s]) for s in range(nspecies)]
    print(f'[projection {nspecies}]', flush=True)
    Qa, _ = _run_vss(backend, cfg0, fs, 1, 3)
    Qb, _ = _run_vss(backend, cfg1, fs, 1, 3)
    cv = vm['cv']
    phi = np.stack([np.ones(cv.shape[1]), cv[0], cv[1], cv[2], (cv ** 2).sum(0)])     # (5, n)
    n = cv.shape[1]
    for d in range(3):
        # constraint rows over the stacked species vector
        rows = []
        for s in range(nspecies):
            r = np.zeros(nspecies * n)
            r[s * n:(s + 1) * n] = 1
            rows.append(r)
        for k in range(1, 5):
            rows.append(np.concatenate([masses[s] * phi[k] for s in range(nspecies)]))
        A = np.array(rows)
        qa = np.concatenate([Qa[s][d] for s in range(nspecies)])
        qb = np.concatenate([Qb[s][d] for s in range(nspecies)])
        scale = np.abs(A) @ np.abs(qa)
        assert np.all(np.abs(A @ qb) <= 1e-11 * scale), (A @ qb, scale)
        delta = qa - qb
        coef, *_ = np.linalg.lstsq(A.T, delta, rcond=None)
        assert np.linalg.norm(A.T @ coef - delta) <= 1e-9 * np.linalg.norm(delta) + 1e-300
        print(nspecies, d, 'defect removed', np.linalg.norm(delta) / np.linalg.norm(qa))
