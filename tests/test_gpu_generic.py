"""Grids without native transform kernels (even N other than 8, 16, 24, 32):
the library-transform path (hipFFT plans plus the pointwise kernels) against
the oracle in both precisions, for VHS, single-species and two-species VSS,
with ragged batches, the conservation projection and HIP-graph replay.

Tolerances as for the native grids: FP64 rel-L2 <= 1e-12 per DOF, mixed
<= 1e-5."""
import numpy as np
import pytest

from tests.cases import oracle_vhs, oracle_vss, rel_l2, vhs_cfg, vss_cfg
from tests.test_gpu_parity import FP64_TOL, MIXED_TOL, _io, _run_vhs, _run_vss

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('N', [4, 6, 12, 20])
@pytest.mark.parametrize('prec,tol', [('double', FP64_TOL), ('mixed', MIXED_TOL)])
def test_vhs_generic_grid(backend, N, prec, tol):
    """VHS omega=0.5, M=6 half sphere, Nrho=4 on an N^3 grid without native kernels."""
    from oracle import fastspectral as fsp
    cfg = vhs_cfg(N, 4, 6, 5.0, precision=prec)
    vm, prm, tab = oracle_vhs(cfg)
    f = fsp.two_maxwellian_batch(vm['cv'], 6, seed=20 + N)
    print(f'[generic vhs N={N} {prec}]', flush=True)
    Q, sm = _run_vhs(backend, cfg, f, 2, 3)
    assert sm._h.paths() == ['library_fft']
    errs = [rel_l2(Q[d], fsp.vhs_fs(f[d], vm, tab, prm['prefac'])) for d in range(len(f))]
    print(prec, 'max rel-L2', max(errs))
    assert max(errs) <= tol


def test_vhs_nodal_generic_grid(backend):
    """The bilinear vhs-gll-nodal form (two sources) on 12^3, FP64."""
    import torch
    from oracle import fastspectral as fsp
    from frfs_amd.scattering import DGFSVHSGLLNodalScatteringModel
    from frfs_amd.velocitymesh import DGFSVelocityMesh
    from tests.test_gpu_parity import _from_io
    cfg = vhs_cfg(12, 3, 6, 5.0, model='vhs-gll-nodal')
    vm, prm, tab = oracle_vhs(cfg)
    f = fsp.two_maxwellian_batch(vm['cv'], 2, seed=91)
    g = fsp.two_maxwellian_batch(vm['cv'], 2, seed=92)
    hvm = DGFSVelocityMesh(backend, cfg)
    sm = DGFSVHSGLLNodalScatteringModel(backend, cfg, hvm)
    assert sm._h.paths() == ['library_fft']
    n = hvm.vsize()
    fi = backend.matrix((1, n, 2), _io(f, 1, 2))
    gi = backend.matrix((1, n, 2), _io(g, 1, 2))
    qo = backend.matrix((1, n, 2))
    print('[generic nodal N=12]', flush=True)
    sm.fs_batch(fi, gi, qo)
    torch.cuda.synchronize()
    Q = _from_io(qo.get())
    errs = [rel_l2(Q[d], fsp.vhs_nodal_fs(f[d], g[d], vm, tab, prm['prefac'])) for d in range(2)]
    assert max(errs) <= FP64_TOL


def test_vhs_48_against_oracle(backend):
    """48^3 (a common grid above the native ones), one DOF of each precision
    against the oracle."""
    from oracle import fastspectral as fsp
    base = vhs_cfg(48, 2, 6, 5.0)
    vm, prm, tab = oracle_vhs(base)
    f = fsp.two_maxwellian_batch(vm['cv'], 1, seed=48)
    ref = fsp.vhs_fs(f[0], vm, tab, prm['prefac'])
    for prec, tol in (('double', FP64_TOL), ('mixed', MIXED_TOL)):
        print(f'[generic vhs N=48 {prec}]', flush=True)
        Q, sm = _run_vhs(backend, vhs_cfg(48, 2, 6, 5.0, precision=prec), f, 1, 1)
        assert sm._h.paths() == ['library_fft']
        e = rel_l2(Q[0], ref)
        print(prec, 'rel-L2', e)
        assert e <= tol


def test_vhs_largest_generic_grid(backend):
    """The largest grid the library path takes (128^3, 2M points; 32 DOFs per
    batch in the 8 GiB budget): Q(2f) = 4 Q(f) in FP64 (Q is quadratic) and
    the mixed mode within 1e-5 of FP64, on 2 DOFs (the oracle takes minutes
    here; 48^3 above is checked against it)."""
    from oracle import fastspectral as fsp
    cfg = vhs_cfg(128, 2, 6, 5.0)
    vm, _, _ = oracle_vhs(cfg)
    f = fsp.two_maxwellian_batch(vm['cv'], 2, seed=128)
    print('[generic vhs N=128]', flush=True)
    Q, sm = _run_vhs(backend, cfg, f, 1, 2)
    assert sm._h.paths() == ['library_fft'] and sm._h.chunk_dofs() == 32
    Q2, _ = _run_vhs(backend, cfg, 2 * f, 1, 2)
    Qm, _ = _run_vhs(backend, vhs_cfg(128, 2, 6, 5.0, precision='mixed'), f, 1, 2)
    for d in range(2):
        assert np.all(np.isfinite(Q[d])) and np.abs(Q[d]).max() > 0
        assert rel_l2(Q2[d], 4 * Q[d]) <= 1e-13
        assert rel_l2(Qm[d], Q[d]) <= MIXED_TOL


@pytest.mark.parametrize('prec,tol', [('double', FP64_TOL), ('mixed', MIXED_TOL)])
def test_vss_generic_grid(backend, prec, tol):
    """VSS omega=0.74 alpha=1.4, M=12 full sphere, Nrho=3, Mpre=32 on 20^3."""
    from oracle import fastspectral as fsp
    cfg = vss_cfg(20, 3, 12, 5.0, precision=prec)
    vm, prm, tab = oracle_vss(cfg, 1)
    f = fsp.two_maxwellian_batch(vm['cv'], 3, seed=33)
    print(f'[generic vss N=20 {prec}]', flush=True)
    (Q,), sm = _run_vss(backend, cfg, [f], 1, 3)
    errs = [rel_l2(Q[d], fsp.vss_fs([f[d]], vm, tab, prm)[0]) for d in range(3)]
    print(prec, 'max rel-L2', max(errs))
    assert max(errs) <= tol


def test_vss_generic_two_species(backend):
    """Two species on 12^3, mass ratio 1:2, four pairs, FP64."""
    from oracle import fastspectral as fsp
    m0 = 39.96e-3
    cfg = vss_cfg(12, 3, 12, 5.0, masses=(m0, 2 * m0),
                  omega={'00': 0.74, '01': 0.775, '10': 0.775, '11': 0.81},
                  dref={'00': 4.17e-10, '01': 4.17e-10 * 1.15, '10': 4.17e-10 * 1.15,
                        '11': 4.17e-10 * 1.3})
    vm, prm, tab = oracle_vss(cfg, 2)
    f0 = fsp.two_maxwellian_batch(vm['cv'], 2, seed=51, mass=vm['masses'][0])
    f1 = fsp.two_maxwellian_batch(vm['cv'], 2, seed=52, mass=vm['masses'][1], dens_range=(0.2, 0.8))
    print('[generic vss-2 N=12]', flush=True)
    (Q0, Q1), sm = _run_vss(backend, cfg, [f0, f1], 1, 2)
    assert sm._h.paths() == ['library_fft']
    for d in range(2):
        r0, r1 = fsp.vss_fs([f0[d], f1[d]], vm, tab, prm)
        assert rel_l2(Q0[d], r0) <= FP64_TOL
        assert rel_l2(Q1[d], r1) <= FP64_TOL


def test_generic_ragged_batches(backend):
    """Batches of at most 3 DOFs (7 DOFs as 3 + 3 + 1 in a 3-slot scratch,
    the library plans transforming all 3 slots) agree with one batch of 7 to
    rounding (the library may pick other kernels for another batch count),
    and with the oracle."""
    from oracle import fastspectral as fsp
    base = vhs_cfg(12, 3, 6, 5.0)
    vm, prm, tab = oracle_vhs(base)
    f = fsp.two_maxwellian_batch(vm['cv'], 7, seed=61)
    print('[generic ragged N=12]', flush=True)
    Q1, sm1 = _run_vhs(backend, base, f, 1, 7)
    Q3, sm3 = _run_vhs(backend, vhs_cfg(12, 3, 6, 5.0, max_chunk=3), f, 1, 7)
    assert sm3._h.chunk_dofs() == 3
    for d in range(len(f)):
        ref = fsp.vhs_fs(f[d], vm, tab, prm['prefac'])
        assert rel_l2(Q3[d], Q1[d]) <= 1e-13
        assert rel_l2(Q3[d], ref) <= FP64_TOL
        assert rel_l2(Q1[d], ref) <= FP64_TOL


def test_generic_projection_conserves(backend):
    """Mixed + projection against FP64 + projection (rounding only) on 20^3,
    and the projected Q conserves mass, momentum and energy."""
    from oracle import fastspectral as fsp
    vm, prm, tab = oracle_vss(vss_cfg(20, 3, 12, 5.0), 1)
    f = fsp.two_maxwellian_batch(vm['cv'], 3, seed=71)
    print('[generic projection N=20]', flush=True)
    (Qm,), _ = _run_vss(backend, vss_cfg(20, 3, 12, 5.0, precision='mixed', project=1), [f], 1, 3)
    (Qd,), _ = _run_vss(backend, vss_cfg(20, 3, 12, 5.0, precision='double', project=1), [f], 1, 3)
    cv = vm['cv']
    phi = np.stack([np.ones(cv.shape[1]), cv[0], cv[1], cv[2], (cv ** 2).sum(0)])
    for d in range(3):
        ref = fsp.vss_fs([f[d]], vm, tab, prm)[0]
        e = rel_l2(Qm[d], Qd[d])
        scale = np.abs(phi) @ np.abs(Qd[d])
        print(d, 'mixed+proj vs fp64+proj', e, 'projection moved Q by', rel_l2(Qd[d], ref))
        assert e <= MIXED_TOL
        assert np.all(np.abs(phi @ Qd[d]) <= 1e-11 * scale)
        assert np.all(np.abs(phi @ Qm[d]) <= 1e-11 * scale)


def test_generic_graph_replay_matches_eager(backend):
    """A captured SSP-RK2 step on a 12^3 grid, replayed three times with
    eager work and synchronisations between the replays, equals eager steps
    bit for bit."""
    import torch
    from frfs_amd.system import DGFSCollisionSystem, STEPPERS
    from oracle import fastspectral as fsp

    def system(graph):
        cfg = vhs_cfg(12, 3, 6, 5.0, omega=1.0)
        sysm = DGFSCollisionSystem(backend, cfg, 1, 2)
        return sysm, STEPPERS['dgfs-tvd-rk2'](sysm, graph=graph)
    sysm_g, stp_g = system(True)
    sysm_e, stp_e = system(False)
    f0 = fsp.bkw(sysm_g.vm.cv(), 5.5)
    init = _io(np.stack([f0, 0.5 * f0]), 1, 2)
    sysm_g.banks[0].set(init)
    sysm_e.banks[0].set(init)
    dt = 0.05 / (4 * np.pi * sysm_g.sm._prefactor)
    print('[generic graph N=12]', flush=True)
    for k in range(3):
        stp_g.step(k * dt, dt)
        torch.cuda.synchronize()
        stp_e.step(k * dt, dt)
        torch.cuda.synchronize()
        junk = torch.randn(1 << 20, device='cuda', dtype=torch.float64).sum()
        torch.cuda.synchronize()
        assert np.isfinite(junk.item())
        assert np.array_equal(sysm_g.banks[0].get(), sysm_e.banks[0].get()), k


@pytest.mark.parametrize('N,Nrho', [(16, 4), (32, 2)])
@pytest.mark.parametrize('prec,tol', [('double', FP64_TOL), ('mixed', MIXED_TOL)])
def test_library_flag_on_native_grid(backend, N, Nrho, prec, tol):
    """FQ_FLAG_LIBRARY_FFT (`library-fft = 1`) runs the library path on a
    grid with native kernels: it matches the oracle and the native path."""
    from oracle import fastspectral as fsp
    cfg = vss_cfg(N, Nrho, 12, 5.0, precision=prec)
    cfg.set('scattering-model', 'library-fft', 1)
    vm, prm, tab = oracle_vss(cfg, 1)
    f = fsp.two_maxwellian_batch(vm['cv'], 3, seed=80 + N)
    print(f'[library flag N={N} {prec}]', flush=True)
    (Ql,), sml = _run_vss(backend, cfg, [f], 1, 3)
    (Qn,), smn = _run_vss(backend, vss_cfg(N, Nrho, 12, 5.0, precision=prec), [f], 1, 3)
    assert sml._h.paths() == ['library_fft'] and 'library_fft' not in smn._h.paths()
    for d in range(3):
        ref = fsp.vss_fs([f[d]], vm, tab, prm)[0]
        assert rel_l2(Ql[d], ref) <= tol
        assert rel_l2(Ql[d], Qn[d]) <= (1e-12 if prec == 'double' else 2e-5)
