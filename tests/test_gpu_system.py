"""GPU tests of the kernels around Q and of the collide/stepper driver:
axnpby and velocity-range copies on AoSoA banks, device moments against the
oracle, BKW relaxation in time (SSP-RK2 against the exact solution) and HIP
graph replay of a whole step against eager execution."""
import numpy as np
import pytest

from oracle import fastspectral as fsp
from tests.cases import ARGON, vhs_cfg

pytestmark = pytest.mark.gpu


def _io(f, nupts, neles):
    ndof, n = f.shape
    return f.reshape(neles, nupts, n).transpose(1, 2, 0).copy()


def test_axnpby_and_copy(backend):
    import torch
    from frfs_amd import _lib
    rng = np.random.default_rng(0)
    shape = (2, 40, 37)
    a, b, c = (backend.matrix(shape, rng.normal(size=shape)) for _ in range(3))
    A, B, C = a.get(), b.get(), c.get()
    st = backend.stream_handle()
    n = a.nrow * a.leaddim
    _lib.axnpby(n, a._as_parameter_, [b._as_parameter_, c._as_parameter_], [0.5, 2.0, -1.5], st)
    torch.cuda.synchronize()
    assert np.allclose(a.get(), 0.5 * A + 2.0 * B - 1.5 * C, rtol=1e-15, atol=1e-15)
    _lib.axnpby(n, a._as_parameter_, [b._as_parameter_], [0.0, 3.0], st)
    torch.cuda.synchronize()
    assert np.array_equal(a.get(), 3.0 * B)
    _lib.axnpby(n, a._as_parameter_, [c._as_parameter_], [1.0, 1.0], st)
    torch.cuda.synchronize()
    assert np.allclose(a.get(), 3.0 * B + C, rtol=1e-15, atol=1e-15)
    # velocity-range copy between banks of different variable counts
    reg = backend.matrix((2, 8, 37))
    _lib.copy_dgfs(2, 37, 8, b._as_parameter_, b.leaddim, 40, 16, reg._as_parameter_, reg.leaddim, 8, 0, st)
    torch.cuda.synchronize()
    assert np.array_equal(reg.get(), B[:, 16:24, :])
    _lib.copy_dgfs(2, 37, 8, reg._as_parameter_, reg.leaddim, 8, 0, c._as_parameter_, c.leaddim, 40, 32, st)
    torch.cuda.synchronize()
    want = C.copy()
    want[:, 32:40, :] = B[:, 16:24, :]
    assert np.array_equal(c.get(), want)


def test_device_moments_match_oracle(backend):
    from frfs_amd.system import bulk_moments
    from frfs_amd.velocitymesh import DGFSVelocityMesh
    cfg = vhs_cfg(16, 4, 6, 5.0)
    vm = DGFSVelocityMesh(backend, cfg)
    ovm = fsp.velocity_mesh(16, 4, 6, dev=5.0, **ARGON)
    f = fsp.two_maxwellian_batch(ovm['cv'], 6, seed=3)
    m = backend.matrix((2, vm.vsize(), 3), _io(f, 2, 3))
    got = bulk_moments(vm, m)
    scale = np.array([vm.rho0(), vm.u0(), vm.u0(), vm.u0(), vm.T0()] + [0.5 * vm.rho0() * vm.u0() ** 3] * 3
                     + [0.5 * vm.rho0() * vm.u0() ** 2] * 6)
    for d in range(6):
        want = fsp.bulk_moments_3d(f[d], ovm['cv'], ovm['cw']) * scale
        assert np.allclose(got[d, :14], want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def _bkw_system(backend, N=24, Nrho=12, dev=6.0, graph=False):
    from frfs_amd.system import DGFSCollisionSystem, STEPPERS
    cfg = vhs_cfg(N, Nrho, 6, dev, omega=1.0)
    sysm = DGFSCollisionSystem(backend, cfg, 1, 2)
    return sysm, STEPPERS['dgfs-tvd-rk2'](sysm, graph=graph)


def test_bkw_relaxation_ssp_rk2(backend):
    """df/dt = Q(f) from the BKW state at tau0 = 5.5: Q = 4 pi prefac d_tau f,
    so f(t) = f_BKW(tau0 + 4 pi prefac t).  SSP-RK2 error falls with dt."""
    import torch
    errs = []
    tau0, dtau = 5.5, 1.0
    for nsteps in (4, 8):
        sysm, stp = _bkw_system(backend)
        cv = sysm.vm.cv()
        f0 = fsp.bkw(cv, tau0)
        sysm.banks[0].set(_io(np.stack([f0, f0]), 1, 2))
        prefac = sysm.sm._prefactor
        dt = dtau / (4 * np.pi * prefac) / nsteps
        for k in range(nsteps):
            stp.step(k * dt, dt)
        torch.cuda.synchronize()
        f = sysm.banks[0].get()[0].T[0]
        want = fsp.bkw(cv, tau0 + dtau)
        errs.append(np.linalg.norm(f - want) / np.linalg.norm(want))
    print('BKW SSP-RK2 rel errors', errs)
    assert errs[1] < errs[0] and errs[1] < 2e-3


def test_graph_replay_matches_eager(backend):
    """A captured SSP-RK2 step replayed three times, with eager work and
    synchronisations between the replays, equals eager steps bit for bit."""
    import torch
    sysm_g, stp_g = _bkw_system(backend, N=16, Nrho=6, dev=5.0, graph=True)
    sysm_e, stp_e = _bkw_system(backend, N=16, Nrho=6, dev=5.0, graph=False)
    f0 = fsp.bkw(sysm_g.vm.cv(), 5.5)
    init = _io(np.stack([f0, 0.5 * f0]), 1, 2)
    sysm_g.banks[0].set(init)
    sysm_e.banks[0].set(init)
    dt = 0.05 / (4 * np.pi * sysm_g.sm._prefactor)
    for k in range(3):
        stp_g.step(k * dt, dt)
        torch.cuda.synchronize()
        stp_e.step(k * dt, dt)
        torch.cuda.synchronize()
        junk = torch.randn(1 << 20, device='cuda', dtype=torch.float64).sum()   # eager work between replays
        torch.cuda.synchronize()
        assert np.isfinite(junk.item())
        assert np.array_equal(sysm_g.banks[0].get(), sysm_e.banks[0].get()), k


def _two_species_cfg(prec='double', project=0):
    from tests.cases import vss_cfg
    m0 = 39.96e-3
    return vss_cfg(16, 4, 12, 5.0, masses=(m0, 2 * m0), precision=prec, project=project,
                   omega={'00': 0.74, '01': 0.775, '10': 0.775, '11': 0.81},
                   dref={'00': 4.17e-10, '01': 4.17e-10 * 1.15, '10': 4.17e-10 * 1.15,
                         '11': 4.17e-10 * 1.3})


def test_two_species_euler_step_matches_oracle(backend):
    """dgfsbi-euler on the multi-species system: f_s <- f_s + dt Q_s(f)."""
    import torch
    from frfs_amd.system import DGFSBiCollisionSystem, STEPPERS
    from tests.cases import oracle_vss
    cfg = _two_species_cfg()
    sysm = DGFSBiCollisionSystem(backend, cfg, 2, 1, 3)
    vm, prm, tab = oracle_vss(cfg, 2)
    f = [fsp.two_maxwellian_batch(vm['cv'], 3, seed=60 + s, mass=vm['masses'][s]) for s in range(2)]
    for s in range(2):
        sysm.banks[s][0].set(_io(f[s], 1, 3))
    stp = STEPPERS['dgfsbi-euler'](sysm)
    dt = 0.02
    print('[bi euler]', flush=True)
    stp.step(0.0, dt)
    torch.cuda.synchronize()
    for d in range(3):
        q = fsp.vss_fs([f[0][d], f[1][d]], vm, tab, prm)
        for s in range(2):
            got = sysm.banks[s][0].get()[0, :, d]
            want = f[s][d] + dt * q[s]
            assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(dt * q[s]), (d, s)


def test_two_species_rk2_graph_conserves(backend):
    """dgfsbi-tvd-rk2, mixed precision with the projection, captured in a HIP
    graph: every species' mass and the total momentum and energy stay fixed."""
    import torch
    from frfs_amd.system import DGFSBiCollisionSystem, STEPPERS
    from tests.cases import oracle_vss
    cfg = _two_species_cfg('mixed', project=1)
    sysm = DGFSBiCollisionSystem(backend, cfg, 2, 1, 4)
    vm, prm, tab = oracle_vss(cfg, 2)
    ms = vm['masses'][:2]
    for s in range(2):
        f = fsp.two_maxwellian_batch(vm['cv'], 4, seed=70 + s, mass=ms[s])
        sysm.banks[s][0].set(_io(f, 1, 4))
    cv = vm['cv']
    phi = np.stack([np.ones(cv.shape[1]), cv[0], cv[1], cv[2], (cv ** 2).sum(0)])

    def invariants():
        f = [sysm.banks[s][0].get()[0] for s in range(2)]          # (n, ndof)
        mom = [phi @ f[s] for s in range(2)]                       # (5, ndof)
        return np.concatenate([mom[0][:1], mom[1][:1], ms[0] * mom[0][1:] + ms[1] * mom[1][1:]])
    i0 = invariants()
    stp = STEPPERS['dgfsbi-tvd-rk2'](sysm, graph=True)
    print('[bi rk2 graph]', flush=True)
    for k in range(4):
        stp.step(k * 0.01, 0.01)
    torch.cuda.synchronize()
    i1 = invariants()
    assert np.max(np.abs(i1 - i0)) <= 1e-11 * np.max(np.abs(i0))
