"""The benchmark workloads at their own shapes (grid, directions, radial
points, kernel), on a few DOFs so the oracle finishes in seconds:

* the single-species VSS operator on the 32^3 grid with M = 12, Nrho = 16,
  omega 0.74 / alpha 1.4, Mpre = 32 (frfs/solvers/dgfsbi/scattering.py:259-330);
* two species at 32^3 (mass ratio 1:2, omega 0.74 / 0.81, diameter ratio
  1.3): the four pairs of frfs/solvers/dgfsbi/scattering.py:323-330;
* one SSP-RK2 step of FR advection plus collisions at 24^3 with Nrho = 12 on a
  periodic 2 x 2 quad mesh of order 2 (frfs/integrators/std/steppers.py:122-225
  around frfs/solvers/dgfs/system.py:199-218).

FP64 results meet rel-L2 <= 1e-12 per DOF against the oracle, mixed
precision <= 1e-5.  With the conservation projection, the mixed operator is
compared with the FP64 operator under the same projection (both conserve, so
they differ by rounding only)."""
import numpy as np
import pytest

from oracle import fastspectral as fsp
from tests.cases import oracle_vhs, oracle_vss, rel_l2, vhs_cfg, vss_cfg
from tests.test_gpu_parity import FP64_TOL, MIXED_TOL, _run_vss

pytestmark = pytest.mark.gpu

M0 = 39.96e-3


def _vss32(prec, project=0):
    return vss_cfg(32, 16, 12, 6.0, precision=prec, project=project, Mpre=32)


@pytest.fixture(scope='module')
def vss32_case():
    cfg = _vss32('double')
    vm, prm, tab = oracle_vss(cfg, 1)
    f = fsp.two_maxwellian_batch(vm['cv'], 2, seed=2024)
    ref = [fsp.vss_fs([f[d]], vm, tab, prm)[0] for d in range(2)]
    return vm, f, ref


@pytest.mark.parametrize('prec,tol', [('double', FP64_TOL), ('mixed', MIXED_TOL)])
def test_vss32_full_shape(backend, vss32_case, prec, tol):
    vm, f, ref = vss32_case
    print(f'[vss32 M=12 Nrho=16 {prec}]', flush=True)
    (Q,), sm = _run_vss(backend, _vss32(prec), [f], 1, 2)
    errs = [rel_l2(Q[d], ref[d]) for d in range(2)]
    print(prec, errs, sm.plan_info())
    assert max(errs) <= tol


def test_vss32_full_shape_projected(backend, vss32_case):
    """Mixed + projection against FP64 + projection (rounding only), and the
    projected Q conserves mass, momentum and energy."""
    vm, f, ref = vss32_case
    print('[vss32 projected]', flush=True)
    (Qm,), _ = _run_vss(backend, _vss32('mixed', 1), [f], 1, 2)
    (Qd,), _ = _run_vss(backend, _vss32('double', 1), [f], 1, 2)
    cv = vm['cv']
    phi = np.stack([np.ones(cv.shape[1]), cv[0], cv[1], cv[2], (cv ** 2).sum(0)])
    for d in range(2):
        e = rel_l2(Qm[d], Qd[d])
        scale = np.abs(phi) @ np.abs(Qd[d])
        print(d, 'mixed+proj vs fp64+proj', e, 'projection moved Q by', rel_l2(Qd[d], ref[d]))
        assert e <= MIXED_TOL
        assert np.all(np.abs(phi @ Qd[d]) <= 1e-11 * scale)
        assert np.all(np.abs(phi @ Qm[d]) <= 1e-11 * scale)


def _two32(prec):
    return vss_cfg(32, 3, 12, 6.0, masses=(M0, 2 * M0), precision=prec,
                   omega={'00': 0.74, '01': 0.775, '10': 0.775, '11': 0.81},
                   dref={'00': 4.17e-10, '01': 4.17e-10 * 1.15, '10': 4.17e-10 * 1.15,
                         '11': 4.17e-10 * 1.3})


@pytest.fixture(scope='module')
def two32_case():
    vm, prm, tab = oracle_vss(_two32('double'), 2)
    f0 = fsp.two_maxwellian_batch(vm['cv'], 2, seed=71, mass=vm['masses'][0])
    f1 = fsp.two_maxwellian_batch(vm['cv'], 2, seed=72, mass=vm['masses'][1], dens_range=(0.2, 0.8))
    ref = [fsp.vss_fs([f0[d], f1[d]], vm, tab, prm) for d in range(2)]
    return f0, f1, ref


@pytest.mark.parametrize('prec,tol', [('double', FP64_TOL), ('mixed', MIXED_TOL)])
def test_two_species_32(backend, two32_case, prec, tol):
    """Cross-species products (items from different sources, mass-ratio
    phases) and same-species conjugate products in one plan at 32^3."""
    f0, f1, ref = two32_case
    print(f'[vss-2 32^3 {prec}]', flush=True)
    (Q0, Q1), sm = _run_vss(backend, _two32(prec), [f0, f1], 1, 2)
    for d in range(2):
        e0, e1 = rel_l2(Q0[d], ref[d][0]), rel_l2(Q1[d], ref[d][1])
        print(prec, d, e0, e1)
        assert e0 <= tol and e1 <= tol


# ---------------------------------------------------------------- SSP-RK2 step, 24^3
def _rk2_system(backend, cfg, graph):
    from frfs_amd.fr import PeriodicQuadAdvection
    from frfs_amd.system import DGFSCollisionSystem, STEPPERS
    nelx = nely = 2
    hx = hy = 0.5
    sysm = DGFSCollisionSystem(backend, cfg, 9, nelx * nely)
    adv = PeriodicQuadAdvection(backend, sysm.vm, nelx, nely, hx, hy, order=2)
    sysm.rhs = lambda t, uin, fout: adv(sysm.banks[uin], sysm.banks[fout])
    return sysm, adv, STEPPERS['dgfs-tvd-rk2'](sysm, graph=graph), (nelx, nely, hx, hy)


def _shear(cv, ploc, amp, seed):
    rng = np.random.default_rng(seed)
    phase = rng.uniform(0, 2 * np.pi)
    y = ploc[:, 1, :]
    Ly = y.max() + y.min()
    ux = amp * np.sin(2 * np.pi * y / Ly + phase)
    r2 = (cv[0][None, :, None] - ux[:, None, :]) ** 2 + cv[1][None, :, None] ** 2 + cv[2][None, :, None] ** 2
    return np.pi ** -1.5 * np.exp(-r2)


def _rk2_cfg(prec, project=0):
    # M = 6 directions of the 12-point design's half sphere (its 6 antipodal pairs)
    return vhs_cfg(24, 12, 6, 5.0, precision=prec, project=project)


def test_rk2_step_24_matches_oracle(backend):
    import torch
    from oracle import fr as ofr
    cfg = _rk2_cfg('double')
    sysm, adv, stp, (nelx, nely, hx, hy) = _rk2_system(backend, cfg, graph=False)
    f0 = _shear(sysm.vm.cv(), adv.ploc(), 0.05, seed=13)
    sysm.banks[0].set(f0)
    dt = 0.002
    print('[rk2 24^3 fp64]', flush=True)
    stp.step(0.0, dt)
    torch.cuda.synchronize()
    got = sysm.banks[0].get()
    vm, prm, tab = oracle_vhs(cfg)
    ops = ofr.quad_operators(2)
    cv = vm['cv']

    def L(f):
        q = np.empty_like(f)
        for e in range(f.shape[2]):
            for u in range(f.shape[0]):
                q[u, :, e] = fsp.vhs_fs(f[u, :, e], vm, tab, prm['prefac'])
        return q + ofr.fr_rhs(f, cv[0], cv[1], nelx, nely, hx, hy, ops)
    f1 = f0 + dt * L(f0)
    want = 0.5 * f0 + 0.5 * f1 + 0.5 * dt * L(f1)
    err = np.linalg.norm(got - want) / np.linalg.norm(want - f0)
    print('24^3 SSP-RK2 step vs oracle, relative to the increment:', err)
    assert err < 1e-10


def test_rk2_step_24_mixed_projected_graph_conserves(backend):
    """Mixed precision + FP64 projection, every stage replayed from a HIP
    graph: the global mass, momentum and energy stay constant to rounding."""
    import torch
    cfg = _rk2_cfg('mixed', project=1)
    sysm, adv, stp, (nelx, nely, hx, hy) = _rk2_system(backend, cfg, graph=True)
    sysm.banks[0].set(_shear(sysm.vm.cv(), adv.ploc(), 0.05, seed=17))
    x, w = np.polynomial.legendre.leggauss(3)
    wq = np.array([w[i] * w[j] for j in range(3) for i in range(3)]) * hx * hy / 4
    cv, cw = sysm.vm.cv(), sysm.vm.cw()
    phis = np.stack([np.ones(cv.shape[1]), cv[0], cv[1], cv[2], (cv ** 2).sum(0)])

    def totals():
        return np.einsum('u,uve,kv->k', wq, sysm.banks[0].get(), phis) * cw

    print('[rk2 24^3 mixed projected graph]', flush=True)
    m0 = totals()
    for k in range(3):
        stp.step(k * 0.002, 0.002)
    torch.cuda.synchronize()
    m1 = totals()
    print('global moments', m0, m1)
    assert np.max(np.abs(m1 - m0)) < 1e-11 * np.abs(m0).max()
