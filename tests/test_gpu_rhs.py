"""Full explicit right-hand side on a periodic quad mesh: FR upwind advection
plus the collision operator inside one SSP-RK2 step (the dgfs-tvd-rk2 step
of frfs/integrators/std/steppers.py:122-225 with the collide of
frfs/solvers/dgfs/system.py:199-218), against the CPU oracle, and global
conservation with the FP64 projection."""
import numpy as np
import pytest

from oracle import fastspectral as fsp
from oracle import fr as ofr
from tests.cases import oracle_vhs, vhs_cfg

pytestmark = pytest.mark.gpu


def _system(backend, cfg, nelx, nely, hx, hy, graph=False):
    from frfs_amd.fr import PeriodicQuadAdvection
    from frfs_amd.system import DGFSCollisionSystem, STEPPERS
    sysm = DGFSCollisionSystem(backend, cfg, 9, nelx * nely)
    adv = PeriodicQuadAdvection(backend, sysm.vm, nelx, nely, hx, hy, order=2)
    sysm.rhs = lambda t, uin, fout: adv(sysm.banks[uin], sysm.banks[fout])
    return sysm, adv, STEPPERS['dgfs-tvd-rk2'](sysm, graph=graph)


def _shear_ic(vm_cv, ploc, amp, seed):
    """Maxwellian (rho = 1, T = 1) with a sinusoidal shear u_x = amp sin(2 pi y/Ly + phase)."""
    rng = np.random.default_rng(seed)
    phase = rng.uniform(0, 2 * np.pi)
    y = ploc[:, 1, :]
    Ly = y.max() + y.min()
    ux = amp * np.sin(2 * np.pi * y / Ly + phase)                       # (nupts, neles)
    c = vm_cv
    r2 = (c[0][None, :, None] - ux[:, None, :]) ** 2 + c[1][None, :, None] ** 2 + c[2][None, :, None] ** 2
    return np.pi ** -1.5 * np.exp(-r2)                                  # (nupts, nvar, neles)


def test_ssp_rk2_step_matches_oracle(backend):
    import torch
    cfg = vhs_cfg(8, 3, 6, 4.0)
    nelx, nely, hx, hy = 4, 3, 0.5, 0.5
    sysm, adv, stp = _system(backend, cfg, nelx, nely, hx, hy)
    f0 = _shear_ic(sysm.vm.cv(), adv.ploc(), 0.3, seed=7)
    sysm.banks[0].set(f0)
    dt = 0.01
    stp.step(0.0, dt)
    torch.cuda.synchronize()
    got = sysm.banks[0].get()
    # oracle step
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
    print('SSP-RK2 step vs oracle, rel to increment:', err)
    assert err < 1e-10


def test_rhs_step_conserves_with_projection(backend):
    """Mixed precision + FP64 conservation projection: collisions conserve
    mass/momentum/energy per DOF and FR advection conserves every velocity
    integral, so the global moments are constant to rounding."""
    import torch
    from frfs_amd.fr import gll_rule    # noqa: F401  (weights below are Gauss-Legendre)
    cfg = vhs_cfg(16, 6, 6, 5.0, precision='mixed', project=1)
    nelx, nely, hx, hy = 6, 6, 1.0 / 6, 1.0 / 6
    sysm, adv, stp = _system(backend, cfg, nelx, nely, hx, hy, graph=True)
    sysm.banks[0].set(_shear_ic(sysm.vm.cv(), adv.ploc(), 0.05, seed=3))
    x, w = np.polynomial.legendre.leggauss(3)
    wq = np.array([w[i] * w[j] for j in range(3) for i in range(3)]) * hx * hy / 4
    cv, cw = sysm.vm.cv(), sysm.vm.cw()
    phis = np.stack([np.ones(cv.shape[1]), cv[0], cv[1], cv[2], (cv ** 2).sum(0)])

    def totals():
        f = sysm.banks[0].get()
        return np.einsum('u,uve,kv->k', wq, f, phis) * cw

    m0 = totals()
    for k in range(5):
        stp.step(k * 0.01, 0.01)
    torch.cuda.synchronize()
    m1 = totals()
    scale = np.abs(m0).max()
    print('global moments before/after', m0, m1)
    assert np.max(np.abs(m1 - m0)) < 1e-11 * scale
