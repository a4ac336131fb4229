"""Flux-reconstruction advection operator: host operator construction against
the oracle's independent construction, and the oracle itself pinned by
constant preservation, discrete conservation, upwinding and convergence;
the HIP kernel against the oracle on the GPU."""
import numpy as np
import pytest

from oracle import fr as ofr


@pytest.mark.parametrize('order', [1, 2, 3])
@pytest.mark.parametrize('pts', ['gauss-legendre', 'gauss-legendre-lobatto'])
def test_operators_match_oracle(order, pts):
    from frfs_amd.fr import QuadFROperators
    a = QuadFROperators(order, pts, pts)
    b = ofr.quad_operators(order, pts, pts)
    for k in ('M0', 'M1x', 'M1y', 'M3'):
        assert np.allclose(getattr(a, k), b[k], atol=1e-12), k
    assert np.allclose(a.upts, b['upts']) and np.allclose(a.fpts, b['fpts'])


def _weights(order, pts):
    if pts == 'gauss-legendre':
        x, w = np.polynomial.legendre.leggauss(order + 1)
    else:
        from frfs_amd.fr import gll_rule
        x, w = gll_rule(order + 1)
    return np.array([w[i] * w[j] for j in range(order + 1) for i in range(order + 1)])


def _rand_problem(order=2, nelx=5, nely=4, nvar=6, seed=0):
    rng = np.random.default_rng(seed)
    ops = ofr.quad_operators(order)
    u = rng.normal(size=((order + 1) ** 2, nvar, nelx * nely))
    cx, cy = rng.normal(size=nvar), rng.normal(size=nvar)
    return ops, u, cx, cy


def test_constant_preserved():
    ops, u, cx, cy = _rand_problem()
    r = ofr.fr_rhs(np.ones_like(u), cx, cy, 5, 4, 0.3, 0.7, ops)
    assert np.max(np.abs(r)) < 1e-12


@pytest.mark.parametrize('pts', ['gauss-legendre', 'gauss-legendre-lobatto'])
def test_discrete_conservation(pts):
    ops = ofr.quad_operators(2, pts, pts)
    rng = np.random.default_rng(1)
    u = rng.normal(size=(9, 5, 20))
    cx, cy = rng.normal(size=5), rng.normal(size=5)
    r = ofr.fr_rhs(u, cx, cy, 5, 4, 0.3, 0.7, ops)
    w = _weights(2, pts) * (0.3 * 0.7 / 4)
    tot = np.einsum('i,ive->v', w, r)
    scale = np.einsum('i,ive->v', w, np.abs(r))
    assert np.all(np.abs(tot) < 1e-13 * scale)


def test_upwind_dependence():
    """cx > 0, cy = 0: an element's right-hand side ignores its right neighbour."""
    ops, u, _, _ = _rand_problem(nvar=1)
    cx, cy = np.array([1.3]), np.array([0.0])
    r0 = ofr.fr_rhs(u, cx, cy, 5, 4, 0.3, 0.7, ops)
    u2 = u.copy()
    u2[:, :, 2] += 1.0                     # element (2, 0): right neighbour of (1, 0)
    r1 = ofr.fr_rhs(u2, cx, cy, 5, 4, 0.3, 0.7, ops)
    assert np.allclose(r0[:, :, 1], r1[:, :, 1], atol=1e-14)
    assert not np.allclose(r0[:, :, 3], r1[:, :, 3])   # left neighbour of the change sees it


def test_convergence_order():
    errs = []
    for nel in (8, 16):
        ops = ofr.quad_operators(2)
        h = 1.0 / nel
        up = ops['upts']
        e = np.arange(nel * nel)
        x = (e % nel + 0.5 * (up[:, 0:1] + 1)) * h
        y = (e // nel + 0.5 * (up[:, 1:2] + 1)) * h
        u = np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y)
        cx, cy = np.array([0.7]), np.array([-0.4])
        r = ofr.fr_rhs(u[:, None, :], cx, cy, nel, nel, h, h, ops)[:, 0, :]
        exact = -(0.7 * 2 * np.pi * np.cos(2 * np.pi * x) * np.cos(2 * np.pi * y)
                  - 0.4 * -2 * np.pi * np.sin(2 * np.pi * x) * np.sin(2 * np.pi * y))
        errs.append(np.sqrt(np.mean((r - exact) ** 2)))
    assert errs[0] / errs[1] > 3.5, errs


@pytest.mark.gpu
@pytest.mark.parametrize('order,pts', [(1, 'gauss-legendre'), (2, 'gauss-legendre'), (3, 'gauss-legendre'),
                                       (2, 'gauss-legendre-lobatto'), (3, 'gauss-legendre-lobatto')])
def test_gpu_kernel_matches_oracle(backend, order, pts):
    import torch
    from frfs_amd.fr import PeriodicQuadAdvection
    from frfs_amd.velocitymesh import DGFSVelocityMesh
    from tests.cases import vhs_cfg
    vm = DGFSVelocityMesh(backend, vhs_cfg(8, 2, 6, 4.0))
    nelx, nely, hx, hy = 6, 5, 0.25, 0.4
    op = PeriodicQuadAdvection(backend, vm, nelx, nely, hx, hy, order=order, upts=pts, fpts=pts)
    rng = np.random.default_rng(2)
    u = rng.normal(size=((order + 1) ** 2, vm.vsize(), nelx * nely))
    ub = backend.matrix(u.shape, u)
    ob = backend.matrix(u.shape)
    print(f'[fr order={order} {pts}]', flush=True)
    op(ub, ob)
    torch.cuda.synchronize()
    want = ofr.fr_rhs(u, vm.cv()[0], vm.cv()[1], nelx, nely, hx, hy, ofr.quad_operators(order, pts, pts))
    got = ob.get()
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
