"""Pair packing of the product kernels (24^3 and 32^3 grids): the two item
transforms of a product run as one complex transform plus the Nyquist-plane
corrections of k_nyq (FP32: every product; FP64: conjugate pairs).  Mixed
cases are compared with the same kernel run unpacked (FQ_FLAG_NO_PACK) and
with the CPU oracle: both are FP32 transforms with FP64 accumulation, so they
differ by rounding only, rel-L2 <= 2e-6 between them and <= 1e-5 against the
oracle.  FP64 cases meet the FP64 bar (1e-12).  Batching and DOF sub-ranges
must not change a bit.
"""
import numpy as np
import pytest

from tests.cases import oracle_vhs, oracle_vss, rel_l2, vhs_cfg, vss_cfg
from tests.test_gpu_parity import MIXED_TOL, _from_io, _io, _run_vhs, _run_vss

pytestmark = pytest.mark.gpu

PACK_VS_UNPACKED = 2e-6


def _two_species(N, pack):
    m0 = 39.96e-3
    return vss_cfg(N, 3, 12, 5.0, masses=(m0, 2 * m0), precision='mixed', pack=pack,
                   omega={'00': 0.74, '01': 0.775, '10': 0.775, '11': 0.81},
                   dref={'00': 4.17e-10, '01': 4.17e-10 * 1.15, '10': 4.17e-10 * 1.15,
                         '11': 4.17e-10 * 1.3})


def test_two_species_24(backend):
    """Sources differ between the items of the cross-species products (the
    unshared-load path) and the phases are scaled by the mass ratios."""
    from oracle import fastspectral as fsp
    vm, prm, tab = oracle_vss(_two_species(24, 1), 2)
    f0 = fsp.two_maxwellian_batch(vm['cv'], 2, seed=51, mass=vm['masses'][0])
    f1 = fsp.two_maxwellian_batch(vm['cv'], 2, seed=52, mass=vm['masses'][1], dens_range=(0.2, 0.8))
    print('[pack vss-2 24^3]', flush=True)
    qp, _ = _run_vss(backend, _two_species(24, 1), [f0, f1], 1, 2)
    qu, _ = _run_vss(backend, _two_species(24, 0), [f0, f1], 1, 2)
    for d in range(2):
        ref = fsp.vss_fs([f0[d], f1[d]], vm, tab, prm)
        for s in range(2):
            assert rel_l2(qp[s][d], qu[s][d]) <= PACK_VS_UNPACKED
            assert rel_l2(qp[s][d], ref[s]) <= MIXED_TOL


@pytest.mark.parametrize('N,Nrho,dev', [(24, 4, 5.0), (32, 3, 6.0)])
def test_vhs_groups(backend, N, Nrho, dev):
    """Several products per group (the accumulating kernel)."""
    from oracle import fastspectral as fsp
    cfg = vhs_cfg(N, Nrho, 6, dev, precision='mixed')
    vm, prm, tab = oracle_vhs(cfg)
    f = fsp.two_maxwellian_batch(vm['cv'], 3, seed=N + 1)
    print(f'[pack vhs {N}^3]', flush=True)
    qp, _ = _run_vhs(backend, cfg, f, 1, 3)
    qu, _ = _run_vhs(backend, vhs_cfg(N, Nrho, 6, dev, precision='mixed', pack=0), f, 1, 3)
    for d in range(3):
        assert rel_l2(qp[d], qu[d]) <= PACK_VS_UNPACKED
        assert rel_l2(qp[d], fsp.vhs_fs(f[d], vm, tab, prm['prefac'])) <= MIXED_TOL


def test_nodal_24(backend):
    """Bilinear Q(f, g): the items of a product read different sources and
    the products carry the coefficient 1/2."""
    import torch
    from oracle import fastspectral as fsp
    from frfs_amd.scattering import DGFSVHSGLLNodalScatteringModel
    from frfs_amd.velocitymesh import DGFSVelocityMesh
    out = {}
    for pack in (1, 0):
        cfg = vhs_cfg(24, 3, 6, 5.0, model='vhs-gll-nodal', precision='mixed', pack=pack)
        vm, prm, tab = oracle_vhs(cfg)
        f = fsp.two_maxwellian_batch(vm['cv'], 2, seed=61)
        g = fsp.two_maxwellian_batch(vm['cv'], 2, seed=62)
        hvm = DGFSVelocityMesh(backend, cfg)
        sm = DGFSVHSGLLNodalScatteringModel(backend, cfg, hvm)
        n = hvm.vsize()
        fi = backend.matrix((1, n, 2), _io(f, 1, 2))
        gi = backend.matrix((1, n, 2), _io(g, 1, 2))
        qo = backend.matrix((1, n, 2))
        print(f'[pack nodal 24^3 pack={pack}]', flush=True)
        sm.fs_batch(fi, gi, qo)
        torch.cuda.synchronize()
        out[pack] = _from_io(qo.get())
    for d in range(2):
        assert rel_l2(out[1][d], out[0][d]) <= PACK_VS_UNPACKED
        assert rel_l2(out[1][d], fsp.vhs_nodal_fs(f[d], g[d], vm, tab, prm['prefac'])) <= MIXED_TOL


def test_single_species_32_ragged_batches(backend):
    """The Nyquist corrections are laid out per (batch slot, product): a
    batch size that does not divide the DOF count changes nothing."""
    from oracle import fastspectral as fsp
    cfg = vss_cfg(32, 3, 12, 6.0, precision='mixed')
    vm, prm, tab = oracle_vss(cfg, 1)
    f = fsp.two_maxwellian_batch(vm['cv'], 5, seed=71)
    print('[pack vss-1 32^3 ragged]', flush=True)
    qa, _ = _run_vss(backend, cfg, [f], 1, 5)
    qb, _ = _run_vss(backend, vss_cfg(32, 3, 12, 6.0, precision='mixed', max_chunk=2), [f], 1, 5)
    assert np.array_equal(qa[0], qb[0])
    for d in range(5):
        assert rel_l2(qa[0][d], fsp.vss_fs([f[d]], vm, tab, prm)[0]) <= MIXED_TOL


def test_subrange_and_in_place_24(backend):
    """A DOF sub-range and an in-place call give the DOFs of a full call bit
    for bit: the corrections are indexed by batch slot, not by DOF."""
    import torch
    from frfs_amd.inputs import mixture_device, mixture_params, to_aosoa
    from frfs_amd.scattering import DGFSBiVSSGLLScatteringModel
    from frfs_amd.velocitymesh import DGFSBiVelocityMesh
    cfg = vss_cfg(24, 3, 12, 5.0, precision='mixed', max_chunk=7)
    vm = DGFSBiVelocityMesh(backend, cfg, 1)
    sm = DGFSBiVSSGLLScatteringModel(backend, cfg, vm)
    nupts, neles = 2, 9
    fin = backend.matrix((nupts, vm.vsize(), neles))
    to_aosoa(mixture_device(vm.cv(), mixture_params(nupts * neles, 81)), fin)
    full = backend.matrix(fin.ioshape)
    print('[pack subrange 24^3]', flush=True)
    sm.fs_batch(fin, full)
    part = backend.matrix(fin.ioshape, np.full(fin.ioshape, 3.5))
    first, n = 5, 9
    sm.fs_batch(fin, part, first_dof=first, ndof=n)
    torch.cuda.synchronize()
    got, want = part.get(), full.get()
    for dof in range(nupts * neles):
        e, u = divmod(dof, nupts)
        if first <= dof < first + n:
            assert np.array_equal(got[u, :, e], want[u, :, e]), dof
        else:
            assert np.all(got[u, :, e] == 3.5), dof
    sm.fs_batch(fin, fin)
    torch.cuda.synchronize()
    assert np.array_equal(fin.get(), want)


@pytest.mark.parametrize('N,dev', [(24, 5.0), (32, 6.0)])
def test_fp64_ragged_batches(backend, N, dev):
    """FP64 packs conjugate pairs too; batches of 3 DOFs equal one batch and
    the oracle to FP64 rounding."""
    from oracle import fastspectral as fsp
    cfg = vss_cfg(N, 3, 12, dev)
    vm, prm, tab = oracle_vss(cfg, 1)
    f = fsp.two_maxwellian_batch(vm['cv'], 4, seed=90 + N)
    print(f'[pack fp64 {N}^3 ragged]', flush=True)
    qa, _ = _run_vss(backend, cfg, [f], 1, 4)
    qb, _ = _run_vss(backend, vss_cfg(N, 3, 12, dev, max_chunk=3), [f], 1, 4)
    assert np.array_equal(qa[0], qb[0])
    for d in range(4):
        assert rel_l2(qa[0][d], fsp.vss_fs([f[d]], vm, tab, prm)[0]) <= 1e-12


def test_nodal_32(backend):
    """Bilinear form on the 32^3 grid: general pairs with several products per
    group, packed against unpacked and the oracle."""
    import torch
    from oracle import fastspectral as fsp
    from frfs_amd.scattering import DGFSVHSGLLNodalScatteringModel
    from frfs_amd.velocitymesh import DGFSVelocityMesh
    out = {}
    for pack in (1, 0):
        cfg = vhs_cfg(32, 2, 6, 6.0, model='vhs-gll-nodal', precision='mixed', pack=pack)
        vm, prm, tab = oracle_vhs(cfg)
        f = fsp.two_maxwellian_batch(vm['cv'], 2, seed=63)
        g = fsp.two_maxwellian_batch(vm['cv'], 2, seed=64)
        hvm = DGFSVelocityMesh(backend, cfg)
        sm = DGFSVHSGLLNodalScatteringModel(backend, cfg, hvm)
        n = hvm.vsize()
        fi = backend.matrix((1, n, 2), _io(f, 1, 2))
        gi = backend.matrix((1, n, 2), _io(g, 1, 2))
        qo = backend.matrix((1, n, 2))
        print(f'[pack nodal 32^3 pack={pack}]', flush=True)
        sm.fs_batch(fi, gi, qo)
        torch.cuda.synchronize()
        out[pack] = _from_io(qo.get())
    for d in range(2):
        assert rel_l2(out[1][d], out[0][d]) <= PACK_VS_UNPACKED
        assert rel_l2(out[1][d], fsp.vhs_nodal_fs(f[d], g[d], vm, tab, prm['prefac'])) <= MIXED_TOL
