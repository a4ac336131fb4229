"""Edge cases of the batched boundary (fq_eval through the plugin classes):
ragged batches, DOF sub-ranges, empty ranges, in-place evaluation, AoSoA
padding with several solution points, and range errors.  Every DOF is
computed independently, so batching must not change a single bit."""
import numpy as np
import pytest

from tests.cases import vhs_cfg, vss_cfg

pytestmark = pytest.mark.gpu


def _model(backend, cfg, vss=False):
    if vss:
        from frfs_amd.scattering import DGFSBiVSSGLLScatteringModel
        from frfs_amd.velocitymesh import DGFSBiVelocityMesh
        vm = DGFSBiVelocityMesh(backend, cfg, 1)
        return vm, DGFSBiVSSGLLScatteringModel(backend, cfg, vm)
    from frfs_amd.scattering import DGFSVHSGLLScatteringModel
    from frfs_amd.velocitymesh import DGFSVelocityMesh
    vm = DGFSVelocityMesh(backend, cfg)
    return vm, DGFSVHSGLLScatteringModel(backend, cfg, vm)


def _inputs(backend, vm, nupts, neles, seed):
    from frfs_amd.inputs import mixture_device, mixture_params, to_aosoa
    fin = backend.matrix((nupts, vm.vsize(), neles))
    to_aosoa(mixture_device(vm.cv(), mixture_params(nupts * neles, seed)), fin)
    return fin


@pytest.mark.parametrize('prec', ['double', 'mixed'])
def test_ragged_batches_bitwise(backend, prec):
    """37 DOFs in batches of at most 8 (five uneven batches) equal one batch."""
    import torch
    vm, one = _model(backend, vhs_cfg(16, 4, 6, 5.0, precision=prec))
    _, small = _model(backend, vhs_cfg(16, 4, 6, 5.0, precision=prec, max_chunk=8))
    assert small._h.chunk_dofs() == 8
    fin = _inputs(backend, vm, 1, 37, seed=5)
    qa, qb = backend.matrix(fin.ioshape), backend.matrix(fin.ioshape)
    print(f'[ragged {prec}]', flush=True)
    one.fs_batch(fin, qa)
    small.fs_batch(fin, qb)
    torch.cuda.synchronize()
    assert np.array_equal(qa.get(), qb.get())


def test_subrange_empty_and_untouched(backend):
    """A DOF sub-range writes exactly those DOFs; an empty range writes nothing."""
    import torch
    vm, sm = _model(backend, vhs_cfg(16, 4, 6, 5.0))
    nupts, neles = 3, 37                       # AoSoA padding: 37 elements -> 64 columns
    fin = _inputs(backend, vm, nupts, neles, seed=6)
    full = backend.matrix(fin.ioshape)
    print('[subrange]', flush=True)
    sm.fs_batch(fin, full)
    sentinel = np.full(fin.ioshape, 7.25)
    part = backend.matrix(fin.ioshape, sentinel)
    sm.fs_batch(fin, part, first_dof=0, ndof=0)
    torch.cuda.synchronize()
    assert np.array_equal(part.get(), sentinel)
    first, n = 17, 23                          # DOF = elem * nupts + upt
    sm.fs_batch(fin, part, first_dof=first, ndof=n)
    torch.cuda.synchronize()
    got, want = part.get(), full.get()
    for dof in range(nupts * neles):
        e, u = divmod(dof, nupts)
        if first <= dof < first + n:
            assert np.array_equal(got[u, :, e], want[u, :, e]), dof
        else:
            assert np.all(got[u, :, e] == 7.25), dof


def test_in_place_equals_out_of_place(backend):
    """Q may alias f: every DOF reads its own f before its Q is written."""
    import torch
    cfg = vss_cfg(16, 4, 12, 5.0, precision='mixed')
    vm, sm = _model(backend, cfg, vss=True)
    fin = _inputs(backend, vm, 2, 20, seed=8)
    out = backend.matrix(fin.ioshape)
    print('[in-place]', flush=True)
    sm.fs_batch(fin, out)
    sm.fs_batch(fin, fin)
    torch.cuda.synchronize()
    assert np.array_equal(fin.get(), out.get())


def test_range_errors(backend):
    from frfs_amd._lib import FQError
    vm, sm = _model(backend, vhs_cfg(16, 4, 6, 5.0))
    fin = _inputs(backend, vm, 1, 4, seed=9)
    q = backend.matrix(fin.ioshape)
    with pytest.raises(FQError):
        sm.fs_batch(fin, q, first_dof=2, ndof=3)     # past nupts * neles
    with pytest.raises(FQError):
        sm.fs_batch(fin, q, first_dof=-1, ndof=1)


@pytest.mark.parametrize('N,Nrho,dev,per_dof', [(16, 5, 5.0, True), (8, 4, 4.0, True), (24, 3, 5.0, False)])
def test_staged_path_odd_grids(backend, N, Nrho, dev, per_dof):
    """The staged product path (fused = 0) on launch grids that are not a
    multiple of 8: per-DOF fs (one DOF per call) at 16^3 with Nrho = 5 and
    at 8^3 with Nrho = 4, and an odd DOF count times groups at 24^3.  Every
    (DOF, group, plane) unit must run exactly once."""
    from oracle import fastspectral as fsp
    from tests.cases import oracle_vhs, rel_l2
    from tests.test_gpu_parity import _run_vhs
    cfg = vhs_cfg(N, Nrho, 6, dev, fused=0)
    vm, prm, tab = oracle_vhs(cfg)
    f = fsp.two_maxwellian_batch(vm['cv'], 3, seed=N + Nrho)
    print(f'[staged N={N} Nrho={Nrho} per_dof={per_dof}]', flush=True)
    Q, _ = _run_vhs(backend, cfg, f, 1, 3, per_dof=per_dof)
    errs = [rel_l2(Q[d], fsp.vhs_fs(f[d], vm, tab, prm['prefac'])) for d in range(3)]
    print(N, Nrho, max(errs))
    assert max(errs) <= 1e-12
