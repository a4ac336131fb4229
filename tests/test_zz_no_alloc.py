"""fq_eval_noalloc (the entry point that allocates nothing; its batch scratch
is a static device array) against fq_eval (the handle's own scratch),
through the C ABI: two calls in a row on each handle, the second on a
smaller input, must give fq_eval's result bit for bit (the same kernels;
a DOF's result does not depend on the batch it is evaluated in).  Also a
handle created with FQ_FLAG_NO_SCRATCH, which only this entry point can
evaluate."""
import numpy as np
import pytest

from tests.cases import vhs_cfg, vss_cfg

pytestmark = pytest.mark.gpu


def _model(backend, cfg, vss):
    if vss:
        from frfs_amd.scattering import DGFSBiVSSGLLScatteringModel
        from frfs_amd.velocitymesh import DGFSBiVelocityMesh
        vm = DGFSBiVelocityMesh(backend, cfg, 1)
        return vm, DGFSBiVSSGLLScatteringModel(backend, cfg, vm)
    from frfs_amd.scattering import DGFSVHSGLLScatteringModel
    from frfs_amd.velocitymesh import DGFSVelocityMesh
    vm = DGFSVelocityMesh(backend, cfg)
    return vm, DGFSVHSGLLScatteringModel(backend, cfg, vm)


def _inputs(backend, vm, neles, seed):
    from frfs_amd.inputs import mixture_device, mixture_params, to_aosoa
    fin = backend.matrix((1, vm.vsize(), neles))
    to_aosoa(mixture_device(vm.cv(), mixture_params(neles, seed)), fin)
    return fin


def _call(h, fn, fin, qout, ndof):
    nrow, ldim, _ = fin.traits
    return fn(h, fin._as_parameter_, qout._as_parameter_, nrow, ldim, fin.ioshape[2], 0, ndof, 0)


# (label, config, vss?, DOFs of the first and of the second call).  The 32^3
# mixed case (Nrho = 2, group pairs) needs 3.85 MiB of scratch per DOF, so 66
# DOFs fit in a batch of the static scratch and its first call of 150 DOFs
# runs as three batches of 50.  At Nrho = 16 with the projection (the
# headline workload's plan) 13 DOFs fit, so 30 DOFs run as three batches of 10.
CASES = [
    ('vhs16-fp64', lambda: vhs_cfg(16, 8, 6, 5.0), False, 70, 33),
    ('vhs24-mixed-proj', lambda: vhs_cfg(24, 4, 6, 5.0, precision='mixed', project=1), False, 40, 7),
    ('vss32-mixed', lambda: vss_cfg(32, 2, 12, 6.0, precision='mixed'), True, 150, 5),
    ('vss32-mixed-proj-nrho16', lambda: vss_cfg(32, 16, 12, 6.0, precision='mixed', project=1), True, 30, 7),
]


@pytest.mark.parametrize('label,mk,vss,n1,n2', CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_eval(backend, label, mk, vss, n1, n2):
    import torch
    from frfs_amd import _lib
    lib = _lib.load()
    vm, sm = _model(backend, mk(), vss)
    h = sm._h._h
    for call, nd in ((1, n1), (2, n2)):
        fin = _inputs(backend, vm, nd, seed=40 + call)
        qa, qb = backend.matrix(fin.ioshape), backend.matrix(fin.ioshape, np.full(fin.ioshape, 3.5))
        print(f'[noalloc {label} call {call}: {nd} DOFs]', flush=True)
        _lib.check(_call(h, lib.fq_eval, fin, qa, nd))
        _lib.check(_call(h, lib.fq_eval_noalloc, fin, qb, nd))
        torch.cuda.synchronize()
        a, b = qa.get(), qb.get()
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b), f'{label} call {call}: max |diff| {np.max(np.abs(a - b))}'


def test_noalloc_on_scratchless_handle(backend):
    """A handle created with FQ_FLAG_NO_SCRATCH holds no scratch; fq_eval
    refuses it with a clear error and fq_eval_noalloc evaluates it."""
    import torch
    from frfs_amd import _lib
    lib = _lib.load()
    vm, sm = _model(backend, vhs_cfg(16, 8, 6, 5.0), False)
    prm = sm._params
    prm.c.flags |= _lib.FLAG_NO_SCRATCH
    h = _lib.Handle(prm)
    try:
        assert h.scratch_bytes() == 0
        for call, nd in ((1, 50), (2, 9)):
            fin = _inputs(backend, vm, nd, seed=50 + call)
            qa, qb = backend.matrix(fin.ioshape), backend.matrix(fin.ioshape)
            print(f'[noalloc scratchless call {call}: {nd} DOFs]', flush=True)
            sm.fs_batch(fin, qa)                      # the model's own handle (fq_eval)
            _lib.check(_call(h._h, lib.fq_eval_noalloc, fin, qb, nd))
            torch.cuda.synchronize()
            assert np.array_equal(qa.get(), qb.get())
        with pytest.raises(_lib.FQError, match='FQ_FLAG_NO_SCRATCH'):
            _lib.check(_call(h._h, lib.fq_eval, fin, qb, nd))
    finally:
        h.close()


def test_noalloc_rejects_oversized_plan(backend):
    """A plan whose single DOF needs more than the static scratch is refused
    with a clear error: vhs-gll on the 32^3 grid, FP64, unfused, M = 16,
    Nrho = 20 stages 640 inverse transforms of 512 KiB per DOF (320 MiB)."""
    from frfs_amd import _lib
    lib = _lib.load()
    vm, sm = _model(backend, vhs_cfg(32, 20, 16, 6.0, fused=0, max_chunk=1), False)
    assert sm._h.info()['n_items'] == 640
    fin = _inputs(backend, vm, 1, seed=60)
    q = backend.matrix(fin.ioshape)
    print('[noalloc oversized plan]', flush=True)
    rc = _call(sm._h._h, lib.fq_eval_noalloc, fin, q, 1)
    assert rc == _lib.ERR_UNSUPPORTED
    assert 'static scratch' in lib.fq_last_error().decode()
