"""Schedule autotuner: selection, caching and failure handling on the host
with stand-in handles, and on the GPU that the tuned handle computes the same
Q as the default schedule."""
import numpy as np
import pytest


class _FakeHandle:
    made = []

    def __init__(self, params):
        fused = not (params.c.flags & 2)
        if params.c.max_chunk == 1024 and fused:
            from frfs_amd._lib import ERR_ALLOC, FQError
            raise FQError(ERR_ALLOC, 'no memory')
        self.cost = 0.001 if not fused else (0.003 if params.c.flags & 4 else 0.005)
        self.closed = False
        _FakeHandle.made.append(self)

    def chunk_dofs(self):
        return 7

    def close(self):
        self.closed = True


class _P:
    def __init__(self):
        self.c = type('fq_params_stub', (), {})()
        self.c.flags, self.c.max_chunk = 1, 0       # FLAG_PROJECT set, automatic chunk


def test_tune_picks_fastest_and_caches(monkeypatch):
    import time
    from frfs_amd import _lib, autotune
    monkeypatch.setattr(_lib, 'Handle', _FakeHandle)
    monkeypatch.setattr(autotune, '_sync', lambda: None)
    autotune.clear_cache()
    _FakeHandle.made = []
    p = _P()
    h, rep = autotune.tune(p, lambda h: time.sleep(h.cost), key=('k', 1), reps=1)
    assert rep['choice'] == {'fused': 0, 'max_chunk': 0, 'two_lane': 0}
    assert not h.closed
    # every timed handle is released before the next is built; the winner is rebuilt
    # the fake handle cannot build the fused 1024-DOF candidates
    ntimed = len([c for c in autotune.CANDIDATES if not (c['max_chunk'] == 1024 and c['fused'])])
    assert [x.closed for x in _FakeHandle.made] == [True] * ntimed + [False]
    assert _FakeHandle.made[-1] is h
    assert any('error' in t for t in rep['timings'])                 # failed candidate reported
    assert p.c.flags == 1 | 2 and p.c.max_chunk == 0                 # params keep the choice
    h2, rep2 = autotune.tune(_P(), lambda h: 1 / 0, key=('k', 1))    # cached: nothing is run
    assert rep2['cached'] and rep2['choice'] == rep['choice']
    autotune.clear_cache()


def test_tune_raises_when_nothing_builds(monkeypatch):
    from frfs_amd import _lib, autotune

    def boom(params):
        raise _lib.FQError(_lib.ERR_ALLOC, 'no memory')
    monkeypatch.setattr(_lib, 'Handle', boom)
    autotune.clear_cache()
    p = _P()
    with pytest.raises(_lib.FQError):
        autotune.tune(p, lambda h: None, key=('none',))
    assert p.c.flags == 1 and p.c.max_chunk == 0                     # restored


def test_user_max_chunk_bounds_every_candidate(monkeypatch):
    """[scattering-model] max-chunk is an upper bound: unbounded candidates
    take it, larger ones are capped, and candidates it makes equal run once."""
    from frfs_amd import _lib, autotune
    seen = []

    class H(_FakeHandle):
        def __init__(self, params):
            seen.append((params.c.flags & 6, params.c.max_chunk))
            self.cost, self.closed = 0.0, False

    monkeypatch.setattr(_lib, 'Handle', H)
    monkeypatch.setattr(autotune, '_sync', lambda: None)
    autotune.clear_cache()
    p = _P()
    p.c.max_chunk = 256
    h, rep = autotune.tune(p, lambda h: None, key=('bounded',), reps=1)
    assert all(mc == min(256, c['max_chunk'] or 256)
               for (_, mc), c in zip(seen, [c for c in autotune.CANDIDATES if c['max_chunk'] not in (1024, 2048)]))
    # fused/auto and fused/1024 collapse into one candidate under the bound
    assert len(rep['timings']) == len(autotune.CANDIDATES) - 1
    assert p.c.max_chunk == 256
    autotune.clear_cache()


@pytest.mark.gpu
def test_tuned_schedule_matches_default(backend):
    import torch
    from frfs_amd import autotune
    from frfs_amd.inputs import mixture_device, mixture_params, to_aosoa
    from frfs_amd.scattering import DGFSBiVSSGLLScatteringModel
    from frfs_amd.velocitymesh import DGFSBiVelocityMesh
    from tests.cases import vss_cfg
    autotune.clear_cache()
    cfg = vss_cfg(16, 6, 12, 5.0)
    vm = DGFSBiVelocityMesh(backend, cfg, 1)
    n = vm.vsize()
    ndof = 40
    fin = backend.matrix((1, n, ndof))
    q0 = backend.matrix((1, n, ndof))
    q1 = backend.matrix((1, n, ndof))
    to_aosoa(mixture_device(vm.cv(), mixture_params(ndof, 11)), fin)
    ref = DGFSBiVSSGLLScatteringModel(backend, cfg, vm)
    ref.fs_batch(fin, q0)
    tuned = DGFSBiVSSGLLScatteringModel(backend, cfg, vm)
    rep = tuned.autotune(fin, q1)
    print('autotune report', rep)
    # every candidate runs once: FP64 skips the FP32-only one, and batches
    # capped at the 40 DOFs of the call merge
    want = []
    for c in autotune.CANDIDATES:
        if c.get('mixed_only'):
            continue
        c = {k: v for k, v in c.items() if k != 'mixed_only'}
        c = {**c, 'max_chunk': min(c['max_chunk'], ndof)}
        if c not in want:
            want.append(c)
    assert not rep['cached'] and len(rep['timings']) == len(want)
    q1.set(np.zeros(q1.ioshape))
    tuned.fs_batch(fin, q1)
    torch.cuda.synchronize()
    a, b = q0.get(), q1.get()
    assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(a)
    again = DGFSBiVSSGLLScatteringModel(backend, cfg, vm)
    assert again.autotune(fin, q1)['cached']
    autotune.clear_cache()


def test_schedule_keys_map_to_flags():
    """Every [scattering-model] schedule key reaches its library flag, so a
    schedule the autotuner picks can be given explicitly."""
    from frfs_amd import _lib
    from frfs_amd.scattering import _flags
    from tests.cases import vss_cfg
    assert _flags(vss_cfg()) == 0
    cfg = vss_cfg(project=1, fused=0, lanes=2, pack=0, hpair=0, overlap=1)
    cfg.set('scattering-model', 'accum-quad', 1)
    cfg.set('scattering-model', 'library-fft', 1)
    assert _flags(cfg) == (_lib.FLAG_PROJECT | _lib.FLAG_UNFUSED | _lib.FLAG_TWO_LANE | _lib.FLAG_NO_PACK
                           | _lib.FLAG_NO_HPAIR | _lib.FLAG_OVERLAP | _lib.FLAG_ACCUM_QUAD
                           | _lib.FLAG_LIBRARY_FFT)


def test_mixed_only_candidates_skip_fp64(monkeypatch):
    """Candidates marked for FP32 transforms are not built for an FP64 plan."""
    from frfs_amd import _lib, autotune
    built = []

    class H(_FakeHandle):
        def __init__(self, params):
            built.append(params.c.max_chunk)
            self.cost, self.closed = 0.0, False

    monkeypatch.setattr(_lib, 'Handle', H)
    monkeypatch.setattr(autotune, '_sync', lambda: None)
    for prec, want in ((_lib.PREC_FP64, False), (_lib.PREC_MIXED, True)):
        autotune.clear_cache()
        built.clear()
        p = _P()
        p.c.precision = prec
        autotune.tune(p, lambda h: None, key=('prec', prec), reps=1)
        assert (2048 in built) == want
    autotune.clear_cache()


def test_explicit_batches_capped_at_the_dof_count(monkeypatch):
    """With the DOF count known, no candidate's explicit batch exceeds it
    (the automatic one stays automatic), and candidates the cap makes equal
    run once."""
    from frfs_amd import _lib, autotune
    built = []

    class H(_FakeHandle):
        def __init__(self, params):
            built.append(params.c.max_chunk)
            self.cost, self.closed = 0.0, False

    monkeypatch.setattr(_lib, 'Handle', H)
    monkeypatch.setattr(autotune, '_sync', lambda: None)
    autotune.clear_cache()
    p = _P()
    h, rep = autotune.tune(p, lambda h: None, key=('ndof',), reps=1, ndof=100)
    assert max(built) == 100 and 0 in built
    chunks = [(t['max_chunk'], t.get('accum_quad', 0)) for t in rep['timings']]
    assert chunks.count((100, 1)) == 1           # the four-slot candidate, capped
    autotune.clear_cache()
