"""Schedule autotuner for the collision operator.

A schedule is (product stage, pencil layout, pair packing, accumulation tile,
DOFs per internal batch): the fused product kernel or the staged path that
writes the inverse z transforms to HBM; for FP32 transforms on the 24^3 and
32^3 grids, one lane per pencil (in-register radix-4/8 transforms) or two
(radix-2 exchange between lanes), and the two items of a product transformed
as one packed volume or separately; on the 32^3 grid, eight or four DOFs per
accumulation workgroup sharing each weight row, and one product volume per
group pair or per group; batches on one stream, or alternating between two
scratch sets and two streams; and the batch the scratch
budget allows, a larger or a smaller one (the largest for FP32 transforms only).
``tune`` builds one handle per candidate on the current device, times each on
the caller's own inputs (one untimed call, then ``reps`` timed calls between
device synchronisations) and releases it before the next is built, so at most
one handle (one scratch allocation) exists at a time; the fastest schedule is
then built once more and returned.  The caller's ``max-chunk`` is an upper
bound on every candidate's batch.  Decisions are cached for the process,
keyed by the plan shape, the DOF count and that bound, so a later model with
the same key builds its handle from the cache without timing.
"""
import time

from . import _lib

CANDIDATES = (
    {'fused': 1, 'max_chunk': 0, 'two_lane': 0},
    {'fused': 1, 'max_chunk': 1024, 'two_lane': 0},
    {'fused': 1, 'max_chunk': 128, 'two_lane': 0},
    # FP32 transforms only: four DOF slots per accumulation workgroup and a
    # 2048-DOF batch (37 GiB of scratch at 32^3), which fills whole rounds of
    # the device in the Nyquist and projection kernels
    {'fused': 1, 'max_chunk': 2048, 'two_lane': 0, 'accum_quad': 1, 'mixed_only': 1},
    {'fused': 1, 'max_chunk': 0, 'two_lane': 0, 'no_pack': 1},
    {'fused': 1, 'max_chunk': 0, 'two_lane': 0, 'no_hpair': 1},
    {'fused': 1, 'max_chunk': 512, 'two_lane': 0, 'overlap': 1},
    {'fused': 1, 'max_chunk': 0, 'two_lane': 1},
    {'fused': 0, 'max_chunk': 0, 'two_lane': 0},
)

# the candidates that run every internal batch on the caller's stream: with two
# streams the kernels of consecutive batches share the device, so per-kernel
# launch times (and a roofline taken from them) no longer describe one kernel
ONE_STREAM = tuple(c for c in CANDIDATES if not c.get('overlap'))

_CACHE = {}


def _sync():
    import torch
    torch.cuda.synchronize()


def _bounded(cand, limit):
    """The candidate with its batch capped by the caller's limit (0: none)."""
    mc = cand['max_chunk']
    if limit > 0:
        mc = limit if mc <= 0 else min(mc, limit)
    return {**cand, 'max_chunk': mc}


def _apply(params, cand):
    c = params.c
    fl = c.flags & ~(_lib.FLAG_UNFUSED | _lib.FLAG_TWO_LANE | _lib.FLAG_NO_PACK | _lib.FLAG_ACCUM_QUAD
                     | _lib.FLAG_NO_HPAIR | _lib.FLAG_OVERLAP)
    if not cand['fused']:
        fl |= _lib.FLAG_UNFUSED
    if cand.get('two_lane'):
        fl |= _lib.FLAG_TWO_LANE
    if cand.get('no_pack'):
        fl |= _lib.FLAG_NO_PACK
    if cand.get('accum_quad'):
        fl |= _lib.FLAG_ACCUM_QUAD
    if cand.get('no_hpair'):
        fl |= _lib.FLAG_NO_HPAIR
    if cand.get('overlap'):
        fl |= _lib.FLAG_OVERLAP
    c.flags = fl
    c.max_chunk = cand['max_chunk']


def tune(params, run, key, reps=3, candidates=CANDIDATES, ndof=0):
    """Return ``(handle, report)``; ``run(handle)`` evaluates the operator on
    the caller's inputs (``ndof`` DOFs when known: no candidate's explicit
    batch then exceeds it, so a small problem never holds scratch for more
    DOFs than it has).  ``params`` is left with the chosen schedule."""
    orig = (params.c.flags, params.c.max_chunk)
    limit = int(orig[1])
    key = (key, limit)
    if key in _CACHE:
        cand = _CACHE[key]
        _apply(params, cand)
        return _lib.Handle(params), {'cached': True, 'choice': dict(cand)}
    best, times, seen = None, [], []
    mixed = getattr(params.c, 'precision', _lib.PREC_MIXED) == _lib.PREC_MIXED
    for cand in candidates:
        if cand.get('mixed_only') and not mixed:
            continue
        cand = _bounded({k: v for k, v in cand.items() if k != 'mixed_only'}, limit)
        if ndof > 0 and cand['max_chunk'] > ndof:
            cand = {**cand, 'max_chunk': int(ndof)}
        if cand in seen:                           # two candidates the limit made equal
            continue
        seen.append(cand)
        _apply(params, cand)
        try:
            h = _lib.Handle(params)
        except _lib.FQError as e:                 # e.g. scratch allocation failed
            times.append({**cand, 'error': str(e)})
            continue
        try:
            run(h)
            _sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                run(h)
            _sync()
            dt = (time.perf_counter() - t0) / reps
            times.append({**cand, 'ms': dt * 1e3, 'chunk': h.chunk_dofs()})
        finally:
            h.close()                              # one scratch allocation at a time
        if best is None or dt < best[0]:
            best = (dt, cand)
    if best is None:
        params.c.flags, params.c.max_chunk = orig
        raise _lib.FQError(_lib.ERR_ALLOC, 'no schedule could be built: ' + str(times))
    _apply(params, best[1])
    h = _lib.Handle(params)
    _CACHE[key] = dict(best[1])
    return h, {'cached': False, 'choice': dict(best[1]), 'timings': times}


def clear_cache():
    _CACHE.clear()
