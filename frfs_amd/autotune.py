"""Schedule autotuner for the collision operator.

A schedule is the triple (product stage, pencil layout, DOFs per internal
batch): the fused product kernel or the staged path that writes the inverse z
transforms to HBM; for FP32 transforms on the 24^3 and 32^3 grids, one lane
per pencil (in-register radix-4/8 transforms) or two (radix-2 exchange
between lanes); and the largest batch the scratch budget allows or a smaller
one.
``tune`` builds one handle per candidate on the current device, times each on
the caller's own inputs (one untimed call, then ``reps`` timed calls between
device synchronisations), keeps the fastest handle and releases the others.
At most two handles exist at a time.  Decisions are cached for the process,
keyed by the plan shape and the DOF count, so a later model with the same key
builds its handle from the cache without timing.
"""
import time

from . import _lib

CANDIDATES = (
    {'fused': 1, 'max_chunk': 0, 'two_lane': 0},
    {'fused': 1, 'max_chunk': 512, 'two_lane': 0},
    {'fused': 1, 'max_chunk': 0, 'two_lane': 1},
    {'fused': 0, 'max_chunk': 0, 'two_lane': 0},
)

_CACHE = {}


def _sync():
    import torch
    torch.cuda.synchronize()


def _apply(params, cand):
    c = params.c
    fl = c.flags & ~(_lib.FLAG_UNFUSED | _lib.FLAG_TWO_LANE)
    if not cand['fused']:
        fl |= _lib.FLAG_UNFUSED
    if cand.get('two_lane'):
        fl |= _lib.FLAG_TWO_LANE
    c.flags = fl
    c.max_chunk = cand['max_chunk']


def tune(params, run, key, reps=2, candidates=CANDIDATES):
    """Return ``(handle, report)``; ``run(handle)`` evaluates the operator on
    the caller's inputs.  ``params`` is left with the chosen schedule."""
    orig = (params.c.flags, params.c.max_chunk)
    if key in _CACHE:
        cand = _CACHE[key]
        _apply(params, cand)
        return _lib.Handle(params), {'cached': True, 'choice': dict(cand)}
    best, times = None, []
    for cand in candidates:
        _apply(params, cand)
        try:
            h = _lib.Handle(params)
        except _lib.FQError as e:                 # e.g. scratch allocation failed
            times.append({**cand, 'error': str(e)})
            continue
        run(h)
        _sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            run(h)
        _sync()
        dt = (time.perf_counter() - t0) / reps
        times.append({**cand, 'ms': dt * 1e3, 'chunk': h.chunk_dofs()})
        if best is None or dt < best[0]:
            if best is not None:
                best[2].close()
            best = (dt, cand, h)
        else:
            h.close()
    if best is None:
        params.c.flags, params.c.max_chunk = orig
        raise _lib.FQError(_lib.ERR_ALLOC, 'no schedule could be built: ' + str(times))
    _apply(params, best[1])
    _CACHE[key] = dict(best[1])
    return best[2], {'cached': False, 'choice': dict(best[1]), 'timings': times}


def clear_cache():
    _CACHE.clear()
