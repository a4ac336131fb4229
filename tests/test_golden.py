"""Golden fixtures (tests/golden/*.npz, written by make_golden.py).

CPU: the oracle reproduces every fixture.  GPU: the HIP operator matches the
fixtures to the FP64 tolerance (rel-L2 <= 1e-12) without running the oracle.
"""
import os

import numpy as np
import pytest

from tests.cases import ARGON, rel_l2, vhs_cfg, vss_cfg

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
NAMES = ['vhs_n16', 'vhs_nodal_n8', 'vss_n16', 'vss2_n8']


def _load(name):
    with np.load(os.path.join(HERE, name + '.npz'), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize('name', NAMES)
def test_oracle_reproduces_fixture(name):
    import sys
    sys.path.insert(0, HERE)
    from make_golden import build
    got = build(name)
    want = _load(name)
    for k, v in got.items():
        assert np.max(np.abs(v - want[k])) <= 1e-13 * np.max(np.abs(want[k])), (name, k)


def _io(f):
    ndof, n = f.shape
    return f.T.reshape(1, n, ndof).copy()


@pytest.mark.gpu
@pytest.mark.parametrize('name', NAMES)
def test_gpu_matches_fixture(backend, name):
    import torch
    from frfs_amd.scattering import (DGFSBiVSSGLLScatteringModel, DGFSVHSGLLNodalScatteringModel,
                                     DGFSVHSGLLScatteringModel)
    from frfs_amd.velocitymesh import DGFSBiVelocityMesh, DGFSVelocityMesh
    z = _load(name)
    f = z['f']
    ndof, n = f.shape
    print(f'[golden {name}] ndof={ndof} n={n}', flush=True)
    mk = lambda a=None: backend.matrix((1, n, ndof), None if a is None else _io(a))
    if name == 'vhs_n16':
        cfg = vhs_cfg(16, 8, 6, 5.0)
        sm = DGFSVHSGLLScatteringModel(backend, cfg, DGFSVelocityMesh(backend, cfg))
        fi, qo = mk(f), mk()
        sm.fs_batch(fi, qo)
        outs = [(qo, z['q'])]
    elif name == 'vhs_nodal_n8':
        cfg = vhs_cfg(8, 4, 6, 4.0, model='vhs-gll-nodal')
        sm = DGFSVHSGLLNodalScatteringModel(backend, cfg, DGFSVelocityMesh(backend, cfg))
        fi, gi, qo = mk(f), mk(z['g']), mk()
        sm.fs_batch(fi, gi, qo)
        outs = [(qo, z['q'])]
    elif name == 'vss_n16':
        cfg = vss_cfg(16, 4, 12, 5.0)
        sm = DGFSBiVSSGLLScatteringModel(backend, cfg, DGFSBiVelocityMesh(backend, cfg, 1))
        fi, qo = mk(f), mk()
        sm.fs_batch(fi, qo)
        outs = [(qo, z['q'])]
    else:
        m0 = ARGON['molarMass0']
        cfg = vss_cfg(8, 3, 12, 4.0, masses=(m0, 2 * m0))
        sm = DGFSBiVSSGLLScatteringModel(backend, cfg, DGFSBiVelocityMesh(backend, cfg, 2))
        f0, f1, q0, q1 = mk(f), mk(z['g']), mk(), mk()
        sm.fs_batch(f0, f1, q0, q1)
        outs = [(q0, z['q']), (q1, z['q1'])]
    torch.cuda.synchronize()
    for mat, want in outs:
        got = mat.get()[0].T
        for d in range(ndof):
            assert rel_l2(got[d], want[d]) <= 1e-12, (name, d)
