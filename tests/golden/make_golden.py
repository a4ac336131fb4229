"""Write the golden fixtures of this directory from the CPU oracle.

The reference ships no test vectors for Q(f,f); these fixtures freeze the
oracle's outputs (complex128 NumPy restatement of the reference's FP64
operator) on small seeded cases so that (a) the oracle cannot drift
silently and (b) the GPU tests have an oracle-free comparison point.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import fastspectral as fsp          # noqa: E402
from tests.cases import ARGON, oracle_vhs, oracle_vss, vhs_cfg, vss_cfg   # noqa: E402

CASES = {
    # name: (builder, description)
    'vhs_n16': 'vhs-gll omega=0.5, 16^3 on [-5,5]^3, M=6 half sphere, Nrho=8, 2 DOFs seed 101',
    'vhs_nodal_n8': 'vhs-gll-nodal, 8^3 on [-4,4]^3, M=6, Nrho=4, 1 DOF pair seeds 102/103',
    'vss_n16': 'vss-gll omega=0.74 alpha=1.4, 16^3 on [-5,5]^3, M=12, Nrho=4, Mpre=32, 1 DOF seed 104',
    'vss2_n8': 'vss-gll two species (mass 1:2), 8^3 on [-4,4]^3, M=12, Nrho=3, Mpre=32, 1 DOF seeds 105/106',
}


def build(name):
    if name == 'vhs_n16':
        cfg = vhs_cfg(16, 8, 6, 5.0)
        vm, prm, tab = oracle_vhs(cfg)
        f = fsp.two_maxwellian_batch(vm['cv'], 2, seed=101)
        q = np.stack([fsp.vhs_fs(x, vm, tab, prm['prefac']) for x in f])
        return dict(f=f, q=q)
    if name == 'vhs_nodal_n8':
        cfg = vhs_cfg(8, 4, 6, 4.0, model='vhs-gll-nodal')
        vm, prm, tab = oracle_vhs(cfg)
        f = fsp.two_maxwellian_batch(vm['cv'], 1, seed=102)
        g = fsp.two_maxwellian_batch(vm['cv'], 1, seed=103)
        return dict(f=f, g=g, q=fsp.vhs_nodal_fs(f[0], g[0], vm, tab, prm['prefac'])[None])
    if name == 'vss_n16':
        cfg = vss_cfg(16, 4, 12, 5.0)
        vm, prm, tab = oracle_vss(cfg, 1)
        f = fsp.two_maxwellian_batch(vm['cv'], 1, seed=104)
        return dict(f=f, q=fsp.vss_fs([f[0]], vm, tab, prm)[0][None])
    if name == 'vss2_n8':
        m0 = ARGON['molarMass0']
        cfg = vss_cfg(8, 3, 12, 4.0, masses=(m0, 2 * m0))
        vm, prm, tab = oracle_vss(cfg, 2)
        f0 = fsp.two_maxwellian_batch(vm['cv'], 1, seed=105, mass=vm['masses'][0])
        f1 = fsp.two_maxwellian_batch(vm['cv'], 1, seed=106, mass=vm['masses'][1], dens_range=(0.2, 0.8))
        q0, q1 = fsp.vss_fs([f0[0], f1[0]], vm, tab, prm)
        return dict(f=f0, g=f1, q=q0[None], q1=q1[None])
    raise KeyError(name)


def main():
    for name, desc in CASES.items():
        arrs = build(name)
        np.savez_compressed(os.path.join(HERE, name + '.npz'), description=np.array(desc), **arrs)
        print(name, {k: v.shape for k, v in arrs.items()})


if __name__ == '__main__':
    main()
