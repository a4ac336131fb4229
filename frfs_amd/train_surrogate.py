"""Train the learned stand-in on the GPU with the library's own kernels.

    python -m frfs_amd.train_surrogate --grid vss32 --pairs 50000 --epochs 20 --out fno32.npz

Training pairs come from the spectral operator itself: seeded mixtures of up
to three sheared Maxwellians (``inputs.shear_mixture_params``), Q(f) from the
FP64 operator with the conservation projection (so the targets conserve
mass, momentum and energy and the stand-in's defect check compares like with
like).  Two teachers:
  * ``vhs16``: ``vhs-gll`` on the 16^3 grid (omega = 0.5, M = 6, Nrho = 8,
    [-5, 5]^3: the operator of frfs/solvers/dgfs/scattering.py:227-286 at the
    reference's hard-sphere settings);
  * ``vss32``: single-species ``vss-gll`` on the 32^3 grid (omega = 0.74,
    alpha = 1.4, M = 12, Nrho = 16, Mpre = 32, [-6, 6]^3: the operator of
    frfs/solvers/dgfsbi/scattering.py:259-330 at the 32^3 workload).
Forward, backward, loss and Adam run in the HIP kernels of
include/frfs_sur.h; PyTorch only holds the device arrays.
"""
from __future__ import annotations

import argparse
import json
import time

from .inifile import Inifile

ARGON = dict(T0=273.0, H0=1e-3, rho0=0.00011140244750324529, molarMass0=39.96e-3)


GRIDS = ('vhs16', 'vss32')


def config(tol=1e-4, weights='', layers=4, grid='vhs16', project=1, nrho=None):
    """The stand-in plugin's configuration: vhs-gll-fno on the 16^3 grid
    (hard spheres, M = 6, Nrho = 8) or vss-gll-fno on the 32^3 grid (VSS
    omega = 0.74, alpha = 1.4, M = 12, Nrho = 16, Mpre = 32); the spectral
    operator in FP64, with the conservation projection unless project = 0."""
    if grid == 'vhs16':
        return Inifile.from_dict({
            'non-dim': ARGON,
            'constants': {'Nv': 16, 'NvBatchSize': 16 ** 3, 'Nrho': nrho or 8},
            'velocity-mesh': {'dev': 5.0, 'cmax': 0, 'Tmax': 273},
            'spherical-design-rule': {'ssrule': 'womersley', 'M': 6},
            'scattering-model': {'type': 'vhs-gll-fno', 'omega': 0.5, 'dRef': 4.17e-10, 'Tref': 273,
                                 'precision': 'double', 'conservation-projection': project,
                                 'surrogate-tol': tol, 'surrogate-weights': weights,
                                 'surrogate-layers': layers}})
    if grid == 'vss32':
        return Inifile.from_dict({
            'non-dim': ARGON,
            'constants': {'Nv': 32, 'NvBatchSize': 32 ** 3, 'Nrho': nrho or 16},
            'velocity-mesh': {'dev': 6.0, 'cmax': 0, 'Tmax': 273, 'molarMass1': ARGON['molarMass0']},
            'spherical-design-rule': {'ssrule': 'womersley', 'M': 12},
            'scattering-model': {'type': 'vss-gll-fno', 'ssrulepre': 'womersley', 'Mpre': 32,
                                 'alpha00': 1.4, 'omega00': 0.74, 'dRef00': 4.17e-10, 'Tref00': 273,
                                 'precision': 'double', 'conservation-projection': project,
                                 'surrogate-tol': tol, 'surrogate-weights': weights,
                                 'surrogate-layers': layers}})
    raise ValueError(f'grid must be one of {GRIDS}')


def model(backend, cfg):
    """The stand-in plugin of a configuration from :func:`config`."""
    from .surrogate import DGFSBiVSSGLLFNOScatteringModel, DGFSVHSGLLFNOScatteringModel
    from .velocitymesh import DGFSBiVelocityMesh, DGFSVelocityMesh
    if cfg.get('scattering-model', 'type') == 'vss-gll-fno':
        vm = DGFSBiVelocityMesh(backend, cfg, 1)
        return DGFSBiVSSGLLFNOScatteringModel(backend, cfg, vm)
    vm = DGFSVelocityMesh(backend, cfg)
    return DGFSVHSGLLFNOScatteringModel(backend, cfg, vm)


def make_pairs(sm, ndof, seed, chunk=8192):
    """(g, Q / rho^2, rho) device tensors for ndof seeded sheared mixtures."""
    import torch
    from .inputs import shear_mixture_device, shear_mixture_params
    from .surrogate import training_pairs
    n = sm.vm.vsize()
    dev = sm.backend.device
    g = torch.empty((ndof, n), dtype=torch.float32, device=dev)
    t = torch.empty_like(g)
    rho = torch.empty(ndof, dtype=torch.float64, device=dev)
    prm = shear_mixture_params(ndof, seed)
    for d0 in range(0, ndof, chunk):
        f = shear_mixture_device(sm.vm.cv(), prm[d0:d0 + chunk], device=dev)
        g[d0:d0 + chunk], t[d0:d0 + chunk], rho[d0:d0 + chunk] = training_pairs(sm, f, chunk)
    return g, t, rho


def rel_l2_rows(p, t):
    return ((p - t).pow(2).sum(1) / t.pow(2).sum(1)).sqrt()


def train(backend, *, pairs=200000, heldout=2048, epochs=20, batch=64, lr=2e-3, seed=0, time_limit=None,
          layers=4, train_acts='fp32', log=print, grid='vhs16', project=1, nrho=None, cons_weight=10.0):
    """Returns (model, report).  The model's network is trained in place;
    train_acts is the activation storage of training ('fp32' or 'bf16':
    BF16 storage with FP32 arithmetic, gradients and weights); cons_weight
    is the weight of the conservation penalty (the output's component along
    the collision invariants, which the projected targets do not have)."""
    import torch
    cfg = config(layers=layers, grid=grid, project=project, nrho=nrho)
    sm = model(backend, cfg)
    sm.net.train_acts = train_acts
    sm.net.cons_weight = cons_weight
    sm.net.cons = (sm.vm.L(), sm.vm.cw(), sm._ginv)
    t0 = time.time()
    g, t, _ = make_pairs(sm, pairs + heldout, seed)
    torch.cuda.synchronize()
    t_gen = time.time() - t0
    q_ref = float(t[:pairs].pow(2).mean().sqrt())
    t /= q_ref
    sm.net.q_ref = q_ref
    before = float(rel_l2_rows(sm.net.predict(g[pairs:]), t[pairs:]).median())
    t0 = time.time()
    hist = sm.net.fit(g[:pairs], t[:pairs], epochs=epochs, batch=batch, lr=lr, seed=seed,
                      time_limit=time_limit, log=log)
    torch.cuda.synchronize()
    t_fit = time.time() - t0
    r = rel_l2_rows(sm.net.predict(g[pairs:]), t[pairs:])
    rep = {'grid': grid, 'targets': 'FP64 operator' + (' with conservation projection' if project else ''),
           'conservation_penalty_weight': cons_weight,
           'pairs': pairs, 'heldout': heldout, 'epochs_run': len(hist), 'batch': batch, 'lr': lr,
           'layers': layers, 'train_acts': train_acts, 'generate_s': round(t_gen, 2),
           'train_s': round(t_fit, 2),
           'train_rel_sq_err': hist[-1] if hist else None,
           'heldout_rel_l2_untrained_median': before,
           'heldout_rel_l2_median': float(r.median()), 'heldout_rel_l2_p90': float(r.quantile(0.9)),
           'heldout_rel_l2_max': float(r.max())}
    del g, t
    return sm, rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--grid', default='vhs16', choices=GRIDS)
    ap.add_argument('--no-projection', action='store_true', help='train on unprojected FP64 targets')
    ap.add_argument('--cons-weight', type=float, default=10.0,
                    help='weight of the conservation penalty (0: plain mean-square loss)')
    ap.add_argument('--pairs', type=int, default=200000)
    ap.add_argument('--heldout', type=int, default=2048)
    ap.add_argument('--epochs', type=int, default=20)
    ap.add_argument('--batch', type=int, default=64)
    ap.add_argument('--lr', type=float, default=2e-3)
    ap.add_argument('--layers', type=int, default=4)
    ap.add_argument('--train-acts', default='fp32', choices=['fp32', 'fp16', 'bf16'])
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--time-limit', type=float, default=None, help='seconds; stops after the running epoch')
    ap.add_argument('--out', required=True)
    a = ap.parse_args()
    from .backend import HIPBackend
    be = HIPBackend()
    sm, rep = train(be, pairs=a.pairs, heldout=a.heldout, epochs=a.epochs, batch=a.batch, lr=a.lr,
                    seed=a.seed, time_limit=a.time_limit, layers=a.layers, train_acts=a.train_acts,
                    log=lambda s: print(s, flush=True), grid=a.grid, project=0 if a.no_projection else 1,
                    cons_weight=a.cons_weight)
    sm.net.save(a.out)
    print(json.dumps(rep), flush=True)


if __name__ == '__main__':
    main()
