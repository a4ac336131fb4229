"""Batched collide driver and the time steppers around it.

The reference advances f with steppers that call ``system.collide`` (a
Python loop over every spatial DOF calling the scattering model's ``fs``,
frfs/solvers/dgfs/system.py:199-218) and then combine registers with
``axnpby`` (frfs/integrators/std/steppers.py:39-225, 354-476).  Here one
``collide`` call evaluates Q for every DOF of every bank through the C ABI,
the register arithmetic runs in ``fq_axnpby`` and a whole step can be
captured once into a HIP graph and replayed.

``rhs`` is an optional callable ``rhs(t, uin, fout)`` that writes the
advection term into bank ``fout`` (the FR spatial operator of the reference);
without it the systems integrate the space-homogeneous equation df/dt = Q(f).
"""
import gc

import numpy as np

from . import _lib
from . import autotune as _autotune
from .scattering import (DGFSBiScatteringModel, DGFSScatteringModel, subclass_where)
from .velocitymesh import DGFSBiVelocityMesh, DGFSVelocityMesh


class DGFSCollisionSystem:
    """Single species: ``nregs`` banks of ioshape (nupts, Nv^3, neles)."""

    def __init__(self, backend, cfg, nupts, neles, nregs=3, rhs=None):
        self.backend = backend
        self.cfg = cfg
        self.vm = DGFSVelocityMesh(backend, cfg)
        smcls = subclass_where(DGFSScatteringModel,
                               scattering_model=cfg.get('scattering-model', 'type'))
        self.sm = smcls(backend, cfg, self.vm)
        self.ioshape = (nupts, self.vm.vsize(), neles)
        self.banks = [backend.matrix(self.ioshape) for _ in range(nregs)]
        self.rhs = rhs
        self.nspcs = 1

    # reference signature: collide(t, uinbank, foutbank)
    def collide(self, t, uinbank, foutbank):
        self.sm.fs_batch(self.banks[uinbank], self.banks[foutbank])

    def prepare(self):
        """With ``[scattering-model] autotune = 1``, tune the collision
        schedule on the solution bank (the last bank is overwritten).  Only
        one-stream schedules: a step may be captured in a HIP graph."""
        if self.cfg.getint('scattering-model', 'autotune', 0) and not getattr(self.sm, '_tuned', False):
            self.sm.autotune(self.banks[0], self.banks[-1], candidates=_autotune.ONE_STREAM)

    def axnpby(self, dst, terms):
        """banks[dst] = a0*banks[dst] + sum(a*banks[i] for a, i in terms[1:])."""
        a0 = terms[0]
        xs = [self.banks[i]._as_parameter_ for _, i in terms[1:]]
        a = [a0] + [c for c, _ in terms[1:]]
        m = self.banks[dst]
        _lib.axnpby(m.nrow * m.leaddim, m._as_parameter_, xs, a, self.backend.stream_handle())

    def moments(self, bank, mass_ratio=1.0):
        return bulk_moments(self.vm, self.banks[bank], mass_ratio)


class DGFSBiCollisionSystem:
    """Multi species: banks[s][r] for species s and register r."""

    def __init__(self, backend, cfg, nspcs, nupts, neles, nregs=3, rhs=None):
        self.backend = backend
        self.cfg = cfg
        self.nspcs = nspcs
        self.vm = DGFSBiVelocityMesh(backend, cfg, nspcs)
        smcls = subclass_where(DGFSBiScatteringModel,
                               scattering_model=cfg.get('scattering-model', 'type'))
        self.sm = smcls(backend, cfg, self.vm)
        self.ioshape = (nupts, self.vm.vsize(), neles)
        self.banks = [[backend.matrix(self.ioshape) for _ in range(nregs)] for _ in range(nspcs)]
        self.rhs = rhs

    def collide(self, t, uinbank, foutbank):
        arr = [b[uinbank] for b in self.banks] + [b[foutbank] for b in self.banks]
        self.sm.fs_batch(*arr)

    def prepare(self):
        if self.cfg.getint('scattering-model', 'autotune', 0) and not getattr(self.sm, '_tuned', False):
            self.sm.autotune(*([b[0] for b in self.banks] + [b[-1] for b in self.banks]),
                             candidates=_autotune.ONE_STREAM)

    def axnpby(self, dst, terms):
        for sb in self.banks:
            a0 = terms[0]
            xs = [sb[i]._as_parameter_ for _, i in terms[1:]]
            a = [a0] + [c for c, _ in terms[1:]]
            m = sb[dst]
            _lib.axnpby(m.nrow * m.leaddim, m._as_parameter_, xs, a, self.backend.stream_handle())

    def moments(self, bank, species):
        return bulk_moments(self.vm, self.banks[species][bank], self.vm.masses()[species])


def bulk_moments(vm, matrix, mass_ratio=1.0):
    """Dimensional bulk properties per DOF, (ndof, 15):
    rho, ux, uy, uz, T, qx, qy, qz, Pxx, Pxy, Pxz, Pyy, Pyz, Pzz, p
    (frfs/plugins/dgfsmomwriterstd.py:142-214), computed on the device."""
    torch = matrix.backend.torch
    nupts, n, neles = matrix.ioshape
    ndof = nupts * neles
    out = torch.empty((ndof, _lib.NMOMENTS), dtype=torch.float64, device=matrix.backend.device)
    _lib.moments(vm.Nv(), vm.L(), vm.cw(), mass_ratio, matrix._as_parameter_, nupts, matrix.leaddim,
                 neles, 0, ndof, out.data_ptr(), matrix.backend.stream_handle())
    m = out.cpu().numpy()
    rho0, u0, T0 = vm.rho0(), vm.u0(), vm.T0()
    scale = np.array([rho0, u0, u0, u0, T0] + [0.5 * rho0 * u0 ** 3] * 3 + [0.5 * rho0 * u0 ** 2] * 6)
    dim = m * scale
    p = (mass_ratio * vm.R0 / vm.molarMass0()) * dim[:, 0] * dim[:, 4]
    return np.concatenate([dim, p[:, None]], axis=1)


# ---------------------------------------------------------------- steppers
class BaseStepper:
    stepper_name = None
    nregs = 3

    def __init__(self, system, graph=False):
        self.system = system
        self.graph = graph
        self._graphs = {}
        self.idxcurr = 0

    def step(self, t, dt):
        if not self.graph:
            return self._step(t, dt)
        torch = self.system.backend.torch
        g = self._graphs.get(dt)
        if g is None:
            if not getattr(self.system.sm, 'capture_safe', True):
                raise ValueError(f'scattering model {self.system.sm.scattering_model!r} synchronises with '
                                 'the host inside fs_batch and cannot be captured in a HIP graph; '
                                 'use graph=False')
            # one-time schedule tuning runs eagerly, never inside a capture
            self.system.prepare()
            # release unreachable handles now rather than from a collection
            # that could run in the middle of the capture
            gc.collect()
            # capture one full step (fixed banks, in-place update of r0); the
            # library allocates nothing and issues no memsets during a step
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._step(t, dt)
            self._graphs[dt] = g
        g.replay()
        return self.idxcurr


class StdEulerDGFSStepper(BaseStepper):
    """dgfs-euler (frfs/integrators/std/steppers.py:39-111):
    r1 = Q(r0); r2 = advection(r0); r0 = r0 + dt*r2 + dt*r1."""
    stepper_name = 'dgfs-euler'
    nregs = 3

    def _step(self, t, dt):
        sysm = self.system
        sysm.collide(t, 0, 1)
        if sysm.rhs is not None:
            sysm.rhs(t, 0, 2)
            sysm.axnpby(0, [1.0, (dt, 2)])
        sysm.axnpby(0, [1.0, (dt, 1)])
        return 0


class StdTVDRK2DGFSStepper(BaseStepper):
    """dgfs-tvd-rk2 / SSP-RK2 (frfs/integrators/std/steppers.py:122-225):
    f1 = f0 + dt L(f0);  f0 <- f0/2 + f1/2 + dt/2 L(f1),  L = Q + advection."""
    stepper_name = 'dgfs-tvd-rk2'
    nregs = 3

    def _step(self, t, dt):
        sysm = self.system
        sysm.collide(t, 0, 1)                       # r1 = Q(r0)
        if sysm.rhs is not None:
            sysm.rhs(t, 0, 2)                       # r2 = advection(r0)
            sysm.axnpby(1, [1.0, (1.0, 2)])
        sysm.axnpby(1, [dt, (1.0, 0)])              # r1 = r0 + dt*r1
        sysm.collide(t + dt, 1, 2)                  # r2 = Q(r1)
        sysm.axnpby(0, [0.5, (0.5 * dt, 2)])        # r0 = r0/2 + dt/2 r2
        if sysm.rhs is not None:
            sysm.rhs(t + dt, 1, 2)                  # r2 = advection(r1)
            sysm.axnpby(0, [1.0, (0.5 * dt, 2)])
        sysm.axnpby(0, [1.0, (0.5, 1)])             # r0 += r1/2
        return 0


STEPPERS = {c.stepper_name: c for c in (StdEulerDGFSStepper, StdTVDRK2DGFSStepper)}
# the multi-species steppers (dgfsbi-euler / dgfsbi-tvd-rk2,
# frfs/integrators/std/steppers.py:230-476) are the same schemes over banks[s][r]
STEPPERS['dgfsbi-euler'] = StdEulerDGFSStepper
STEPPERS['dgfsbi-tvd-rk2'] = StdTVDRK2DGFSStepper
