"""Velocity meshes of the DGFS solvers (host NumPy, init only).

DGFSVelocityMesh   : frfs/solvers/dgfs/velocitymesh.py:7-160 (single species,
                     half-sphere design of M points taken from a 2M design)
DGFSBiVelocityMesh : frfs/solvers/dgfsbi/velocitymesh.py:8-179 (species
                     masses, mass-ratio dependent S, full-sphere design)

Configuration keys are the reference's: [non-dim] H0 T0 rho0 molarMass0;
[constants] Nv NvBatchSize Nrho; [velocity-mesh] cmax Tmax dev quad-rule
quad-rule-alpha quad-rule-beta (+ molarMass1..n); [spherical-design-rule]
ssrule M.
"""
import numpy as np

from .quadratures import zwgj, zwglj, zwgrjm, zwgrjp
from .sphericaldesign import get_sphquadrule

_QUADS = {'jacobi': zwgj, 'lobatto': zwglj, 'radau-m': zwgrjm, 'radau-p': zwgrjp}


class _VelocityMeshBase:
    R0 = 8.3144598
    NA = 6.0221409e+23

    def __init__(self, backend, cfg):
        self.backend = backend
        self.cfg = cfg
        self._construct_velocity_mesh()
        self._load_quadrature()
        self._construct_spherical_mesh()

    def _nondim(self):
        cfg = self.cfg
        self._H0 = cfg.getfloat('non-dim', 'H0')
        self._T0 = cfg.getfloat('non-dim', 'T0')
        self._rho0 = cfg.getfloat('non-dim', 'rho0')
        self._molarMass0 = cfg.getfloat('non-dim', 'molarMass0')
        self._n0 = self._rho0 / self._molarMass0 * self.NA
        self._u0 = np.sqrt(2 * self.R0 / self._molarMass0 * self._T0)
        self._Nv = cfg.getint('constants', 'Nv')
        self._NvBatchSize = cfg.getint('constants', 'NvBatchSize', self._Nv ** 3)
        self._vsize = self._Nv ** 3
        if self._vsize < self._NvBatchSize:
            raise ValueError('NvBatchSize must not exceed Nv^3')
        self._NvBatches = -(-self._vsize // self._NvBatchSize)
        if self._NvBatches * self._NvBatchSize != self._vsize:
            raise ValueError('Nv^3 must be a multiple of NvBatchSize')
        self._Nrho = cfg.getint('constants', 'Nrho', int(self._Nv / 2))
        cmax = cfg.getfloat('velocity-mesh', 'cmax') / self._u0
        Tmax = cfg.getfloat('velocity-mesh', 'Tmax') / self._T0
        dev = cfg.getfloat('velocity-mesh', 'dev')
        self._L = cmax + dev * np.sqrt(Tmax)

    def _grid(self):
        N, L = self._Nv, self._L
        self._cw = (2.0 * L / N) ** 3
        c0 = np.linspace(-L + L / N, L - L / N, N)
        idx = np.arange(self._vsize)
        self._cv = np.stack([c0[idx // (N * N)], c0[(idx // N) % N], c0[idx % N]])

    def _load_quadrature(self):
        rule = self.cfg.get('velocity-mesh', 'quad-rule', 'jacobi')
        if rule not in _QUADS:
            raise ValueError('Valid quads: ' + str(sorted(_QUADS)))
        a = self.cfg.getfloat('velocity-mesh', 'quad-rule-alpha', 0.0)
        b = self.cfg.getfloat('velocity-mesh', 'quad-rule-beta', 0.0)
        z, w = _QUADS[rule](self._Nrho, a, b)
        self._qz = (self._R / 2.0) * (1.0 + z)
        self._qw = ((self._R - 0.0) / 2.0) * w

    def Nv(self): return self._Nv
    def Nrho(self): return self._Nrho
    def NvBatches(self): return self._NvBatches
    def NvBatchSize(self): return self._NvBatchSize
    def M(self): return self._M
    def L(self): return self._L
    def S(self): return self._S
    def R(self): return self._R
    def vsize(self): return self._vsize
    def qz(self): return self._qz
    def qw(self): return self._qw
    def cv(self): return self._cv
    def cw(self): return self._cw
    def sz(self): return self._sz
    def sw(self): return self._sw
    def H0(self): return self._H0
    def T0(self): return self._T0
    def rho0(self): return self._rho0
    def n0(self): return self._n0
    def u0(self): return self._u0
    def molarMass0(self): return self._molarMass0


class DGFSVelocityMesh(_VelocityMeshBase):
    def _construct_velocity_mesh(self):
        self._nondim()
        self._S = self._L * 2.0 / (3.0 + np.sqrt(2.0))
        self._R = 2 * self._S
        self._grid()

    def _construct_spherical_mesh(self):
        self._ssrule = self.cfg.get('spherical-design-rule', 'ssrule')
        self._M = self.cfg.getint('spherical-design-rule', 'M')
        srule = get_sphquadrule('symmetric', rule=self._ssrule, npts=2 * self._M)
        self._sz = srule.pts[0:self._M, :]      # half sphere
        self._sw = 2 * np.pi / self._M


class DGFSBiVelocityMesh(_VelocityMeshBase):
    def __init__(self, backend, cfg, nspcs):
        self._nspcs = nspcs
        super().__init__(backend, cfg)

    def _construct_velocity_mesh(self):
        self._nondim()
        self._masses = [self.cfg.getfloat('velocity-mesh', 'molarMass' + str(p + 1))
                        / self._molarMass0 for p in range(self._nspcs)]
        mr = self._mr = max(self._masses)
        L = self._L
        self._S = min(L * 2.0 / (max(4. * mr / (1. + mr), 2.) + np.sqrt(1. + mr) + 1.),
                      L * 2.0 / (max(4. * 1. / (1. + mr), 2.) + np.sqrt(1. + 1.0 / mr) + 1.0))
        self._R = 2 * self._S
        self._grid()

    def _construct_spherical_mesh(self):
        self._ssrule = self.cfg.get('spherical-design-rule', 'ssrule')
        self._M = self.cfg.getint('spherical-design-rule', 'M')
        srule = get_sphquadrule('symmetric', rule=self._ssrule, npts=self._M)
        self._sz = srule.pts                     # full sphere
        self._sw = 4 * np.pi / self._M

    def masses(self): return self._masses
    def nspcs(self): return self._nspcs
