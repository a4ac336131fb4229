"""Scattering-model plugins for the fast spectral operator Q(f,f) on MI355X.

Same plugin shape as the reference (frfs/solvers/dgfs/scattering.py:16-41,
frfs/solvers/dgfsbi/scattering.py:17-43): the constructor takes
``(backend, cfg, velocitymesh)`` and calls ``load_parameters()`` then
``perform_precomputation()``; the class attribute ``scattering_model`` is the
``[scattering-model] type`` key looked up with :func:`subclass_where`.

``fs`` keeps the reference's per-DOF signature; ``fs_batch`` evaluates a
range of DOFs in one call (the batched replacement for the per-DOF loop in
DGFSSystem.collide, frfs/solvers/dgfs/system.py:217-218).  All work runs in
``libfrfs_fq.so``; there is no CPU path.

Extra configuration keys (absent from the reference, whose operator is
always FP64): ``[scattering-model] precision = double | mixed`` and
``conservation-projection = 0 | 1``; ``max-chunk`` caps the DOFs per
internal batch; ``fused = 0`` selects the three-kernel path that stages the
inverse transforms through HBM (for comparison); ``pencil-lanes = 2`` selects
the two-lanes-per-pencil FP32 product kernel on the 24^3 and 32^3 grids
(default 1: one lane per pencil); ``pair-pack = 0`` makes the one-lane kernel
transform the two items of every product separately instead of as one packed
complex volume (default 1); ``autotune = 1`` makes the
first ``fs_batch`` call time the candidate schedules on its own inputs and
keep the fastest (``frfs_amd.autotune``; ``autotune()`` does it on demand).
"""
from abc import ABCMeta, abstractmethod
from math import gamma as _gammafn

import numpy as np

from . import _lib
from . import autotune as _autotune
from .sphericaldesign import get_sphquadrule


def subclasses(cls, just_leaf=False):
    sc = cls.__subclasses__()
    ssc = [g for s in sc for g in subclasses(s, just_leaf)]
    return [s for s in sc if not just_leaf or not s.__subclasses__()] + ssc


def subclass_where(cls, **kwargs):
    """Plugin lookup by class attribute (frfs/util.py:151-159)."""
    k, v = next(iter(kwargs.items()))
    for s in subclasses(cls):
        if hasattr(s, k) and getattr(s, k) == v:
            return s
    raise KeyError("No subclasses of {0} with cls.{1} == '{2}'".format(cls.__name__, k, v))


def _precision(cfg):
    p = cfg.get('scattering-model', 'precision', 'double')
    if p not in ('double', 'mixed'):
        raise ValueError("[scattering-model] precision must be 'double' or 'mixed'")
    return _lib.PREC_MIXED if p == 'mixed' else _lib.PREC_FP64


def _flags(cfg):
    fl = _lib.FLAG_PROJECT if cfg.getint('scattering-model', 'conservation-projection', 0) else 0
    if not cfg.getint('scattering-model', 'fused', 1):
        fl |= _lib.FLAG_UNFUSED
    if cfg.getint('scattering-model', 'pencil-lanes', 1) == 2:
        fl |= _lib.FLAG_TWO_LANE
    if not cfg.getint('scattering-model', 'pair-pack', 1):
        fl |= _lib.FLAG_NO_PACK
    if not cfg.getint('scattering-model', 'group-pairs', 1):
        fl |= _lib.FLAG_NO_HPAIR
    if cfg.getint('scattering-model', 'overlap', 0):
        fl |= _lib.FLAG_OVERLAP
    if cfg.getint('scattering-model', 'accum-quad', 0):
        fl |= _lib.FLAG_ACCUM_QUAD
    if cfg.getint('scattering-model', 'library-fft', 0):
        fl |= _lib.FLAG_LIBRARY_FFT
    return fl


def _stream(backend):
    try:
        return backend.stream_handle()
    except AttributeError:
        return 0


class DGFSScatteringModel(metaclass=ABCMeta):
    # fs_batch only enqueues device work (no host syncs, allocations or
    # host-side branches on results), so a step may be captured in a HIP graph
    capture_safe = True

    def __init__(self, backend, cfg, velocitymesh):
        self.backend = backend
        self.cfg = cfg
        self.vm = velocitymesh
        self.load_parameters()
        self.perform_precomputation()

    @abstractmethod
    def load_parameters(self):
        pass

    @abstractmethod
    def perform_precomputation(self):
        pass

    @abstractmethod
    def fs(self, d_arr_in, d_arr_out, elem, upt):
        pass

    # -- helpers shared by the VHS models -----------------------------------
    def _vhs_params(self, model):
        vm = self.vm
        return _lib.Params(
            model=model, precision=_precision(self.cfg), flags=_flags(self.cfg),
            Nv=vm.Nv(), Nrho=vm.Nrho(), M=vm.M(), L=vm.L(), sw=vm.sw(), cw=vm.cw(),
            qz=vm.qz(), qw=vm.qw(), sz=vm.sz(), gamma=self._gamma,
            prefac=self._prefactor,
            max_chunk=self.cfg.getint('scattering-model', 'max-chunk', 0))

    def plan_info(self):
        return self._h.info()

    # -- schedule autotuning (shared by every model) --------------------------
    def _tune(self, run, ndof, candidates=None):
        info = self._h.info()
        cands = _autotune.CANDIDATES if candidates is None else tuple(candidates)
        key = (type(self).__name__, self._params.c.precision, self._params.c.flags & _lib.FLAG_PROJECT,
               self._params.c.Nv, info['n_items'], info['n_groups'], info['n_targets'], int(ndof),
               tuple(tuple(sorted(c.items())) for c in cands))
        self._h.close()
        self._h, self.tune_report = _autotune.tune(self._params, run, key, candidates=cands, ndof=int(ndof))
        self._tuned = True

    def _maybe_tune(self, run, ndof):
        if not getattr(self, '_tuned', False) and self.cfg.getint('scattering-model', 'autotune', 0):
            self._tune(run, ndof)


class DGFSVHSGLLScatteringModel(DGFSScatteringModel):
    """vhs-gll (frfs/solvers/dgfs/scattering.py:44-286)."""
    scattering_model = 'vhs-gll'
    _model = _lib.MODEL_VHS

    def load_parameters(self):
        # frfs/solvers/dgfs/scattering.py:50-65 (alpha hard-coded to 1)
        alpha = 1.0
        omega = self.cfg.getfloat('scattering-model', 'omega')
        self._gamma = 2.0 * (1 - omega)
        dRef = self.cfg.getfloat('scattering-model', 'dRef')
        Tref = self.cfg.getfloat('scattering-model', 'Tref')
        invKn = self.vm.H0() * np.sqrt(2.0) * np.pi * self.vm.n0() * dRef * dRef * pow(
            Tref / self.vm.T0(), omega - 0.5)
        self._Kn = 1.0 / invKn
        self._prefactor = invKn * alpha / (
            pow(2.0, 2 - omega + alpha) * _gammafn(2.5 - omega) * np.pi)

    def perform_precomputation(self):
        self._params = self._vhs_params(self._model)
        self._h = _lib.Handle(self._params)

    def fs(self, d_arr_in, d_arr_out, elem, upt):
        nrow, ldim, _ = d_arr_in.traits
        nvar, neles = d_arr_in.ioshape[1:]
        self._h.eval(d_arr_in._as_parameter_, d_arr_out._as_parameter_, nrow, ldim, neles,
                     elem * nrow + upt, 1, _stream(self.backend))

    def _runner(self, d_arr_in, d_arr_out, first_dof, ndof):
        nrow, ldim, _ = d_arr_in.traits
        nvar, neles = d_arr_in.ioshape[1:]
        ndof = nrow * neles - first_dof if ndof is None else ndof

        def run(h):
            h.eval(d_arr_in._as_parameter_, d_arr_out._as_parameter_, nrow, ldim, neles,
                   first_dof, ndof, _stream(self.backend))
        return run, ndof

    def fs_batch(self, d_arr_in, d_arr_out, first_dof=0, ndof=None):
        run, ndof = self._runner(d_arr_in, d_arr_out, first_dof, ndof)
        self._maybe_tune(run, ndof)
        run(self._h)

    def autotune(self, d_arr_in, d_arr_out, first_dof=0, ndof=None, candidates=None):
        """Time the candidate schedules (default ``autotune.CANDIDATES``) on
        these arrays and keep the fastest (``d_arr_out`` is overwritten).
        Returns the tuning report."""
        run, ndof = self._runner(d_arr_in, d_arr_out, first_dof, ndof)
        self._tune(run, ndof, candidates)
        return self.tune_report


class DGFSVHSGLLNodalScatteringModel(DGFSVHSGLLScatteringModel):
    """vhs-gll-nodal: bilinear Q(f, g) (frfs/solvers/dgfs/scattering.py:291-414)."""
    scattering_model = 'vhs-gll-nodal'
    _model = _lib.MODEL_VHS_NODAL

    def fs(self, d_arr_in1, d_arr_in2, d_arr_out, elem, upt):
        nrow, ldim, _ = d_arr_in1.traits
        nvar, neles = d_arr_in1.ioshape[1:]
        self._h.eval_bilinear(d_arr_in1._as_parameter_, d_arr_in2._as_parameter_,
                              d_arr_out._as_parameter_, nrow, ldim, neles,
                              elem * nrow + upt, 1, _stream(self.backend))

    def _runner(self, d_arr_in1, d_arr_in2, d_arr_out, first_dof, ndof):
        nrow, ldim, _ = d_arr_in1.traits
        nvar, neles = d_arr_in1.ioshape[1:]
        ndof = nrow * neles - first_dof if ndof is None else ndof

        def run(h):
            h.eval_bilinear(d_arr_in1._as_parameter_, d_arr_in2._as_parameter_,
                            d_arr_out._as_parameter_, nrow, ldim, neles,
                            first_dof, ndof, _stream(self.backend))
        return run, ndof

    def fs_batch(self, d_arr_in1, d_arr_in2, d_arr_out, first_dof=0, ndof=None):
        run, ndof = self._runner(d_arr_in1, d_arr_in2, d_arr_out, first_dof, ndof)
        self._maybe_tune(run, ndof)
        run(self._h)

    def autotune(self, d_arr_in1, d_arr_in2, d_arr_out, first_dof=0, ndof=None, candidates=None):
        run, ndof = self._runner(d_arr_in1, d_arr_in2, d_arr_out, first_dof, ndof)
        self._tune(run, ndof, candidates)
        return self.tune_report


class DGFSBiScatteringModel(metaclass=ABCMeta):
    capture_safe = True

    def __init__(self, backend, cfg, velocitymesh):
        self.backend = backend
        self.cfg = cfg
        self.vm = velocitymesh
        self._nspcs = self.vm.nspcs()
        self.load_parameters()
        self.perform_precomputation()

    @abstractmethod
    def load_parameters(self):
        pass

    @abstractmethod
    def perform_precomputation(self):
        pass

    @abstractmethod
    def fs(self, elem, upt, *args):
        pass

    def plan_info(self):
        return self._h.info()

    _tune = DGFSScatteringModel._tune
    _maybe_tune = DGFSScatteringModel._maybe_tune


class DGFSBiVSSGLLScatteringModel(DGFSBiScatteringModel):
    """vss-gll (frfs/solvers/dgfsbi/scattering.py:46-330)."""
    scattering_model = 'vss-gll'

    def load_parameters(self):
        # frfs/solvers/dgfsbi/scattering.py:52-91
        ssrulepre = self.cfg.get('scattering-model', 'ssrulepre')
        self._Mpre = self.cfg.getint('scattering-model', 'Mpre')
        srule = get_sphquadrule('symmetric', rule=ssrulepre, npts=self._Mpre)
        self._szpre = srule.pts
        self._swpre = 4 * np.pi / self._Mpre
        ns = self._nspcs
        masses = self.vm.masses()
        self._eta = np.zeros((ns, ns))
        self._gamma = np.zeros((ns, ns))
        self._prefactor = np.zeros((ns, ns))
        for p in range(ns):
            for q in range(ns):
                pq = str(p) + str(q)
                alpha = self.cfg.getfloat('scattering-model', 'alpha' + pq)
                omega = self.cfg.getfloat('scattering-model', 'omega' + pq)
                dRef = self.cfg.getfloat('scattering-model', 'dRef' + pq)
                Tref = self.cfg.getfloat('scattering-model', 'Tref' + pq)
                self._eta[p, q] = alpha - 1.
                self._gamma[p, q] = 2.0 * (1. - omega)
                invKn = (np.sqrt(1. + masses[p] / masses[q]) * np.pi * self.vm.n0()
                         * dRef * dRef * pow(Tref / self.vm.T0(), omega - 0.5) * self.vm.H0())
                mu = masses[p] * masses[q] / (masses[p] + masses[q])
                self._prefactor[p, q] = invKn * alpha / (
                    np.sqrt(1. + masses[p] / masses[q]) * pow(mu, omega - 0.5)
                    * pow(2., 1 + alpha) * _gammafn(2.5 - omega) * np.pi)

    def perform_precomputation(self):
        vm = self.vm
        self._params = _lib.Params(
            model=_lib.MODEL_VSS, precision=_precision(self.cfg), flags=_flags(self.cfg),
            Nv=vm.Nv(), Nrho=vm.Nrho(), M=vm.M(), L=vm.L(), sw=vm.sw(), cw=vm.cw(),
            qz=vm.qz(), qw=vm.qw(), sz=vm.sz(), masses=np.asarray(vm.masses(), float),
            eta=self._eta, gammapq=self._gamma, prefacpq=self._prefactor,
            szpre=self._szpre, swpre=self._swpre,
            max_chunk=self.cfg.getint('scattering-model', 'max-chunk', 0))
        self._h = _lib.Handle(self._params)

    def _runner(self, arrs, first_dof, ndof):
        ns = self._nspcs
        if len(arrs) != 2 * ns:
            raise ValueError('Need to be equal')
        nrow, ldim, _ = arrs[0].traits
        nvar, neles = arrs[0].ioshape[1:]
        for a in arrs[1:]:
            if a.traits[:2] != (nrow, ldim) or tuple(a.ioshape) != tuple(arrs[0].ioshape):
                raise ValueError('species matrices must share one AoSoA geometry')
        ndof = nrow * neles - first_dof if ndof is None else ndof
        fp = [a._as_parameter_ for a in arrs[:ns]]
        qp = [a._as_parameter_ for a in arrs[ns:]]

        def run(h):
            if ns == 1:
                h.eval(fp[0], qp[0], nrow, ldim, neles, first_dof, ndof, _stream(self.backend))
            else:
                h.eval_multi(fp, qp, nrow, ldim, neles, first_dof, ndof, _stream(self.backend))
        return run, ndof

    def fs(self, elem, upt, *args):
        nrow = args[0].traits[0]
        run, _ = self._runner(args, elem * nrow + upt, 1)
        run(self._h)

    def fs_batch(self, *args, first_dof=0, ndof=None):
        run, ndof = self._runner(args, first_dof, ndof)
        self._maybe_tune(run, ndof)
        run(self._h)

    def autotune(self, *args, first_dof=0, ndof=None, candidates=None):
        """Time the candidate schedules (default ``autotune.CANDIDATES``) on
        these arrays (f_0..f_{S-1}, Q_0..Q_{S-1}; the Q arrays are overwritten)
        and keep the fastest."""
        run, ndof = self._runner(args, first_dof, ndof)
        self._tune(run, ndof, candidates)
        return self.tune_report
