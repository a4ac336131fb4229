"""Learned stand-in for the collision operator: a Fourier-layer neural
operator evaluated, trained and checked by the HIP kernels declared in
include/frfs_sur.h.

The reference has no learned operator.  The stand-in takes the plugin place
of ``vhs-gll`` (frfs/solvers/dgfs/scattering.py:227-286, type key
``vhs-gll-fno``) or of single-species ``vss-gll``
(frfs/solvers/dgfsbi/scattering.py:259-330, type key ``vss-gll-fno``) on the
16^3 or 32^3 grid and keeps the spectral operator both as its teacher
(training pairs are (f, Q(f)) from the spectral ``fs_batch``) and as its
fallback: ``fs_batch`` runs the network on every DOF, measures each output's
conservation defect (the largest cosine between Q and a collision invariant
1, c, |c|^2), writes the conservation-projected network output where the
defect is within ``tol`` and evaluates the spectral operator on the rest.
A sampled spectral check guards each batch of up to 8192 DOFs as well: the
spectral operator also evaluates every ``check_every``-th DOF that the
network served, and when the median rel-L2 between the two exceeds
``check_tol`` the whole batch is evaluated by the spectral operator.

Normalisation: Q is quadratic in f, so with the density rho of f the network
maps g = f / (rho f_ref) to Q / (rho^2 q_ref) (f_ref = pi^{-3/2}, the peak of
the unit Maxwellian; q_ref is the RMS of Q / rho^2 over the training set).

Extra configuration keys: ``[scattering-model] surrogate-weights`` (an .npz
written by :meth:`FourierSurrogate.save`), ``surrogate-tol`` (default 1e-4),
``surrogate-check-every`` (default 64; 0 turns the sampled check off),
``surrogate-check-tol`` (default 3e-2), ``surrogate-activations`` (``fp32``,
the default, ``fp16`` or ``bf16``: the storage of the activations between the
kernels of the inference pass, arithmetic in FP32) and ``surrogate-layers``
(default 4).
Without weights the network is the seeded initialisation and every DOF
falls back.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np

from . import _lib
from .scattering import DGFSBiVSSGLLScatteringModel, DGFSVHSGLLScatteringModel, _stream

F_REF = math.pi ** -1.5
WIDTH, MODES = 32, 8
GRIDS = (16, 32)                               # grids the kernels are built for
ACTS = {'fp32': 0, 'fp16': 1, 'bf16': 2}      # FQ_SUR_ACTS_*


def gram(cv, cw):
    """Gram matrix of the collision invariants (1, cx, cy, cz, |c|^2) in the
    cw-weighted inner product: (inverse, diagonal)."""
    phi = np.stack([np.ones(cv.shape[1]), cv[0], cv[1], cv[2], (cv ** 2).sum(axis=0)])
    G = phi @ phi.T * cw
    return np.ascontiguousarray(np.linalg.inv(G)), np.ascontiguousarray(np.diag(G).copy())


def init_params(layers, seed, width=WIDTH, modes=MODES):
    """Seeded initial parameter vector (float64, theta layout of frfs_sur.h):
    P ~ U(-1, 1), p0 ~ U(-0.1, 0.1), R ~ U(-1, 1)/W^2 per real part,
    Wm, b, Qw ~ U(-1/sqrt(W), 1/sqrt(W)), q0 = 0."""
    rng = np.random.default_rng(seed)
    m3, W = modes ** 3, width
    parts = [rng.uniform(-1, 1, W), rng.uniform(-0.1, 0.1, W)]
    lim = 1.0 / math.sqrt(W)
    for _ in range(layers):
        parts.append(rng.uniform(-1, 1, 2 * W * W * m3) / (W * W))
        parts.append(rng.uniform(-lim, lim, W * W))
        parts.append(rng.uniform(-lim, lim, W))
    parts.append(rng.uniform(-lim, lim, W))
    parts.append(np.zeros(1))
    return np.concatenate(parts)


class FourierSurrogate:
    """The network, its parameters and optimiser state on the current device."""

    def __init__(self, backend, layers=4, seed=0, theta=None, q_ref=1.0, nv=16):
        import torch
        if nv not in GRIDS:
            raise ValueError(f'the learned stand-in is built for the grids {GRIDS}, not {nv}^3')
        self.torch = torch
        self.backend = backend
        self.dev = backend.device
        self.nv = nv
        self.shape = _lib.fq_sur_shape(_lib.SUR_ABI_VERSION, nv, WIDTH, MODES, layers)
        self.layers = layers
        self.nparams = self._sizes(0, False)[0]
        th = init_params(layers, seed) if theta is None else np.asarray(theta, dtype=np.float64)
        if th.shape != (self.nparams,):
            raise ValueError(f'theta has {th.size} entries, the network has {self.nparams}')
        self.theta = torch.tensor(th, dtype=torch.float32, device=self.dev)
        self.q_ref = float(q_ref)
        self._adam = None
        self._bufs = {}
        self._fused = None           # folded inference weights, valid for theta version _fused_ver
        self._ver, self._fused_ver = 0, -1
        self.acts = 'fp32'           # activation storage of inference: fp32, fp16 or bf16
        self.train_acts = 'fp32'     # activation storage of training (arithmetic and gradients FP32)
        # conservation penalty of training: weight lambda and (L, cw, device inverse Gram
        # matrix of the invariants) of the grid; lambda = 0 leaves the plain mean-square loss
        self.cons_weight = 0.0
        self.cons = None

    def invalidate(self):
        """Call after changing theta in place (train_step does it itself)."""
        self._ver += 1

    # -- sizes and buffers -------------------------------------------------
    def _sizes(self, batch, training):
        npar, act, work = C.c_int64(), C.c_int64(), C.c_int64()
        _lib.check(_lib.load().fq_sur_sizes(C.byref(self.shape), batch, int(training), C.byref(npar),
                                            C.byref(act), C.byref(work)))
        return npar.value, act.value, work.value

    def _buf(self, name, n, dtype=None):
        torch = self.torch
        dtype = dtype or torch.float32
        b = self._bufs.get(name)
        if b is None or b.numel() < n or b.dtype != dtype:
            b = torch.empty(max(n, 1), dtype=dtype, device=self.dev)
            self._bufs[name] = b
        return b

    def _stream(self):
        return _stream(self.backend)

    # -- network -----------------------------------------------------------
    def forward(self, g, out, training=False, fused=True, acts=None):
        """out[b] = network(g[b]) for (batch, N^3) float32 device tensors.
        Inference (training False) with two or more layers runs the folded
        first and last layers (fq_sur_forward_fused) unless fused is False;
        acts is the activation storage, 'fp32', 'fp16' or 'bf16' (default:
        self.train_acts for training, self.acts for inference)."""
        acts_type = ACTS[acts if acts is not None else (self.train_acts if training else self.acts)]
        batch = g.shape[0]
        _, act, _ = self._sizes(batch, training)
        buf = self._buf('acts_train' if training else 'acts', act)
        lib = _lib.load()
        if not training and fused and self.layers >= 2:
            if self._fused_ver != self._ver:
                if self._fused is None:
                    self._fused = self.torch.empty(_lib.SUR_FUSED_FLOATS, dtype=self.torch.float32, device=self.dev)
                _lib.check(lib.fq_sur_fuse(C.byref(self.shape), self.theta.data_ptr(), self._fused.data_ptr(),
                                           self._stream()))
                self._fused_ver = self._ver
            _lib.check(lib.fq_sur_forward_fused(C.byref(self.shape), self.theta.data_ptr(), self._fused.data_ptr(),
                                                g.data_ptr(), batch, acts_type, buf.data_ptr(),
                                                out.data_ptr(), self._stream()))
            return buf
        _lib.check(lib.fq_sur_forward(C.byref(self.shape), self.theta.data_ptr(), g.data_ptr(), batch,
                                      int(training), acts_type, buf.data_ptr(), out.data_ptr(), self._stream()))
        return buf

    def backward(self, g, acts, gout, grad, acts_kind=None):
        """grad += gradient, from the activations of forward(training=True)
        with the same storage (default self.train_acts)."""
        batch = g.shape[0]
        _, _, work = self._sizes(batch, True)
        w = self._buf('work', work)
        kind = ACTS[self.train_acts if acts_kind is None else acts_kind]
        _lib.check(_lib.load().fq_sur_backward(C.byref(self.shape), self.theta.data_ptr(), g.data_ptr(), batch,
                                               kind, acts.data_ptr(), gout.data_ptr(), grad.data_ptr(),
                                               w.data_ptr(), self._stream()))

    def gradient(self, g, target):
        """(loss sums [sum d^2, sum t^2, sum |u|^2], grad) of the mean-square
        loss plus, with cons_weight > 0, the conservation penalty
        cons_weight |u|^2 / count of the output's invariant component u."""
        torch = self.torch
        out = torch.empty_like(g)
        acts = self.forward(g, out, training=True)
        gout = torch.empty_like(g)
        sums = torch.zeros(3, dtype=torch.float64, device=self.dev)
        grad = torch.zeros(self.nparams, dtype=torch.float32, device=self.dev)
        lib = _lib.load()
        _lib.check(lib.fq_sur_loss(g.numel(), out.data_ptr(), target.data_ptr(), gout.data_ptr(),
                                   sums.data_ptr(), self._stream()))
        if self.cons_weight > 0 and self.cons is not None:
            L, cw, ginv = self.cons
            _lib.check(lib.fq_sur_cons(self.nv, L, cw, self.cons_weight, g.shape[0], out.data_ptr(),
                                       ginv.data_ptr(), gout.data_ptr(), sums.data_ptr(), self._stream()))
        self.backward(g, acts, gout, grad)
        return sums, grad

    def train_step(self, g, target, lr, betas=(0.9, 0.999), eps=1e-8):
        """One Adam step on a batch; returns the device loss sums."""
        torch = self.torch
        if self._adam is None:
            self._adam = [torch.zeros_like(self.theta), torch.zeros_like(self.theta), 0]
        sums, grad = self.gradient(g, target)
        self._adam[2] += 1
        m, v, step = self._adam
        _lib.check(_lib.load().fq_sur_adam(self.nparams, self.theta.data_ptr(), grad.data_ptr(), m.data_ptr(),
                                           v.data_ptr(), lr, betas[0], betas[1], eps, step, self._stream()))
        self.invalidate()
        return sums

    def fit(self, g_all, t_all, *, epochs=1, batch=64, lr=1e-3, lr_final=None, seed=0, time_limit=None,
            log=None):
        """Adam over shuffled mini-batches of the (g, target) rows; cosine
        decay from lr to lr_final.  Returns the per-epoch mean relative
        squared error."""
        torch = self.torch
        n = g_all.shape[0]
        nb = max(n // batch, 1)
        total = epochs * nb
        lr_final = lr * 0.05 if lr_final is None else lr_final
        gen = torch.Generator(device='cpu').manual_seed(seed)
        hist, step, t0 = [], 0, time.time()
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen).to(self.dev)
            acc = torch.zeros(3, dtype=torch.float64, device=self.dev)
            for j in range(nb):
                idx = perm[j * batch:(j + 1) * batch]
                a = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * step / total))
                acc += self.train_step(g_all[idx].contiguous(), t_all[idx].contiguous(), a)
                step += 1
            r = float(acc[0] / acc[1])
            hist.append(r)
            if log:
                log(f'epoch {ep}: relative squared error {r:.4e} ({time.time() - t0:.1f} s)')
            if time_limit is not None and time.time() - t0 > time_limit:
                break
        return hist

    def default_chunk(self):
        """DOFs per forward call: 8192 on the 16^3 grid, 1024 on the 32^3 grid
        (inference activations of 3 x 32 channels x N^3 floats per DOF)."""
        return 8192 if self.nv == 16 else 1024

    def predict(self, g, chunk=None, fused=True, acts=None, out=None):
        """Network output for (batch, N^3) float32 rows, in chunks."""
        chunk = chunk or self.default_chunk()
        out = self.torch.empty_like(g) if out is None else out
        for b0 in range(0, g.shape[0], chunk):
            self.forward(g[b0:b0 + chunk], out[b0:b0 + chunk], fused=fused, acts=acts)
        return out

    # -- persistence -------------------------------------------------------
    def save(self, path):
        np.savez(path, theta=self.theta.cpu().numpy(), q_ref=self.q_ref, layers=self.layers, nv=self.nv)

    @classmethod
    def load(cls, backend, path):
        with np.load(path, allow_pickle=False) as z:
            nv = int(z['nv']) if 'nv' in z.files else 16
            return cls(backend, layers=int(z['layers']), theta=z['theta'].astype(np.float64),
                       q_ref=float(z['q_ref']), nv=nv)


def training_pairs(model, f, chunk=8192):
    """(g, target, rho) float32/float64 device tensors for the rows of f
    (ndof, N^3): g = f / (rho f_ref), target = Q(f) / (rho^2 q) with q = 1
    (the caller rescales by q_ref), Q from the model's spectral operator."""
    import torch
    from .inputs import to_aosoa
    be, vm = model.backend, model.vm
    n = vm.vsize()
    nd = f.shape[0]
    g = torch.empty((nd, n), dtype=torch.float32, device=be.device)
    t = torch.empty_like(g)
    rho = torch.empty(nd, dtype=torch.float64, device=be.device)
    lib = _lib.load()
    for d0 in range(0, nd, chunk):
        m = min(chunk, nd - d0)
        fin = be.matrix((1, n, m))
        q = be.matrix((1, n, m))
        to_aosoa(f[d0:d0 + m], fin)
        model.fs_spectral(fin, q)
        st = _stream(be)
        nv = vm.Nv()
        _lib.check(lib.fq_sur_gather(nv, vm.cw(), F_REF, 1, fin._as_parameter_, 1, fin.leaddim, m, 0, m, None,
                                     rho[d0:].data_ptr(), g[d0:].data_ptr(), st))
        _lib.check(lib.fq_sur_gather(nv, vm.cw(), 1.0, 2, q._as_parameter_, 1, q.leaddim, m, 0, m,
                                     rho[d0:].data_ptr(), None, t[d0:].data_ptr(), st))
    return g, t, rho


class _FNOStandIn:
    """The stand-in's logic, shared by its plugins: network in front, the
    plugin's spectral operator (``fs_spectral``) as fallback and judge."""
    # fs_batch reads fallback flags on the host and branches on them
    capture_safe = False

    def _setup_standin(self):
        nv = self.vm.Nv()
        if nv not in GRIDS:
            raise ValueError(f'the learned stand-in is built for the grids {GRIDS}, not {nv}^3')
        path = self.cfg.get('scattering-model', 'surrogate-weights', '')
        layers = self.cfg.getint('scattering-model', 'surrogate-layers', 4)
        self.tol = self.cfg.getfloat('scattering-model', 'surrogate-tol', 1e-4)
        self.check_every = self.cfg.getint('scattering-model', 'surrogate-check-every', 64)
        self.check_tol = self.cfg.getfloat('scattering-model', 'surrogate-check-tol', 3e-2)
        self.inplace_runs = 4        # more runs of flagged DOFs than this: one compact batch
        self.net = (FourierSurrogate.load(self.backend, path) if path
                    else FourierSurrogate(self.backend, layers, nv=nv))
        if self.net.nv != nv:
            raise ValueError(f'surrogate-weights are for the {self.net.nv}^3 grid, the mesh is {nv}^3')
        acts = self.cfg.get('scattering-model', 'surrogate-activations', 'fp32')
        if acts not in ACTS:
            raise ValueError("[scattering-model] surrogate-activations must be 'fp32', 'fp16' or 'bf16'")
        self.net.acts = acts
        gi, gd = gram(self.vm.cv(), self.vm.cw())
        torch = self.net.torch
        self._ginv = torch.tensor(gi.reshape(-1), dtype=torch.float64, device=self.backend.device)
        self._gdiag = torch.tensor(gd, dtype=torch.float64, device=self.backend.device)
        self.last_stats = None

    def _standin_batch(self, d_arr_in, d_arr_out, first_dof=0, ndof=None, chunk=None):
        if d_arr_in._as_parameter_ == d_arr_out._as_parameter_:
            # the fallback and the sampled check read f after the network's Q
            # has been written, so Q may not alias f here
            raise ValueError(f'{self.scattering_model}: d_arr_out must not alias d_arr_in (in-place '
                             'evaluation is not supported by the learned stand-in)')
        torch = self.net.torch
        chunk = chunk or self.net.default_chunk()
        nrow, ldim, _ = d_arr_in.traits
        neles = d_arr_in.ioshape[2]
        ndof = nrow * neles - first_dof if ndof is None else ndof
        lib = _lib.load()
        st = _stream(self.backend)
        n = self.vm.vsize()
        nv = self.vm.Nv()
        cw, L = self.vm.cw(), self.vm.L()
        flags_all, defects, checks = [], [], []
        for d0 in range(first_dof, first_dof + ndof, chunk):
            m = min(chunk, first_dof + ndof - d0)
            g = self.net._buf('g', m * n)[:m * n].view(m, n)
            q = self.net._buf('q', m * n)[:m * n].view(m, n)
            rho = self.net._buf('rho', m, torch.float64)
            dfc = self.net._buf('defect', m, torch.float64)
            flag = self.net._buf('flag', m, torch.int32)
            _lib.check(lib.fq_sur_gather(nv, cw, F_REF, 1, d_arr_in._as_parameter_, nrow, ldim, neles, d0, m,
                                         None, rho.data_ptr(), g.data_ptr(), st))
            self.net.forward(g, q)
            _lib.check(lib.fq_sur_finish(nv, L, cw, self.net.q_ref, self.tol, q.data_ptr(), rho.data_ptr(),
                                         self._ginv.data_ptr(), self._gdiag.data_ptr(), d_arr_out._as_parameter_,
                                         nrow, ldim, neles, d0, m, dfc.data_ptr(), flag.data_ptr(), st))
            fl = flag[:m].cpu().numpy().copy()
            if self.check_every > 0 and not fl.all():
                med = self._sampled_check(d_arr_in, d_arr_out, d0, m, fl)
                checks.append(med)
                if med > self.check_tol:
                    fl[:] = 1                 # the whole chunk goes to the spectral operator
            flags_all.append(fl)
            defects.append(dfc[:m].cpu().numpy().copy())
        flags = np.concatenate(flags_all) if flags_all else np.zeros(0, np.int32)
        # spectral operator on the flagged DOFs: in place for a few contiguous
        # runs, else gathered into one compact batch (each DOF's result does
        # not depend on the batch it is evaluated in)
        bad = np.flatnonzero(flags)
        runs = np.split(bad, np.flatnonzero(np.diff(bad) != 1) + 1) if bad.size else []
        if len(runs) <= self.inplace_runs:
            for r in runs:
                self.fs_spectral(d_arr_in, d_arr_out, first_dof + int(r[0]), int(r.size))
        else:
            from .inputs import from_aosoa, to_aosoa
            idx = self._aosoa_index(d_arr_in, first_dof + bad)
            sin = self.backend.matrix((1, n, bad.size))
            sout = self.backend.matrix((1, n, bad.size))
            to_aosoa(d_arr_in.data[idx], sin)
            self.fs_spectral(sin, sout)
            d_arr_out.data[idx] = from_aosoa(sout)
        self.last_stats = {'ndof': int(ndof), 'fallback': int(bad.size),
                           'defect': np.concatenate(defects) if defects else np.zeros(0),
                           'check_rel_l2_median': checks,
                           'check_failed_chunks': int(sum(c > self.check_tol for c in checks))}
        return self.last_stats

    def _aosoa_index(self, d_arr, dofs):
        """(len(dofs), N^3) element offsets of the given DOFs in an AoSoA
        matrix of this grid (dof = elem * nupts + upt)."""
        torch = self.net.torch
        nrow, ldim, _ = d_arr.traits
        n = self.vm.vsize()
        soa = self.backend.soasz
        dev = self.backend.device
        d = torch.as_tensor(np.asarray(dofs), dtype=torch.int64, device=dev)
        e, u = d // nrow, d % nrow
        base = u * ldim + (e // soa) * (n * soa) + e % soa
        return base[:, None] + torch.arange(n, device=dev)[None, :] * soa

    def _sampled_check(self, d_arr_in, d_arr_out, d0, m, fl):
        """Median rel-L2, over every check_every-th of the chunk's DOFs that
        the network served, between the written Q and the spectral operator's."""
        from .inputs import from_aosoa, to_aosoa
        n = self.vm.vsize()
        pick = np.flatnonzero(fl == 0)[::self.check_every]
        idx = self._aosoa_index(d_arr_in, d0 + pick)
        sin = self.backend.matrix((1, n, len(pick)))
        sout = self.backend.matrix((1, n, len(pick)))
        to_aosoa(d_arr_in.data[idx], sin)
        self.fs_spectral(sin, sout)
        qs = from_aosoa(sout)
        qn = d_arr_out.data[idx]
        rel = ((qn - qs).pow(2).sum(1) / qs.pow(2).sum(1).clamp_min(1e-300)).sqrt()
        return float(rel.median())


class DGFSVHSGLLFNOScatteringModel(_FNOStandIn, DGFSVHSGLLScatteringModel):
    """vhs-gll with the learned stand-in in front of the spectral operator."""
    scattering_model = 'vhs-gll-fno'

    def perform_precomputation(self):
        DGFSVHSGLLScatteringModel.perform_precomputation(self)
        self._setup_standin()

    def fs(self, d_arr_in, d_arr_out, elem, upt):
        nrow = d_arr_in.traits[0]
        self.fs_batch(d_arr_in, d_arr_out, elem * nrow + upt, 1)

    def fs_spectral(self, d_arr_in, d_arr_out, first_dof=0, ndof=None):
        """The spectral operator alone (the teacher and the fallback)."""
        DGFSVHSGLLScatteringModel.fs_batch(self, d_arr_in, d_arr_out, first_dof, ndof)

    def fs_batch(self, d_arr_in, d_arr_out, first_dof=0, ndof=None, chunk=None):
        return self._standin_batch(d_arr_in, d_arr_out, first_dof, ndof, chunk)


class DGFSBiVSSGLLFNOScatteringModel(_FNOStandIn, DGFSBiVSSGLLScatteringModel):
    """Single-species vss-gll with the learned stand-in in front of the
    spectral operator; the plugin signatures of vss-gll
    (frfs/solvers/dgfsbi/scattering.py:323-330) with args = [f, Q]."""
    scattering_model = 'vss-gll-fno'

    def perform_precomputation(self):
        if self._nspcs != 1:
            raise ValueError('vss-gll-fno: the learned stand-in serves one species')
        DGFSBiVSSGLLScatteringModel.perform_precomputation(self)
        self._setup_standin()

    def fs(self, elem, upt, *args):
        nrow = args[0].traits[0]
        self._standin_batch(args[0], args[1], elem * nrow + upt, 1)

    def fs_spectral(self, d_arr_in, d_arr_out, first_dof=0, ndof=None):
        """The spectral operator alone (the teacher and the fallback)."""
        DGFSBiVSSGLLScatteringModel.fs_batch(self, d_arr_in, d_arr_out, first_dof=first_dof, ndof=ndof)

    def fs_batch(self, *args, first_dof=0, ndof=None, chunk=None):
        if len(args) != 2:
            raise ValueError('vss-gll-fno: fs_batch(f, Q)')
        return self._standin_batch(args[0], args[1], first_dof, ndof, chunk)
