"""ctypes binding of ``libfrfs_fq.so`` (declared in ``include/frfs_fq.h``).

Errors from the library raise :class:`FQError` carrying the library's
message, the way the reference maps cuFFT status codes to exceptions
(frfs/solvers/dgfs/cufft.py:48-117).  A missing library raises at load time;
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIBPATH = os.environ.get('FRFS_FQ_LIB') or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), 'libfrfs_fq.so')

ABI_VERSION = 1

SUCCESS, ERR_INVALID, ERR_UNSUPPORTED, ERR_HIP, ERR_ALLOC = 0, -1, -2, -3, -4
NKERNELS = 10
AXNPBY_MAX = 4
NMOMENTS = 14
MODEL_VHS, MODEL_VHS_NODAL, MODEL_VSS = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
FLAG_PROJECT = 1
FLAG_UNFUSED = 2
FLAG_TWO_LANE = 4
FLAG_NO_PACK = 8
FLAG_NO_SCRATCH = 16
FLAG_ACCUM_QUAD = 32
FLAG_NO_HPAIR = 64
FLAG_OVERLAP = 128
FLAG_LIBRARY_FFT = 256
PATHS = ((1, 'fused'), (2, 'one_lane'), (4, 'pair_pack'), (8, 'group_pairs'), (16, 'overlap'),
         (32, 'library_fft'))
NOALLOC_SCRATCH_BYTES = 256 << 20

_dp = C.POINTER(C.c_double)


class FQError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f'frfs_fq error {code}: {msg}')
        self.code = code


class fq_params(C.Structure):
    _fields_ = [
        ('abi_version', C.c_int32), ('model', C.c_int32), ('precision', C.c_int32),
        ('flags', C.c_int32), ('Nv', C.c_int32), ('Nrho', C.c_int32), ('M', C.c_int32),
        ('nspcs', C.c_int32), ('Mpre', C.c_int32), ('soasz', C.c_int32),
        ('max_chunk', C.c_int64),
        ('L', C.c_double), ('sw', C.c_double), ('cw', C.c_double),
        ('qz', _dp), ('qw', _dp), ('sz', _dp),
        ('gamma', C.c_double), ('prefac', C.c_double),
        ('masses', _dp), ('eta', _dp), ('gammapq', _dp), ('prefacpq', _dp),
        ('szpre', _dp), ('swpre', C.c_double),
    ]


SUR_ABI_VERSION = 1
SUR_MAX_LAYERS = 8
SUR_MAX_BATCH = 65536
SUR_FUSED_FLOATS = 65697


class fq_sur_shape(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ('abi_version', 'Nv', 'width', 'modes', 'layers')]


class fq_plan_info(C.Structure):
    _fields_ = [(n, C.c_int32) for n in (
        'n_items', 'n_products', 'n_groups', 'n_contribs', 'n_loss',
        'n_sources', 'n_targets', 'n_fft_ref', 'n_fft')]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


# Every symbol the header declares, with its signature.
_i32, _i64, _vp = C.c_int32, C.c_int64, C.c_void_p
SIGNATURES = {
    'fq_create': (C.c_int, [C.POINTER(fq_params), C.POINTER(_vp)]),
    'fq_destroy': (C.c_int, [_vp]),
    'fq_last_error': (C.c_char_p, []),
    'fq_get_plan_info': (C.c_int, [_vp, C.POINTER(fq_plan_info)]),
    'fq_scratch_bytes': (C.c_int, [_vp, C.POINTER(_i64)]),
    'fq_eval': (C.c_int, [_vp, _vp, _vp, _i32, _i64, _i32, _i64, _i64, _vp]),
    'fq_eval_noalloc': (C.c_int, [_vp, _vp, _vp, _i32, _i64, _i32, _i64, _i64, _vp]),
    'fq_eval_multi': (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp), _i32, _i64, _i32, _i64, _i64, _vp]),
    'fq_eval_bilinear': (C.c_int, [_vp, _vp, _vp, _vp, _i32, _i64, _i32, _i64, _i64, _vp]),
    'fq_profile_enable': (C.c_int, [_vp, C.c_int]),
    'fq_profile_read': (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(_i64)]),
    'fq_kernel_name': (C.c_char_p, [C.c_int]),
    'fq_chunk_dofs': (C.c_int, [_vp, C.POINTER(_i64)]),
    'fq_get_paths': (C.c_int, [_vp, C.POINTER(C.c_int32)]),
    'fq_axnpby': (C.c_int, [_i64, _vp, _i32, C.POINTER(_vp), C.POINTER(C.c_double), _vp]),
    'fq_copy_dgfs': (C.c_int, [_i32, _i32, _i32, _vp, _i64, _i32, _i32, _vp, _i64, _i32, _i32, _vp]),
    'fq_moments': (C.c_int, [_i32, C.c_double, C.c_double, C.c_double, _vp, _i32, _i64, _i32, _i64, _i64,
                             _vp, _vp]),
    'fq_fr_rhs_quad': (C.c_int, [_i32, _i32, _i32, C.c_double, C.c_double, _vp, _vp, _vp, _vp, _vp, _i64,
                                 _i32, _vp]),
    'fq_plan_create': (C.c_int, [C.POINTER(fq_params), C.POINTER(_vp)]),
    'fq_plan_destroy': (C.c_int, [_vp]),
    'fq_plan_get_info': (C.c_int, [_vp, C.POINTER(fq_plan_info)]),
    'fq_plan_export': (C.c_int, [_vp] + [_vp] * 10),
    # include/frfs_sur.h
    'fq_sur_sizes': (C.c_int, [_vp, _i64, _i32, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)]),
    'fq_sur_forward': (C.c_int, [_vp, _vp, _vp, _i64, _i32, _i32, _vp, _vp, _vp]),
    'fq_sur_fuse': (C.c_int, [_vp, _vp, _vp, _vp]),
    'fq_sur_forward_fused': (C.c_int, [_vp, _vp, _vp, _vp, _i64, _i32, _vp, _vp, _vp]),
    'fq_sur_backward': (C.c_int, [_vp, _vp, _vp, _i64, _i32, _vp, _vp, _vp, _vp, _vp]),
    'fq_sur_loss': (C.c_int, [_i64, _vp, _vp, _vp, _vp, _vp]),
    'fq_sur_cons': (C.c_int, [_i32, C.c_double, C.c_double, C.c_double, _i64, _vp, _vp, _vp, _vp, _vp]),
    'fq_sur_adam': (C.c_int, [_i64, _vp, _vp, _vp, _vp, C.c_float, C.c_float, C.c_float, C.c_float,
                              _i32, _vp]),
    'fq_sur_gather': (C.c_int, [_i32, C.c_double, C.c_double, _i32, _vp, _i32, _i64, _i32, _i64, _i64,
                                _vp, _vp, _vp, _vp]),
    'fq_sur_finish': (C.c_int, [_i32, C.c_double, C.c_double, C.c_double, C.c_double, _vp, _vp, _vp, _vp,
                                _vp, _i32, _i64, _i32, _i64, _i64, _vp, _vp, _vp]),
}

_lib = None


def load():
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise FQError(ERR_UNSUPPORTED,
                          f'{LIBPATH} not built; run python -m frfs_amd.build')
        lib = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(code):
    if code != SUCCESS:
        raise FQError(code, load().fq_last_error().decode())


def _arr(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


class Params:
    """Owns the numpy arrays an :class:`fq_params` points into."""

    def __init__(self, *, model, Nv, Nrho, M, L, sw, cw, qz, qw, sz,
                 precision=PREC_FP64, flags=0, gamma=0.0, prefac=1.0,
                 masses=None, eta=None, gammapq=None, prefacpq=None,
                 szpre=None, swpre=0.0, max_chunk=0):
        self._keep = []
        p = fq_params()
        p.abi_version = ABI_VERSION
        p.model, p.precision, p.flags = model, precision, flags
        p.Nv, p.Nrho, p.M = Nv, Nrho, M
        p.soasz = 32
        p.max_chunk = max_chunk
        p.L, p.sw, p.cw = L, sw, cw
        p.qz = self._ptr(qz)
        p.qw = self._ptr(qw)
        p.sz = self._ptr(np.asarray(sz).reshape(-1))
        p.gamma, p.prefac = gamma, prefac
        if masses is not None:
            ns = len(masses)
            p.nspcs = ns
            p.masses = self._ptr(masses)
            p.eta = self._ptr(np.asarray(eta).reshape(ns * ns))
            p.gammapq = self._ptr(np.asarray(gammapq).reshape(ns * ns))
            p.prefacpq = self._ptr(np.asarray(prefacpq).reshape(ns * ns))
            p.szpre = self._ptr(np.asarray(szpre).reshape(-1))
            p.Mpre = len(np.asarray(szpre).reshape(-1, 3))
            p.swpre = swpre
        else:
            p.nspcs = 1
        self.c = p

    def _ptr(self, a):
        a, ptr = _arr(a)
        self._keep.append(a)
        return ptr


class Plan:
    """Host-only plan (no GPU needed)."""

    def __init__(self, params: Params):
        lib = load()
        self._h = C.c_void_p()
        check(lib.fq_plan_create(C.byref(params.c), C.byref(self._h)))

    def info(self):
        i = fq_plan_info()
        check(load().fq_plan_get_info(self._h, C.byref(i)))
        return i.as_dict()

    def export(self):
        i = self.info()
        ni, npr, nc, nl = i['n_items'], i['n_products'], i['n_contribs'], i['n_loss']
        out = dict(item_src=np.zeros(ni, np.int32), item_theta=np.zeros((ni, 3)),
                   prod_i1=np.zeros(npr, np.int32), prod_i2=np.zeros(npr, np.int32),
                   prod_coef=np.zeros(npr), prod_group=np.zeros(npr, np.int32),
                   contrib_group=np.zeros(nc, np.int32), contrib_target=np.zeros(nc, np.int32),
                   loss_src=np.zeros(nl, np.int32), loss_target=np.zeros(nl, np.int32))
        keys = ['item_src', 'item_theta', 'prod_i1', 'prod_i2', 'prod_coef', 'prod_group',
                'contrib_group', 'contrib_target', 'loss_src', 'loss_target']
        check(load().fq_plan_export(self._h, *[out[k].ctypes.data_as(C.c_void_p) for k in keys]))
        return out

    def __del__(self):
        if getattr(self, '_h', None) and _lib is not None:
            _lib.fq_plan_destroy(self._h)
            self._h = None


class Handle:
    """Device handle: tables + scratch on the current HIP device."""

    def __init__(self, params: Params):
        lib = load()
        self._h = C.c_void_p()
        check(lib.fq_create(C.byref(params.c), C.byref(self._h)))
        self.model = params.c.model
        self.nspcs = params.c.nspcs

    def info(self):
        i = fq_plan_info()
        check(load().fq_get_plan_info(self._h, C.byref(i)))
        return i.as_dict()

    def scratch_bytes(self):
        b = C.c_int64()
        check(load().fq_scratch_bytes(self._h, C.byref(b)))
        return b.value

    def chunk_dofs(self):
        b = C.c_int64()
        check(load().fq_chunk_dofs(self._h, C.byref(b)))
        return b.value

    def paths(self):
        """Names of the kernel paths this handle runs (fq_get_paths)."""
        b = C.c_int32()
        check(load().fq_get_paths(self._h, C.byref(b)))
        return [name for bit, name in PATHS if b.value & bit]

    def profile(self, on=True):
        check(load().fq_profile_enable(self._h, 1 if on else 0))

    def profile_read(self):
        """{kernel name: (total ms, launches)} since the last read."""
        ms = (C.c_double * NKERNELS)()
        n = (C.c_int64 * NKERNELS)()
        check(load().fq_profile_read(self._h, ms, n))
        lib = load()
        return {lib.fq_kernel_name(k).decode(): (ms[k], n[k]) for k in range(NKERNELS)}

    def eval(self, f_ptr, q_ptr, nupts, ldim, neles, first_dof, ndof, stream=0):
        check(load().fq_eval(self._h, f_ptr, q_ptr, nupts, ldim, neles, first_dof, ndof, stream))

    def eval_noalloc(self, f_ptr, q_ptr, nupts, ldim, neles, first_dof, ndof, stream=0):
        check(load().fq_eval_noalloc(self._h, f_ptr, q_ptr, nupts, ldim, neles, first_dof, ndof, stream))

    def eval_multi(self, f_ptrs, q_ptrs, nupts, ldim, neles, first_dof, ndof, stream=0):
        n = len(f_ptrs)
        fa = (C.c_void_p * n)(*f_ptrs)
        qa = (C.c_void_p * n)(*q_ptrs)
        check(load().fq_eval_multi(self._h, fa, qa, nupts, ldim, neles, first_dof, ndof, stream))

    def eval_bilinear(self, f_ptr, g_ptr, q_ptr, nupts, ldim, neles, first_dof, ndof, stream=0):
        check(load().fq_eval_bilinear(self._h, f_ptr, g_ptr, q_ptr, nupts, ldim, neles,
                                      first_dof, ndof, stream))

    def close(self):
        if getattr(self, '_h', None) and _lib is not None:
            _lib.fq_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


def axnpby(n, x0, xs, a, stream=0):
    """x0 = a[0] x0 + sum a[k+1] xs[k] over n doubles (device pointers)."""
    nv = len(xs)
    xa = (C.c_void_p * max(nv, 1))(*xs)
    aa = (C.c_double * (nv + 1))(*a)
    check(load().fq_axnpby(n, x0, nv, xa, aa, stream))


def copy_dgfs(nrow, neles, count, src, ldim_src, nvar_src, sidx_src, dst, ldim_dst, nvar_dst, sidx_dst,
              stream=0):
    check(load().fq_copy_dgfs(nrow, neles, count, src, ldim_src, nvar_src, sidx_src, dst, ldim_dst,
                              nvar_dst, sidx_dst, stream))


def moments(Nv, L, cw, mass_ratio, f, nupts, ldim, neles, first_dof, ndof, out, stream=0):
    check(load().fq_moments(Nv, L, cw, mass_ratio, f, nupts, ldim, neles, first_dof, ndof, out, stream))
