"""CPU oracle for the fast spectral collision operator Q(f,f).

TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library imports this file:
only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` use it, and only as the checker.

This is a literal NumPy (complex128) restatement of the reference's FP64
operator.  Every function cites the reference file:line it follows
(paths relative to the jaisw7/frfs source tree).  The reference has no CPU
implementation of Q; its operator is CUDA C generated from Mako templates plus
cuFFT.  The restatement keeps the reference's stage order and quirks:

* cuFFT FORWARD (e^{-2 pi i}, unnormalised)  == ``np.fft.fftn``;
  cuFFT INVERSE (e^{+2 pi i}, unnormalised)  == ``n * np.fft.ifftn``
  (frfs/solvers/dgfs/cufft.py:182-183);
* the in-place ``F /= n`` inside the phase kernel (vhs-gll.mako:120);
* ``sinc(s) = sin(s + 1e-15) / (s + 1e-15)`` (vhs-gll.mako:11-14);
* complex (not real) transforms of real data, the Nyquist frequency stored
  as ``-N/2`` (frfs/solvers/dgfs/scattering.py:79);
* the real part is taken only at the output (vhs-gll.mako:246-266);
* VSS accumulates Q_p over the species pairs in ``product`` order, the
  first pair of each p assigning (vss-gll.mako:294-298).

Pins (see tests/test_oracle_*.py): Gauss-Jacobi vs Legendre/SciPy,
spherical-design invariants, FFT conventions vs a brute-force DFT, the sinc
tables vs spherical quadrature of plane waves, the exact identities
Q(cf)=c^2 Q(f) and shift equivariance, the half-sphere cos/sin form vs the
full-sphere product form, VSS(alpha=1) vs VHS, the collision invariants,
Q(Maxwellian) ~ 0, the BKW exact solution's time derivative, and for the
bilinear Q(f, g) its reduction to Q(f) at g = f and bilinearity.
"""
from __future__ import annotations

import itertools as it
import os
import re
from math import gamma as _gammafn

import numpy as np

_DESIGN_DIR = os.path.join(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__))), 'frfs_amd', 'sphericaldesign', 'symmetric')

R0 = 8.3144598        # frfs/solvers/dgfs/velocitymesh.py:8
NA = 6.0221409e+23    # frfs/solvers/dgfs/velocitymesh.py:9


# ---------------------------------------------------------------------------
# Quadrature (frfs/quadratures.py:6-55): Golub-Welsch for Gauss-Jacobi
# ---------------------------------------------------------------------------
def rjacobi(n, a, b):
    """Recurrence coefficients of Jacobi(a, b) (frfs/quadratures.py:6-26)."""
    ra, rb = np.zeros(n), np.zeros(n)
    apbp2 = 2. + a + b
    ra[0] = (b - a) / apbp2
    rb[0] = np.power(2., a + b + 1) * (
        _gammafn(a + 1.) * _gammafn(b + 1.) / _gammafn(apbp2))
    rb[1] = (4. * (a + 1.) * (b + 1.) / ((apbp2 + 1.) * apbp2 * apbp2))
    apbp2 += 2
    for i in range(1, n - 1):
        ra[i] = (b * b - a * a) / ((apbp2 - 2.) * apbp2)
        rb[i + 1] = (4. * (i + 1) * (i + 1 + a) * (i + 1 + b) * (i + 1 + a + b)
                     / ((apbp2 * apbp2 - 1) * apbp2 * apbp2))
        apbp2 += 2
    ra[n - 1] = (b * b - a * a) / ((apbp2 - 2.) * apbp2)
    return ra, rb


def gauss(n, ra, rb):
    """Eigen-solve of the Jacobi matrix (frfs/quadratures.py:28-38)."""
    rb = rb.copy()
    scal = rb[0]
    rb[:-1] = np.sqrt(rb[1:])
    z, V = np.linalg.eigh(np.diag(ra) + np.diag(rb[:-1], -1))
    zidx = np.argsort(z)
    z = np.sort(z)
    V = V[:, zidx]
    w = V[0, :]
    w = scal * (w ** 2)
    return z, w


def zwgj(n, a, b):
    """Gauss-Jacobi nodes/weights on [-1,1] (frfs/quadratures.py:40-43)."""
    ra, rb = rjacobi(n, a, b)
    return gauss(n, ra, rb)


def zwglj(n, a, b):
    """Gauss-Lobatto-Jacobi (frfs/quadratures.py:45-55)."""
    N = n - 2
    z, w = rjacobi(n, a, b)
    apb1 = a + b + 1.
    z[n - 1] = (a - b) / (2. * N + apb1 + 1.)
    w[n - 1] = (4. * (N + a + 1.) * (N + b + 1.) * (N + apb1)
                / ((2. * N + apb1) * np.power(2 * N + apb1 + 1, 2.)))
    return gauss(n, z, w)


# ---------------------------------------------------------------------------
# Spherical designs (frfs/sphericaldesign/__init__.py:45-107)
# ---------------------------------------------------------------------------
def load_design(rule, npts):
    """Stored symmetric design named ``rule`` with exactly ``npts`` points;
    among matches the highest degree wins (frfs/sphericaldesign/__init__.py:59-81)."""
    best = None
    for fn in sorted(os.listdir(_DESIGN_DIR)):
        m = re.match(r'([a-zA-Z0-9\-~+]+)-ss(\d+)(?:-m(\d+))?\.txt$', fn)
        if not m:
            continue
        rname, rqdeg, rnpts = m.group(1), int(m.group(2)), int(m.group(3) or -1)
        if (not rule or rule == rname) and npts == rnpts:
            if best is None or rqdeg > best[2]:
                best = (fn, rnpts, rqdeg)
    if best is None:
        raise ValueError('No suitable spherical quadrature rule found')
    return np.loadtxt(os.path.join(_DESIGN_DIR, best[0]), dtype=np.float64,
                      ndmin=2)


# ---------------------------------------------------------------------------
# Velocity meshes
# ---------------------------------------------------------------------------
def velocity_mesh(Nv, Nrho, M, *, dev, cmax=0.0, Tmax=273.0, T0=273.0,
                  H0=1e-3, rho0=0.00011140244750324529,
                  molarMass0=39.96e-3, quad='jacobi', qa=0.0, qb=0.0,
                  ssrule='womersley', masses=None):
    """Single species: frfs/solvers/dgfs/velocitymesh.py:25-135 (half-sphere).
    With ``masses`` (species molar masses): frfs/solvers/dgfsbi/velocitymesh.py:27-151
    (full sphere, mass-ratio-dependent S).  The bi mesh's shipped quadrature
    table references undefined names (velocitymesh.py:127-132); the intended
    Gauss-Jacobi / Lobatto rules are used here."""
    n0 = rho0 / molarMass0 * NA
    u0 = np.sqrt(2 * R0 / molarMass0 * T0)
    vsize = Nv ** 3
    _cmax = cmax / u0
    _Tmax = Tmax / T0
    L = _cmax + dev * np.sqrt(_Tmax)
    if masses is None:
        S = L * 2.0 / (3.0 + np.sqrt(2.0))
        ms = None
    else:
        ms = [m / molarMass0 for m in masses]
        mr = max(ms)
        S = min(L * 2.0 / (max(4. * mr / (1. + mr), 2.) + np.sqrt(1. + mr) + 1.),
                L * 2.0 / (max(4. * 1. / (1. + mr), 2.) + np.sqrt(1. + 1.0 / mr) + 1.0))
    R = 2 * S
    cw = (2.0 * L / Nv) ** 3
    c0 = np.linspace(-L + L / Nv, L - L / Nv, Nv)
    idx = np.arange(vsize)
    I, J, K = idx // (Nv * Nv), (idx % (Nv * Nv)) // Nv, (idx % (Nv * Nv)) % Nv
    cv = np.stack([c0[I], c0[J], c0[K]])

    rule = {'jacobi': zwgj, 'lobatto': zwglj}[quad]
    qz, qw = rule(Nrho, qa, qb)
    qz = (R / 2.0) * (1.0 + qz)
    qw = ((R - 0.0) / 2.0) * qw

    if masses is None:
        srule = load_design(ssrule, 2 * M)
        sz = srule[0:M, :]
        sw = 2 * np.pi / M
    else:
        sz = load_design(ssrule, M)
        sw = 4 * np.pi / M
    return dict(Nv=Nv, Nrho=Nrho, M=M, vsize=vsize, L=L, S=S, R=R, cw=cw,
                cv=cv, qz=qz, qw=qw, sz=sz, sw=sw, n0=n0, u0=u0, H0=H0, T0=T0,
                rho0=rho0, molarMass0=molarMass0, masses=ms)


def lvec(Nv):
    """Integer frequency vectors l (3, N^3) in C order, Nyquist stored as -N/2
    (frfs/solvers/dgfs/scattering.py:79-89)."""
    l0 = np.concatenate((np.arange(0, Nv / 2), np.arange(-Nv / 2, 0)))
    idx = np.arange(Nv ** 3)
    I, J, K = idx // (Nv * Nv), (idx % (Nv * Nv)) // Nv, (idx % (Nv * Nv)) % Nv
    return np.stack([l0[I], l0[J], l0[K]]).astype(np.float64)


def dft_fwd(x, axes):
    """cuFFT CUFFT_FORWARD: unnormalised e^{-2 pi i} transform."""
    return np.fft.fftn(x, axes=axes)


def dft_inv(x, axes):
    """cuFFT CUFFT_INVERSE: unnormalised e^{+2 pi i} transform."""
    n = 1
    for a in axes:
        n *= x.shape[a]
    return np.fft.ifftn(x, axes=axes) * n


def _sinc(s):
    """vhs-gll.mako:11-14 (offset sinc)."""
    return np.sin(s + 1e-15) / (s + 1e-15)


# ---------------------------------------------------------------------------
# VHS single species (vhs-gll): frfs/solvers/dgfs/scattering.py:44-286
# ---------------------------------------------------------------------------
def vhs_parameters(vm, omega, dRef, Tref):
    """frfs/solvers/dgfs/scattering.py:50-65 (alpha hard-coded to 1)."""
    alpha = 1.0
    gam = 2.0 * (1 - omega)
    invKn = vm['H0'] * np.sqrt(2.0) * np.pi * vm['n0'] * dRef * dRef * pow(
        Tref / vm['T0'], omega - 0.5)
    prefac = invKn * alpha / (pow(2.0, 2 - omega + alpha) * _gammafn(2.5 - omega) * np.pi)
    return dict(gamma=gam, prefac=prefac, invKn=invKn)


def vhs_tables(vm, gamma):
    """precompute_aa / precompute_bb (vhs-gll.mako:45-101); constants are formed
    in Python exactly as the template renders them (scattering.py:152-158)."""
    import math
    Nv, Nrho, M, L = vm['Nv'], vm['Nrho'], vm['M'], vm['L']
    qz, qw, sz = vm['qz'], vm['qw'], vm['sz']
    l = lvec(Nv)
    n = Nv ** 3
    a = np.empty((Nrho, M, n))
    bb1 = np.empty((Nrho, n))
    bb2 = np.zeros(n)
    cSqr = np.sqrt(l[0] * l[0] + l[1] * l[1] + l[2] * l[2])
    for p in range(Nrho):
        fac = math.pi / L * qz[p] / 2.
        for q in range(M):
            a[p, q] = fac * ((sz[q, 0] * l[0]) + (sz[q, 1] * l[1]) + (sz[q, 2] * l[2]))
        bb1[p] = (pow(qz[p], gamma + 2.) * 4 * math.pi) * _sinc((math.pi / L * qz[p] / 2.) * cSqr)
        bb2 += (qw[p] * pow(qz[p], gamma + 2.) * 16 * math.pi * math.pi) * _sinc((math.pi / L * qz[p]) * cSqr)
    return dict(a=a, bb1=bb1, bb2=bb2)


def vhs_fs(f, vm, tables, prefac):
    """One DOF of DGFSVHSGLLScatteringModel.fs (scattering.py:227-286).
    ``f``: real (N^3,) in C order.  Returns Q (N^3,)."""
    Nv, Nrho, M = vm['Nv'], vm['Nrho'], vm['M']
    n = Nv ** 3
    sw = vm['sw']
    qw = vm['qw']
    a, bb1, bb2 = tables['a'], tables['bb1'], tables['bb2']
    shp = (Nv, Nv, Nv)
    fC = f.astype(np.complex128)                              # r2z
    FTf = dft_fwd(fC.reshape(shp), (0, 1, 2)).reshape(n)       # FFT fwd
    FTf = FTf / n                                              # cosSinMul scaling
    ca, sa = np.cos(a), np.sin(a)                              # (Nrho, M, n)
    t1 = ca * FTf
    t2 = sa * FTf
    t3 = dft_inv(t1.reshape(Nrho, M, *shp), (2, 3, 4))
    t4 = dft_inv(t2.reshape(Nrho, M, *shp), (2, 3, 4))
    G = t3 * t3 + t4 * t4                                      # magSqr (complex)
    t1 = dft_fwd(G, (2, 3, 4)).reshape(Nrho, M, n)
    QG = np.zeros(n, dtype=np.complex128)
    for p in range(Nrho):                                      # computeQG
        intw_p = (2. * qw[p] * sw / n) * bb1[p]
        for q in range(M):
            QG += intw_p * t1[p, q]
    QG = dft_inv(QG.reshape(shp), (0, 1, 2)).reshape(n)        # gain
    FTf = FTf * bb2                                            # ax
    fCl = dft_inv(FTf.reshape(shp), (0, 1, 2)).reshape(n)      # loss conv
    return prefac * (QG.real - f * fCl.real)                   # output


def vhs_nodal_fs(f, g, vm, tables, prefac):
    """Bilinear Q(f,g) of DGFSVHSGLLNodalScatteringModel.fs (scattering.py:327-414)."""
    Nv, Nrho, M = vm['Nv'], vm['Nrho'], vm['M']
    n = Nv ** 3
    sw, qw = vm['sw'], vm['qw']
    a, bb1, bb2 = tables['a'], tables['bb1'], tables['bb2']
    shp = (Nv, Nv, Nv)
    FTf = dft_fwd(f.astype(np.complex128).reshape(shp), (0, 1, 2)).reshape(n) / n
    FTg = dft_fwd(g.astype(np.complex128).reshape(shp), (0, 1, 2)).reshape(n) / n
    ca, sa = np.cos(a), np.sin(a)
    t3 = dft_inv((ca * FTf).reshape(Nrho, M, *shp), (2, 3, 4))
    t4 = dft_inv((ca * FTg).reshape(Nrho, M, *shp), (2, 3, 4))
    t5 = t3 * t4
    t3 = dft_inv((sa * FTf).reshape(Nrho, M, *shp), (2, 3, 4))
    t4 = dft_inv((sa * FTg).reshape(Nrho, M, *shp), (2, 3, 4))
    t5 = t5 + t3 * t4
    t1 = dft_fwd(t5, (2, 3, 4)).reshape(Nrho, M, n)
    QG = np.zeros(n, dtype=np.complex128)
    for p in range(Nrho):
        intw_p = (2. * qw[p] * sw / n) * bb1[p]
        for q in range(M):
            QG += intw_p * t1[p, q]
    QG = dft_inv(QG.reshape(shp), (0, 1, 2)).reshape(n)
    fCl = dft_inv((FTg * bb2).reshape(shp), (0, 1, 2)).reshape(n)
    return prefac * (QG.real - f * fCl.real)


# ---------------------------------------------------------------------------
# VSS multi species (vss-gll): frfs/solvers/dgfsbi/scattering.py:46-330
# ---------------------------------------------------------------------------
def vss_parameters(vm, alpha, omega, dRef, Tref, ssrulepre, Mpre):
    """frfs/solvers/dgfsbi/scattering.py:52-91.  ``alpha`` etc. are dicts keyed
    by the pair string 'pq'."""
    szpre = load_design(ssrulepre, Mpre)
    swpre = 4 * np.pi / Mpre
    nspcs = len(vm['masses'])
    cases = list(it.product(range(nspcs), repeat=2))
    masses = vm['masses']
    eta, gam, prefac = {}, {}, {}
    for p, q in cases:
        pq = str(p) + str(q)
        eta[pq] = alpha[pq] - 1.
        gam[pq] = 2.0 * (1. - omega[pq])
        invKn = (np.sqrt(1. + masses[p] / masses[q]) * np.pi * vm['n0']
                 * dRef[pq] * dRef[pq] * pow(Tref[pq] / vm['T0'], omega[pq] - 0.5) * vm['H0'])
        mu = masses[p] * masses[q] / (masses[p] + masses[q])
        prefac[pq] = invKn * alpha[pq] / (
            np.sqrt(1. + masses[p] / masses[q]) * pow(mu, omega[pq] - 0.5)
            * pow(2., 1 + alpha[pq]) * _gammafn(2.5 - omega[pq]) * np.pi)
    return dict(cases=cases, eta=eta, gamma=gam, prefac=prefac,
                szpre=szpre, swpre=swpre, Mpre=Mpre)


def vss_tables(vm, prm):
    """precompute_a and precompute_bc_{pq} (vss-gll.mako:31-95), launched per
    (x, y) in x-major order (frfs/solvers/dgfsbi/scattering.py:172-200)."""
    import math
    Nv, Nrho, M, L = vm['Nv'], vm['Nrho'], vm['M'], vm['L']
    qz, qw, sz, sw = vm['qz'], vm['qw'], vm['sz'], vm['sw']
    szpre, swpre = prm['szpre'], prm['swpre']
    masses = vm['masses']
    l = lvec(Nv)
    n = Nv ** 3
    a = np.empty((Nrho, M, n))
    for x in range(Nrho):
        fac = math.pi / L * qz[x]
        for y in range(M):
            a[x, y] = fac * ((sz[y, 0] * l[0]) + (sz[y, 1] * l[1]) + (sz[y, 2] * l[2]))
    b, c = {}, {}
    for p, q in prm['cases']:
        pq = str(p) + str(q)
        mpq = masses[q] / (masses[p] + masses[q])
        etapq = prm['eta'][pq]
        bpq = np.empty((Nrho, M, n), dtype=np.complex128)
        cpq = np.zeros(n, dtype=np.complex128)
        for x in range(Nrho):
            fac = np.pi / L * qz[x]
            fac_b = swpre * pow(qz[x], prm['gamma'][pq] + 2)
            fac_c = qw[x] * sw * fac_b
            for y in range(M):
                bs = np.zeros(n, dtype=np.complex128)
                cs = np.zeros(n, dtype=np.complex128)
                for k in range(len(szpre)):
                    Bpq = pow(1 + sz[y, 0] * szpre[k, 0] + sz[y, 1] * szpre[k, 1]
                              + sz[y, 2] * szpre[k, 2], etapq)
                    Ld = fac * ((szpre[k, 0] * l[0]) + (szpre[k, 1] * l[1]) + (szpre[k, 2] * l[2]))
                    bs += Bpq * fac_b * np.cos(mpq * Ld) - 1j * (Bpq * fac_b * np.sin(mpq * Ld))
                    cs += Bpq * fac_c * np.cos(Ld) - 1j * (Bpq * fac_c * np.sin(Ld))
                bpq[x, y] = bs
                cpq += cs
        b[pq], c[pq] = bpq, cpq
    return dict(a=a, b=b, c=c)


def vss_Qij(pq, fp, fq, vm, tables, prm):
    """Contribution of pair (p,q) to Q_p (frfs/solvers/dgfsbi/scattering.py:259-320)."""
    Nv, Nrho, M = vm['Nv'], vm['Nrho'], vm['M']
    n = Nv ** 3
    sw, qw = vm['sw'], vm['qw']
    p, q = int(pq[0]), int(pq[1])
    mp, mq = vm['masses'][p], vm['masses'][q]
    a = tables['a']
    shp = (Nv, Nv, Nv)
    FTf = dft_fwd(fp.astype(np.complex128).reshape(shp), (0, 1, 2)).reshape(n) / n
    FTg = dft_fwd(fq.astype(np.complex128).reshape(shp), (0, 1, 2)).reshape(n) / n
    a1 = a * (mq / (mp + mq))
    a2 = a * (mp / (mp + mq))
    t1 = (np.cos(a1) + 1j * np.sin(a1)) * FTf
    t2 = (np.cos(a2) - 1j * np.sin(a2)) * FTg
    t3 = dft_inv(t1.reshape(Nrho, M, *shp), (2, 3, 4))
    t1 = dft_inv(t2.reshape(Nrho, M, *shp), (2, 3, 4))
    t2 = t3 * t1                                                # prod
    t1 = dft_fwd(t2, (2, 3, 4)).reshape(Nrho, M, n)
    b = tables['b'][pq]
    QG = np.zeros(n, dtype=np.complex128)
    for x in range(Nrho):                                       # computeQG
        intw_x = qw[x] * sw / n
        for y in range(M):
            QG += intw_x * (t1[x, y] * b[x, y])
    QG = dft_inv(QG.reshape(shp), (0, 1, 2)).reshape(n)
    L = dft_inv((tables['c'][pq] * FTg).reshape(shp), (0, 1, 2)).reshape(n)   # ax2
    return prm['prefac'][pq] * (QG.real - fp * L.real)


def vss_fs(fs, vm, tables, prm):
    """DGFSBiVSSGLLScatteringModel.fs (scattering.py:323-330): Q_p over all pairs,
    assigned at q == 0 and accumulated after (vss-gll.mako:294-298)."""
    nspcs = len(fs)
    Q = [None] * nspcs
    for p, q in prm['cases']:
        contrib = vss_Qij(str(p) + str(q), fs[p], fs[q], vm, tables, prm)
        Q[p] = contrib if q == 0 else Q[p] + contrib
    return Q


# ---------------------------------------------------------------------------
# AoSoA layout (frfs/backends/base/types.py:26-47,82-102; base.mako:5-6)
# ---------------------------------------------------------------------------
def soa_ix(a, v, nv, soasz=32):
    """SOA_IX(a, v, nv) (frfs/backends/cuda/kernels/base.mako:5-6)."""
    return ((a // soasz) * nv + v) * soasz + a % soasz


def aosoa_shape(nupts, nvar, neles, soasz=32, alignb=128, itemsize=8):
    """(nrow, leaddim) of a 3D matrix of ioshape (nupts, nvar, neles)
    (frfs/backends/base/types.py:26-47)."""
    nparr = neles - neles % -soasz
    ncol = nvar * nparr
    ldmod = alignb // itemsize
    leaddim = ncol - (ncol % -ldmod)
    return nupts, leaddim


def aosoa_pack(ary, soasz=32, alignb=128):
    """ioshape (nupts, nvar, neles) -> padded (nrow, leaddim) buffer
    (MatrixBase._pack, types.py:82-93 + CUDAMatrixBase._set)."""
    nupts, nvar, neles = ary.shape
    nrow, ld = aosoa_shape(nupts, nvar, neles, soasz, alignb)
    n, k = ary.shape[-1], soasz
    a = np.pad(ary, [(0, 0)] * (ary.ndim - 1) + [(0, -n % k)], mode='constant')
    a = a.reshape(a.shape[:-1] + (-1, k)).swapaxes(-2, -3)
    a = a.reshape(nrow, -1)
    buf = np.zeros((nrow, ld), dtype=ary.dtype)
    buf[:, :a.shape[1]] = a
    return buf


def aosoa_unpack(buf, ioshape, soasz=32):
    """Inverse of aosoa_pack (MatrixBase._unpack, types.py:95-102)."""
    nupts, nvar, neles = ioshape
    nparr = neles - neles % -soasz
    a = buf[:, :nvar * nparr].reshape(nupts, nparr // soasz, nvar, soasz)
    a = a.swapaxes(-2, -3).reshape(nupts, nvar, -1)
    return a[..., :neles]


# ---------------------------------------------------------------------------
# Inputs (frfs/solvers/dgfs/initcond.py:56-58, dgfsbi/initcond.py:71-73)
# ---------------------------------------------------------------------------
def maxwellian(cv, rho, u, T, mass=1.0):
    """Non-dimensional (species) Maxwellian."""
    u = np.asarray(u, dtype=np.float64).reshape(3, 1)
    return (rho * (mass / (np.pi * T)) ** 1.5
            * np.exp(-mass * np.sum((cv - u) ** 2, axis=0) / T))


def two_maxwellian_batch(cv, ndof, seed, mass=1.0, dens_range=(0.5, 1.5)):
    """Seeded batch of two-Maxwellian mixtures: per DOF and component,
    n~U(dens_range), u~U(-1,1)^3, T~U(0.5,2).  Returns (ndof, N^3)."""
    rng = np.random.default_rng(seed)
    out = np.empty((ndof, cv.shape[1]))
    for d in range(ndof):
        acc = np.zeros(cv.shape[1])
        for _ in range(2):
            nk = rng.uniform(*dens_range)
            uk = rng.uniform(-1, 1, 3)
            Tk = rng.uniform(0.5, 2)
            acc += maxwellian(cv, nk, uk, Tk, mass)
        out[d] = acc
    return out


def bkw(cv, t0):
    """BKW exact solution for Maxwell molecules (frfs/solvers/dgfs/initcond.py:120-140)."""
    K = 1.0 - np.exp(-t0 / 6.0)
    cSqr = np.sum(cv * cv, axis=0)
    return (1 / (2 * pow(2 * np.pi * K, 1.5)) * np.exp(-cSqr / (2.0 * K))
            * ((5 * K - 3) / K + (1 - K) / (K * K) * cSqr))


def moments(f, cv, cw, mass=1.0):
    """Collision-invariant moments sum f*{1, c, |c|^2}*cw (the quantities of
    frfs/plugins/dgfsmomwriterstd.py:142-214 before normalisation)."""
    return np.array([np.sum(f) * cw,
                     np.sum(f * cv[0]) * cw, np.sum(f * cv[1]) * cw,
                     np.sum(f * cv[2]) * cw,
                     np.sum(f * np.sum(cv * cv, axis=0)) * cw]) * mass


def bulk_moments_3d(f, cv, cw, mr=1.0):
    """Non-dimensional moments of _compute_moments_3D
    (frfs/plugins/dgfsmomwriterstd.py:142-214) for one DOF:
    rho, ux, uy, uz, T, qx, qy, qz, Pxx, Pxy, Pxz, Pyy, Pyz, Pzz."""
    mcw = mr * cw
    rho = np.sum(f) * mcw
    u = [np.dot(f, cv[d]) * mcw / rho for d in range(3)]
    c = [cv[d] - u[d] for d in range(3)]
    c2 = c[0] * c[0] + c[1] * c[1] + c[2] * c[2]
    T = np.sum(f * c2) * (2.0 / 3.0 * mcw * mr) / rho
    q = [mr * np.sum(f * c2 * c[d]) * mcw for d in range(3)]
    P = [2 * mr * np.sum(f * c[i] * c[j]) * mcw for i, j in ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))]
    return np.array([rho, *u, T, *q, *P])
