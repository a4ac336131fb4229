"""CPU oracle for the flux-reconstruction advection operator -div(c f) on a
periodic structured quad mesh.  TEST INFRASTRUCTURE ONLY (used by tests/).

Restates the reference's FR discretisation (PyFR-derived):
  BaseShape.m0/m1/m2/m3 and gbasis_coeffs (frfs/shapes.py:94-110, 244-291),
  QuadShape faces (frfs/shapes.py:368-384), tflux.mako (smats . flux),
  rusanov.mako (upwind common flux), intcflux.mako (|n| scaling, opposite
  signs on the two sides) and negdivconf.mako (-rcpdjac * div).
The operator matrices are built here a second way -- Lagrange polynomials
and a mass-matrix solve for the lifting functions -- rather than from an
orthonormal basis, and the right-hand side is assembled in the reference's
unfolded kernel order, so it checks frfs_amd/fr.py independently.
"""
import numpy as np


def _lagrange(nodes, k, x):
    others = [nodes[m] for m in range(len(nodes)) if m != k]
    out = np.ones_like(np.asarray(x, dtype=float))
    for o in others:
        out = out * (x - o) / (nodes[k] - o)
    return out


def _dlagrange(nodes, k, x):
    x = np.asarray(x, dtype=float)
    n = len(nodes)
    out = np.zeros_like(x)
    for j in range(n):
        if j == k:
            continue
        term = np.ones_like(x) / (nodes[k] - nodes[j])
        for m in range(n):
            if m != k and m != j:
                term = term * (x - nodes[m]) / (nodes[k] - nodes[m])
        out = out + term
    return out


def _points(name, n):
    if name == 'gauss-legendre':
        return np.polynomial.legendre.leggauss(n)[0]
    c = np.zeros(n)
    c[-1] = 1
    inner = np.sort(np.real(np.polynomial.legendre.Legendre(c).deriv().roots()))
    return np.concatenate(([-1.0], inner, [1.0]))


def quad_operators(order, upts='gauss-legendre', fpts='gauss-legendre'):
    n1 = order + 1
    s = _points(upts, n1)
    fs = _points(fpts, n1)
    up = np.array([(s[i], s[j]) for j in range(n1) for i in range(n1)])
    faces = [(lambda t: (t, -1.0), (0.0, -1.0)), (lambda t: (1.0, t), (1.0, 0.0)),
             (lambda t: (t, 1.0), (0.0, 1.0)), (lambda t: (-1.0, t), (-1.0, 0.0))]
    fp = np.array([proj(t) for proj, _ in faces for t in fs])
    nrm = np.array([n for _, n in faces for _ in fs])

    def nodal(pts):       # (npts, nupts): l_i(x) l_j(y), node i + n1 j
        return np.stack([_lagrange(s, i, pts[:, 0]) * _lagrange(s, j, pts[:, 1])
                         for j in range(n1) for i in range(n1)], axis=1)

    M0 = nodal(fp)
    M1x = np.stack([_dlagrange(s, i, up[:, 0]) * _lagrange(s, j, up[:, 1])
                    for j in range(n1) for i in range(n1)], axis=1)
    M1y = np.stack([_lagrange(s, i, up[:, 0]) * _dlagrange(s, j, up[:, 1])
                    for j in range(n1) for i in range(n1)], axis=1)
    # lifting: g_{f,k} in Q_p with  int_elem g phi = sum_q w_q L_k(q) phi(proj q)
    # for all phi in Q_p, face rule = Gauss-Lobatto with p+1 points
    xg, wg = np.polynomial.legendre.leggauss(n1 + 2)
    gq = np.array([(a, b) for b in xg for a in xg])
    gw = np.array([wa * wb for wb in wg for wa in wg])
    Bq = nodal(gq)
    mass = Bq.T @ (gw[:, None] * Bq)
    ql = _points('gauss-legendre-lobatto', n1)
    c = np.zeros(n1)
    c[-1] = 1
    qw = 2.0 / (n1 * (n1 - 1) * np.polynomial.legendre.legval(ql, c) ** 2)
    rhs = np.zeros((len(up), 4 * n1))
    for fi, (proj, _) in enumerate(faces):
        pq = np.array([proj(t) for t in ql])
        Bf = nodal(pq)                                  # (nq, nupts)
        for k in range(n1):
            Lk = _lagrange(fs, k, ql)
            rhs[:, fi * n1 + k] = Bf.T @ (qw * Lk)
    M3 = np.linalg.solve(mass, rhs)                     # nodal values of g at upts
    return dict(upts=up, fpts=fp, norm=nrm, M0=M0, M1x=M1x, M1y=M1y, M3=M3, n1=n1)


def fr_rhs(u, cx, cy, nelx, nely, hx, hy, ops):
    """u: (nupts, nvar, neles) with element e = ex + nelx*ey.  cx, cy: (nvar,).
    Returns -div(c u) in the reference's kernel order."""
    M0, M1x, M1y, M3, nrm, n1 = ops['M0'], ops['M1x'], ops['M1y'], ops['M3'], ops['norm'], ops['n1']
    nupts, nvar, neles = u.shape
    # tflux: transformed flux smats . (c u), smats = detJ J^{-T} = diag(hy/2, hx/2)
    Fx = (0.5 * hy) * cx[None, :, None] * u
    Fy = (0.5 * hx) * cy[None, :, None] * u
    div = np.einsum('ik,kve->ive', M1x, Fx) + np.einsum('ik,kve->ive', M1y, Fy)
    # normal transformed flux at flux points from the interpolated flux (M2)
    Fnf = (nrm[:, 0, None, None] * np.einsum('fk,kve->fve', M0, Fx)
           + nrm[:, 1, None, None] * np.einsum('fk,kve->fve', M0, Fy))
    # common fluxes on each interface (intcflux + rusanov upwind)
    uf = np.einsum('fk,kve->fve', M0, u)
    ex, ey = np.arange(neles) % nelx, np.arange(neles) // nelx
    nb = [ex + nelx * ((ey - 1) % nely), (ex + 1) % nelx + nelx * ey,
          ex + nelx * ((ey + 1) % nely), (ex - 1) % nelx + nelx * ey]
    fstar = np.zeros_like(uf)
    for f in range(4):
        fo = (f + 2) % 4
        n = nrm[f * n1]
        magn = 0.5 * hx if f % 2 == 0 else 0.5 * hy
        cn = (cx * n[0] + cy * n[1])[:, None]
        ul = uf[f * n1:(f + 1) * n1]
        ur = uf[fo * n1:(fo + 1) * n1][:, :, nb[f]]
        fn = 0.5 * cn * (ul + ur) + 0.5 * np.abs(cn) * (ul - ur)
        fstar[f * n1:(f + 1) * n1] = magn * fn
    out = div + np.einsum('if,fve->ive', M3, fstar - Fnf)
    return -(4.0 / (hx * hy)) * out
