"""Flux-reconstruction advection operator -div(c f) on a periodic structured
quadrilateral mesh, for every velocity of the discrete velocity set at once.

The operator follows the reference's FR formulation (PyFR-derived):
  * solution points: tensor-product line points (``gauss-legendre`` or
    ``gauss-legendre-lobatto``) ordered with the first coordinate fastest
    (frfs/quadrules/quad/*.txt);
  * flux points on the faces (s,-1), (1,s), (s,1), (-1,s) in that order
    (QuadShape.faces, frfs/shapes.py:368-384);
  * M0: interpolation to flux points; M1: gradient at solution points;
    M2 = n . M0; M3: correction (lifting) basis from an L2 projection of
    the face nodal bases with a (p+1)-point Gauss-Lobatto face rule
    (BaseShape.m0-m3, gbasis_coeffs, frfs/shapes.py:94-110, 244-291);
  * transformed flux smats.(c f) (tflux.mako), upwind common normal flux
    scaled by |n| (rusanov.mako, intcflux.mako) and -rcpdjac * div
    (negdivconf.mako):
        rhs = -rcpdjac [ (M1 - M3 M2) F~ + M3 F*_common ].
For an affine element of size hx x hy this folds, per velocity (cx, cy),
into rhs = -rcpdjac [ (hy/2) cx Dx u + (hx/2) cy Dy u + M3 F* ] with
Dx, Dy = M1_d - M3 diag(n_d) M0, which the HIP kernel applies.
"""
import numpy as np

from . import _lib


def line_points(name, n):
    if name == 'gauss-legendre':
        return np.polynomial.legendre.leggauss(n)[0]
    if name == 'gauss-legendre-lobatto':
        if n == 1:
            return np.zeros(1)
        c = np.zeros(n)
        c[-1] = 1
        inner = np.polynomial.legendre.Legendre(c).deriv().roots()
        return np.concatenate(([-1.0], np.sort(np.real(inner)), [1.0]))
    raise ValueError('unknown point set ' + name)


def gll_rule(n):
    """n-point Gauss-Lobatto-Legendre nodes and weights on [-1, 1]."""
    x = line_points('gauss-legendre-lobatto', n)
    c = np.zeros(n)
    c[-1] = 1
    Pn1 = np.polynomial.legendre.legval(x, c)
    w = 2.0 / (n * (n - 1) * Pn1 ** 2)
    return x, w


def ortho1d(p, x):
    """Orthonormal Legendre polynomials sqrt((2k+1)/2) P_k(x), k <= p: (p+1, len(x))."""
    x = np.asarray(x, dtype=float)
    out = np.empty((p + 1, x.size))
    for k in range(p + 1):
        c = np.zeros(k + 1)
        c[k] = 1
        out[k] = np.sqrt((2 * k + 1) / 2.0) * np.polynomial.legendre.legval(x, c)
    return out


def dortho1d(p, x):
    x = np.asarray(x, dtype=float)
    out = np.empty((p + 1, x.size))
    for k in range(p + 1):
        c = np.zeros(k + 1)
        c[k] = 1
        out[k] = np.sqrt((2 * k + 1) / 2.0) * np.polynomial.legendre.legval(
            x, np.polynomial.legendre.legder(c))
    return out


class QuadFROperators:
    """Reference-element operator matrices for quads of order p."""

    def __init__(self, order, upts='gauss-legendre', fpts='gauss-legendre'):
        p = self.order = order
        n1 = p + 1
        s = line_points(upts, n1)
        self.upts = np.array([(s[i], s[j]) for j in range(n1) for i in range(n1)])
        fs = line_points(fpts, n1)
        self.fs = fs
        faces = [(lambda t: (t, -1.0), (0.0, -1.0)), (lambda t: (1.0, t), (1.0, 0.0)),
                 (lambda t: (t, 1.0), (0.0, 1.0)), (lambda t: (-1.0, t), (-1.0, 0.0))]
        self.fpts = np.array([proj(t) for proj, _ in faces for t in fs])
        self.norm_fpts = np.array([nrm for _, nrm in faces for _ in fs])
        self.nupts, self.nfpts = n1 * n1, 4 * n1
        # tensor orthonormal basis phi_{a + n1 b}(x, y) = P~_a(x) P~_b(y)
        def phi(pts):
            X, Y = ortho1d(p, pts[:, 0]), ortho1d(p, pts[:, 1])
            return np.einsum('ak,bk->bak', X, Y).reshape(n1 * n1, -1)
        def dphi(pts):
            X, Y = ortho1d(p, pts[:, 0]), ortho1d(p, pts[:, 1])
            dX, dY = dortho1d(p, pts[:, 0]), dortho1d(p, pts[:, 1])
            gx = np.einsum('ak,bk->bak', dX, Y).reshape(n1 * n1, -1)
            gy = np.einsum('ak,bk->bak', X, dY).reshape(n1 * n1, -1)
            return gx, gy
        Vinv = np.linalg.inv(phi(self.upts).T)        # nodal = phi(pts)^T Vinv
        self.M0 = phi(self.fpts).T @ Vinv             # (nfpts, nupts)
        gx, gy = dphi(self.upts)
        self.M1x = gx.T @ Vinv                        # (nupts, nupts) d/dxi
        self.M1y = gy.T @ Vinv
        # lifting basis: S[k, j] = sum_q w_q L_k(q) phi_j(proj(q)), GLL(p+1) face rule
        qx, qw = gll_rule(n1)
        Lq = np.empty((n1, n1))                       # face nodal basis at q: L[q, k]
        for k in range(n1):
            others = [fs[m] for m in range(n1) if m != k]
            Lq[:, k] = np.prod([(qx - o) / (fs[k] - o) for o in others], axis=0) if others else 1.0
        S = []
        for proj, _ in faces:
            pq = np.array([proj(t) for t in qx])
            M = phi(pq)                               # (nbasis, nq)
            S.append(np.einsum('i,ik,ji->kj', qw, Lq, M))
        S = np.vstack(S)                              # (nfpts, nbasis)
        self.M3 = (S @ phi(self.upts)).T             # (nupts, nfpts)
        nx, ny = self.norm_fpts[:, 0], self.norm_fpts[:, 1]
        self.Dx = self.M1x - self.M3 @ (nx[:, None] * self.M0)
        self.Dy = self.M1y - self.M3 @ (ny[:, None] * self.M0)

    def packed(self):
        """Dx, Dy, M0, M3 concatenated (row major) for the device kernel."""
        return np.concatenate([self.Dx.ravel(), self.Dy.ravel(), self.M0.ravel(), self.M3.ravel()])


class PeriodicQuadAdvection:
    """-div(c f) for all velocities on an nelx x nely periodic mesh of
    hx x hy elements (element e = ex + nelx*ey), AoSoA banks of ioshape
    (nupts, Nv^3, nelx*nely)."""

    def __init__(self, backend, vm, nelx, nely, hx, hy, order=2, upts='gauss-legendre',
                 fpts='gauss-legendre'):
        torch = backend.torch
        self.backend = backend
        self.ops = QuadFROperators(order, upts, fpts)
        self.order, self.nelx, self.nely, self.hx, self.hy = order, nelx, nely, hx, hy
        self.d_ops = torch.as_tensor(self.ops.packed(), dtype=torch.float64, device=backend.device)
        cv = vm.cv()
        self.d_cvx = torch.as_tensor(np.ascontiguousarray(cv[0]), dtype=torch.float64, device=backend.device)
        self.d_cvy = torch.as_tensor(np.ascontiguousarray(cv[1]), dtype=torch.float64, device=backend.device)
        self.nvar = vm.vsize()

    def ploc(self):
        """Physical solution-point coordinates (nupts, 2, neles)."""
        ex = np.arange(self.nelx * self.nely) % self.nelx
        ey = np.arange(self.nelx * self.nely) // self.nelx
        u = self.ops.upts
        x = (ex[None, :] + 0.5 * (u[:, 0:1] + 1)) * self.hx
        y = (ey[None, :] + 0.5 * (u[:, 1:2] + 1)) * self.hy
        return np.stack([x, y], axis=1)

    def __call__(self, u, out):
        _lib.check(_lib.load().fq_fr_rhs_quad(
            self.order, self.nelx, self.nely, self.hx, self.hy, self.d_ops.data_ptr(),
            self.d_cvx.data_ptr(), self.d_cvy.data_ptr(), u._as_parameter_, out._as_parameter_,
            u.leaddim, self.nvar, self.backend.stream_handle()))
