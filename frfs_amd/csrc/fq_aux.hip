// Kernels on either side of the collision operator, in the solver's AoSoA
// layout: register arithmetic (axnpby), velocity-range copies between banks
// and per-DOF moments.  Stateless entry points (no handle).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/frfs_fq.h"

int fq_internal_fail(int code, const std::string &msg);   // fq_lib.hip

namespace {

// x0 = a0 x0 + sum_{i>=1} a_i x_i over a whole bank; a0 == 0 never reads x0
// and a0 == 1 accumulates (axnpby_dgfs_full.mako; blasext.py:46-75).
struct AxArgs {
    const double *x[FQ_AXNPBY_MAX];
    double a[FQ_AXNPBY_MAX + 1];
};

__global__ __launch_bounds__(256) void k_axnpby(int64_t n, double *__restrict__ x0, int nv, AxArgs args) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0;
    for (int k = 0; k < nv; ++k) s += args.a[k + 1] * args.x[k][i];
    const double a0 = args.a[0];
    if (a0 == 0.0) x0[i] = s;
    else if (a0 == 1.0) x0[i] += s;
    else x0[i] = a0 * x0[i] + s;
}

// dst[u, v + sidx_dst, e] = src[u, v + sidx_src, e] for v < count
// (copy_dgfs.mako; blasext.py:77-138).  One thread per (u, e, v).
__global__ __launch_bounds__(256) void k_copy_dgfs(int nrow, int neles, int count, const double *__restrict__ src,
                                                   int64_t lds, int nvs, int sidxs, double *__restrict__ dst,
                                                   int64_t ldd, int nvd, int sidxd) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t tot = int64_t(nrow) * neles * count;
    if (t >= tot) return;
    const int e = int(t % neles);
    const int64_t r = t / neles;
    const int v = int(r % count);
    const int u = int(r / count);
    const int64_t is = u * lds + (int64_t(e / 32) * nvs + v + sidxs) * 32 + e % 32;
    const int64_t id = u * ldd + (int64_t(e / 32) * nvd + v + sidxd) * 32 + e % 32;
    dst[id] = src[is];
}

// Non-dimensional moments of one DOF per block, in the order of
// _compute_moments_3D (frfs/plugins/dgfsmomwriterstd.py:142-214):
//   rho, ux, uy, uz, T, qx, qy, qz, Pxx, Pxy, Pxz, Pyy, Pyz, Pzz
// (the caller applies the dimensional factors).  Two passes: mass and
// momentum, then the peculiar-velocity moments.
__device__ __forceinline__ double block_sum(double v, double *red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(256) void k_moments(int N, double L, double cw, double mr, const double *__restrict__ f,
                                                 int nupts, int64_t ldim, int64_t first, double *__restrict__ out,
                                                 int64_t ndof) {
    __shared__ double red[256];
    // workgroups are dealt round-robin to the eight XCDs: each XCD takes a
    // contiguous DOF range, so neighbouring elements (one AoSoA line) share an L2
    const int64_t per = (ndof + 7) / 8;
    const int64_t slot = int64_t(blockIdx.x & 7) * per + int64_t(blockIdx.x >> 3);
    if (slot >= ndof) return;
    const int64_t dof = first + slot;
    const int64_t elem = dof / nupts, upt = dof % nupts;
    const int64_t nvol = int64_t(N) * N * N;
    const int64_t base = upt * ldim + (elem / 32) * nvol * 32 + elem % 32;
    const double h = 2.0 * L / N, c0 = -L + L / N;
    const double mcw = mr * cw;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    for (int64_t v = threadIdx.x; v < nvol; v += blockDim.x) {
        const double fv = f[base + v * 32];
        const double cx = c0 + h * double(v / (int64_t(N) * N)), cy = c0 + h * double((v / N) % N),
                     cz = c0 + h * double(v % N);
        s0 += fv; s1 += fv * cx; s2 += fv * cy; s3 += fv * cz;
    }
    const double rho = block_sum(s0, red) * mcw;
    const double ux = block_sum(s1, red) * mcw / rho;
    const double uy = block_sum(s2, red) * mcw / rho;
    const double uz = block_sum(s3, red) * mcw / rho;
    double m[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int64_t v = threadIdx.x; v < nvol; v += blockDim.x) {
        const double fv = f[base + v * 32];
        const double cx = c0 + h * double(v / (int64_t(N) * N)) - ux, cy = c0 + h * double((v / N) % N) - uy,
                     cz = c0 + h * double(v % N) - uz;
        const double c2 = cx * cx + cy * cy + cz * cz;
        m[0] += fv * c2;
        m[1] += fv * c2 * cx; m[2] += fv * c2 * cy; m[3] += fv * c2 * cz;
        m[4] += fv * cx * cx; m[5] += fv * cx * cy; m[6] += fv * cx * cz;
        m[7] += fv * cy * cy; m[8] += fv * cy * cz; m[9] += fv * cz * cz;
    }
    double r[10];
    for (int k = 0; k < 10; ++k) r[k] = block_sum(m[k], red);
    if (threadIdx.x == 0) {
        double *o = out + slot * FQ_NMOMENTS;
        o[0] = rho; o[1] = ux; o[2] = uy; o[3] = uz;
        o[4] = r[0] * (2.0 / 3.0 * mcw * mr) / rho;
        o[5] = mr * r[1] * mcw; o[6] = mr * r[2] * mcw; o[7] = mr * r[3] * mcw;
        for (int k = 0; k < 6; ++k) o[8 + k] = 2 * mr * r[4 + k] * mcw;
    }
}

// FR advection right-hand side -div(c f) on a periodic nelx x nely mesh of
// affine hx x hy quads, order P, one thread per (velocity v, element e):
//   rhs = -rcpdjac [ (hy/2) cx Dx u + (hx/2) cy Dy u + M3 F* ],
//   F*_f = |n_f| c_n u_upwind   (upwind common normal flux, rusanov.mako)
// ops = {Dx[NU][NU], Dy[NU][NU], M0[NF][NU], M3[NU][NF]} (row major).
template <int P>
__global__ __launch_bounds__(256) void k_fr_rhs_quad(int nelx, int nely, double hx, double hy,
                                                     const double *__restrict__ ops, const double *__restrict__ cvx,
                                                     const double *__restrict__ cvy, const double *__restrict__ u,
                                                     double *__restrict__ out, int64_t ldim, int nvar) {
    constexpr int N1 = P + 1, NU = N1 * N1, NF = 4 * N1;
    const int neles = nelx * nely;
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= int64_t(nvar) * neles) return;
    const int e = int(t % neles);
    const int v = int(t / neles);
    const double *Dx = ops, *Dy = ops + NU * NU, *M0 = ops + 2 * NU * NU, *M3 = M0 + NF * NU;
    auto at = [&](int upt, int ee) { return upt * ldim + (int64_t(ee / 32) * nvar + v) * 32 + ee % 32; };
    const int ex = e % nelx, ey = e / nelx;
    const int nb[4] = {ex + nelx * ((ey + nely - 1) % nely), (ex + 1) % nelx + nelx * ey,
                       ex + nelx * ((ey + 1) % nely), (ex + nelx - 1) % nelx + nelx * ey};
    const double nx[4] = {0, 1, 0, -1}, ny[4] = {-1, 0, 1, 0};
    const double cx = cvx[v], cy = cvy[v];
    double ue[NU];
#pragma unroll
    for (int k = 0; k < NU; ++k) ue[k] = u[at(k, e)];
    double fstar[NF];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        double un[NU];
#pragma unroll
        for (int k = 0; k < NU; ++k) un[k] = u[at(k, nb[f])];
        const int fo = (f + 2) % 4;
        const double cn = cx * nx[f] + cy * ny[f];
        const double magn = (f % 2 == 0) ? 0.5 * hx : 0.5 * hy;
#pragma unroll
        for (int k = 0; k < N1; ++k) {
            double uo = 0, ui = 0;
#pragma unroll
            for (int m = 0; m < NU; ++m) {
                uo += M0[(f * N1 + k) * NU + m] * ue[m];
                ui += M0[(fo * N1 + k) * NU + m] * un[m];
            }
            fstar[f * N1 + k] = magn * cn * (cn > 0 ? uo : ui);
        }
    }
    const double ax = 0.5 * hy * cx, ay = 0.5 * hx * cy, rcp = -4.0 / (hx * hy);
#pragma unroll
    for (int i = 0; i < NU; ++i) {
        double sx = 0, sy = 0, sc = 0;
#pragma unroll
        for (int k = 0; k < NU; ++k) {
            sx += Dx[i * NU + k] * ue[k];
            sy += Dy[i * NU + k] * ue[k];
        }
#pragma unroll
        for (int j = 0; j < NF; ++j) sc += M3[i * NF + j] * fstar[j];
        out[at(i, e)] = rcp * (ax * sx + ay * sy + sc);
    }
}

}  // namespace

extern "C" {

int fq_fr_rhs_quad(int32_t order, int32_t nelx, int32_t nely, double hx, double hy, const double *ops,
                   const double *cvx, const double *cvy, const double *u, double *out, int64_t ldim, int32_t nvar,
                   void *stream) {
    if (nelx <= 0 || nely <= 0 || !(hx > 0) || !(hy > 0) || !ops || !cvx || !cvy || !u || !out || nvar <= 0)
        return fq_internal_fail(FQ_ERR_INVALID, "fq_fr_rhs_quad: invalid arguments");
    const int64_t nparr = (int64_t(nelx) * nely + 31) / 32 * 32;
    if (ldim < int64_t(nvar) * nparr) return fq_internal_fail(FQ_ERR_INVALID, "fq_fr_rhs_quad: ldim too small");
    if (u == out) return fq_internal_fail(FQ_ERR_INVALID, "fq_fr_rhs_quad: in-place evaluation is not supported");
    const int64_t tot = int64_t(nvar) * nelx * nely;
    const dim3 grid(unsigned((tot + 255) / 256)), blk(256);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    switch (order) {
    case 1: hipLaunchKernelGGL(k_fr_rhs_quad<1>, grid, blk, 0, st, nelx, nely, hx, hy, ops, cvx, cvy, u, out, ldim, nvar); break;
    case 2: hipLaunchKernelGGL(k_fr_rhs_quad<2>, grid, blk, 0, st, nelx, nely, hx, hy, ops, cvx, cvy, u, out, ldim, nvar); break;
    case 3: hipLaunchKernelGGL(k_fr_rhs_quad<3>, grid, blk, 0, st, nelx, nely, hx, hy, ops, cvx, cvy, u, out, ldim, nvar); break;
    default: return fq_internal_fail(FQ_ERR_UNSUPPORTED, "fq_fr_rhs_quad: order must be 1, 2 or 3");
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FQ_SUCCESS : fq_internal_fail(FQ_ERR_HIP, std::string("fq_fr_rhs_quad: ") + hipGetErrorString(e));
}

int fq_axnpby(int64_t n, double *x0, int32_t nv, const double *const *xs, const double *a, void *stream) {
    if (n < 0 || !x0 || !a || nv < 0 || nv > FQ_AXNPBY_MAX || (nv > 0 && !xs))
        return fq_internal_fail(FQ_ERR_INVALID, "fq_axnpby: invalid arguments");
    if (n == 0) return FQ_SUCCESS;
    AxArgs args{};
    for (int k = 0; k < nv; ++k) {
        if (!xs[k]) return fq_internal_fail(FQ_ERR_INVALID, "fq_axnpby: null operand");
        args.x[k] = xs[k];
    }
    for (int k = 0; k <= nv; ++k) args.a[k] = a[k];
    hipLaunchKernelGGL(k_axnpby, dim3(unsigned((n + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), n, x0, nv, args);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FQ_SUCCESS : fq_internal_fail(FQ_ERR_HIP, std::string("fq_axnpby: ") + hipGetErrorString(e));
}

int fq_copy_dgfs(int32_t nrow, int32_t neles, int32_t count, const double *src, int64_t ldim_src, int32_t nvar_src,
                 int32_t sidx_src, double *dst, int64_t ldim_dst, int32_t nvar_dst, int32_t sidx_dst, void *stream) {
    if (nrow <= 0 || neles <= 0 || count < 0 || !src || !dst || sidx_src < 0 || sidx_dst < 0 ||
        sidx_src + count > nvar_src || sidx_dst + count > nvar_dst)
        return fq_internal_fail(FQ_ERR_INVALID, "fq_copy_dgfs: invalid arguments");
    const int64_t nparr = (int64_t(neles) + 31) / 32 * 32;
    if (ldim_src < nvar_src * nparr || ldim_dst < nvar_dst * nparr)
        return fq_internal_fail(FQ_ERR_INVALID, "fq_copy_dgfs: leading dimension too small");
    const int64_t tot = int64_t(nrow) * neles * count;
    if (tot == 0) return FQ_SUCCESS;
    hipLaunchKernelGGL(k_copy_dgfs, dim3(unsigned((tot + 255) / 256)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), nrow, neles, count, src, ldim_src, nvar_src, sidx_src,
                       dst, ldim_dst, nvar_dst, sidx_dst);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FQ_SUCCESS : fq_internal_fail(FQ_ERR_HIP, std::string("fq_copy_dgfs: ") + hipGetErrorString(e));
}

int fq_moments(int32_t Nv, double L, double cw, double mass_ratio, const double *f, int32_t nupts, int64_t ldim,
               int32_t neles, int64_t first_dof, int64_t ndof, double *out, void *stream) {
    if (Nv <= 0 || !(L > 0) || !f || !out || nupts <= 0 || neles <= 0 || first_dof < 0 || ndof < 0 ||
        first_dof + ndof > int64_t(nupts) * neles)
        return fq_internal_fail(FQ_ERR_INVALID, "fq_moments: invalid arguments");
    const int64_t nparr = (int64_t(neles) + 31) / 32 * 32;
    if (ldim < int64_t(Nv) * Nv * Nv * nparr) return fq_internal_fail(FQ_ERR_INVALID, "fq_moments: ldim too small");
    if (ndof == 0) return FQ_SUCCESS;
    hipLaunchKernelGGL(k_moments, dim3(unsigned(8 * ((ndof + 7) / 8))), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), Nv, L, cw, mass_ratio, f, nupts, ldim, first_dof, out,
                       ndof);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? FQ_SUCCESS : fq_internal_fail(FQ_ERR_HIP, std::string("fq_moments: ") + hipGetErrorString(e));
}

}  // extern "C"
