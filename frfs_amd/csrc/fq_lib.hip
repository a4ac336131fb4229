// MI355X (gfx950) fast spectral collision operator: kernels and C ABI.
//
// Pipeline per batch ("chunk") of D DOFs, all volumes stored plane-major
// [plane][I][J] (J fastest, plane = the z index K or vz):
//   k_fwd_z    f (AoSoA, per source)  -> T0  : DFT- along z           (r2z + cuFFT fwd)
//   k_fwd_xy   T0 -> F (FP64)                : DFT- along x, y, / n (+ FP32 copy, mixed;
//                                              + the slice J = N/2 for k_nyq)
//   k_nyq      F -> NY                       : (pair packing) per product, the 2D
//                                              transforms of the Nyquist planes
//   k_prod / k_prod1  F, NY -> P_g (T)       : per group: the items' pruned z
//                                              DFTs straight into LDS, x,y DFT+,
//                                              products, group sum, DFT- x,y; with
//                                              pair packing both items of a product
//                                              as one complex volume
//     (staged alternative, FQ_FLAG_UNFUSED: k_inv_z F -> Z_i through HBM,
//      then k_plane Z -> P_g)
//   k_accum / k_accum_lds  P, F -> Y (FP64)  : DFT- along z, H += W . Ghat (FP64),
//                                              L += Lt . F, DFT+ along z of H and L
//   k_out      Y -> Q (AoSoA, per target)    : DFT+ x,y; Q = Re gain - f Re loss
//   k_project  optional conservation projection (FP64)
// T = double (FQ_PREC_FP64) or float (FQ_PREC_MIXED).
//
// Lane mapping: a "unit" (one plane or one row of N pencils) is LPP = 2*NP
// lanes, NP = N rounded up to a power of two; lane = 2*J + half, the two
// lanes of a pencil exchange through DPP (see fft.hpp).  k_prod1 instead
// puts one pencil on each lane (FP32, N = 24, 32).  Global loads and stores
// walk J across lanes, so every access is a contiguous row segment.
#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/frfs_fq.h"
#include "fft.hpp"
#include "plan.hpp"

// ---------------------------------------------------------------------------
// Static scratch of fq_eval_noalloc (declared in frfs_fq.h): a file-scope
// device array, one copy per GPU, held until the process ends.  Host code
// never uses its name as an address (on the host it is only a stand-in); the
// device address comes from hipGetSymbolAddress, resolved once per device
// with that device current (at fq_create) and kept per device.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[FQ_NOALLOC_SCRATCH_BYTES];

using fq::cpx;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(call)                                                                 \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess)                                                        \
            return fail(FQ_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int pow2ceil(int n) { return n <= 8 ? 8 : n <= 16 ? 16 : n <= 32 ? 32 : 64; }

template <int N> struct Geo {
    static constexpr int E = N / 2;
    static constexpr int NP = pow2ceil(N);
    static constexpr int LPP = 2 * NP;        // lanes per unit
    static constexpr int N2 = N * N;
    static constexpr int VOL = N * N * N;
    static constexpr int LDSROW = N + 1;      // padded LDS row (complex elements)
};

// Workgroups are dealt round-robin to the eight XCDs.  A kernel launched with
// a grid that is a multiple of 8 runs block w as block (w % 8) * (grid / 8) +
// w / 8, so each XCD takes a contiguous block range and neighbouring DOFs,
// whose AoSoA values share cache lines, meet in one L2.
__device__ __forceinline__ int64_t xcd_block() {
    return int64_t(blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
}

constexpr int kPencilThreads = 256;           // pencil kernels (no LDS)
constexpr int kPlaneThreads = 128;            // plane kernels (LDS transpose per unit)

// Buffer view of a workgroup-uniform array of `bytes` bytes.  Loads take a
// per-lane 32-bit byte offset plus a uniform byte offset held in an SGPR, so
// a run of loads at constant strides shares one address register.  An offset
// past the end reads zero (buffer range check) instead of faulting.
struct BufView {
    __amdgpu_buffer_rsrc_t r;
    __device__ __forceinline__ BufView(const void *base, uint32_t bytes)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, int(bytes), 0x00020000)) {}
    template <typename T> __device__ __forceinline__ cpx<T> ld(uint32_t voff, uint32_t soff) const;
    // the same load, non-temporal (cache policy bit nt), for data read once
    template <typename T> __device__ __forceinline__ cpx<T> ld_nt(uint32_t voff, uint32_t soff) const;
};
template <> __device__ __forceinline__ cpx<float> BufView::ld<float>(uint32_t voff, uint32_t soff) const {
    return __builtin_bit_cast(cpx<float>, __builtin_amdgcn_raw_buffer_load_b64(r, int(voff), int(soff), 0));
}
// Two consecutive FP32 complex values in one 16-byte load.
struct cpx2f { cpx<float> a, b; };
__device__ __forceinline__ cpx2f ld2f(const BufView &v, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(cpx2f, __builtin_amdgcn_raw_buffer_load_b128(v.r, int(voff), int(soff), 0));
}
template <> __device__ __forceinline__ cpx<double> BufView::ld<double>(uint32_t voff, uint32_t soff) const {
    return __builtin_bit_cast(cpx<double>, __builtin_amdgcn_raw_buffer_load_b128(r, int(voff), int(soff), 0));
}
template <> __device__ __forceinline__ cpx<double> BufView::ld_nt<double>(uint32_t voff, uint32_t soff) const {
    return __builtin_bit_cast(cpx<double>, __builtin_amdgcn_raw_buffer_load_b128(r, int(voff), int(soff), 2));
}

// clang vector types for __builtin_nontemporal_load / _store
typedef float nt_f2 __attribute__((ext_vector_type(2)));
typedef float nt_f4 __attribute__((ext_vector_type(4)));
typedef double nt_d2 __attribute__((ext_vector_type(2)));

template <typename To, typename Ti> __device__ __forceinline__ cpx<To> cvt(cpx<Ti> v) {
    return {To(v.x), To(v.y)};
}

struct AosoaPtrs {
    const double *f[4];
    double *q[4];
};

struct Dof {
    int64_t first;     // global DOF of chunk slot 0
    int32_t nupts;
    int32_t soasz;
    int64_t ldim;
    int64_t nvol;      // N^3
    __device__ __forceinline__ int64_t base(int slot) const {
        int64_t dof = first + slot;
        int64_t elem = dof / nupts, upt = dof % nupts;
        return upt * ldim + (elem / soasz) * nvol * soasz + elem % soasz;
    }
};

// ---------------------------------------------------------------------------
// Weight-table precompute (FP64), replaces precompute_bb (vhs-gll.mako:71-101)
// and precompute_bc_{pq} (vss-gll.mako:53-95).  Table layout is plane-major
// [K][I][J] with l = (l0[I], l0[J], l0[K]).
// ---------------------------------------------------------------------------
struct DevTerm {
    int kind, x, y, pair;
    double A, B, scale;
};

__device__ __forceinline__ double l0of(int idx, int N) { return double(idx < N / 2 ? idx : idx - N); }

__device__ __forceinline__ double sinc_ref(double s) { return sin(s + 1e-15) / (s + 1e-15); }

__global__ void k_tables(cpx<double> *tabs, const DevTerm *terms, const int *tbeg, int N,
                         double L, double sw, double swpre, int M, int Mpre, int Nrho, int ns,
                         const double *qz, const double *qw, const double *sz,
                         const double *szpre, const double *masses, const double *eta,
                         const double *gam) {
    const int n = N * N * N;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int tab = blockIdx.y;
    if (idx >= n) return;
    const int K = idx / (N * N), I = (idx / N) % N, J = idx % N;
    const double lx = l0of(I, N), ly = l0of(J, N), lz = l0of(K, N);
    const double cSqr = sqrt(lx * lx + ly * ly + lz * lz);
    double acc_re = 0, acc_im = 0;
    for (int t = tbeg[tab]; t < tbeg[tab + 1]; ++t) {
        const DevTerm tm = terms[t];
        if (tm.kind == 0) {
            acc_re += tm.scale * (tm.A * sinc_ref(tm.B * cSqr));
        } else {
            const int p = tm.pair / ns, q = tm.pair % ns;
            const double mpq = masses[q] / (masses[p] + masses[q]);
            const double e = eta[tm.pair], g = gam[tm.pair];
            const int x0 = tm.kind == 1 ? tm.x : 0, x1 = tm.kind == 1 ? tm.x + 1 : Nrho;
            double c_re = 0, c_im = 0;
            for (int x = x0; x < x1; ++x) {
                const double fac = M_PI / L * qz[x];
                const double fac_b = swpre * pow(qz[x], g + 2);
                const double fac_c = qw[x] * sw * fac_b;
                const int y0 = tm.kind == 1 ? tm.y : 0, y1 = tm.kind == 1 ? tm.y + 1 : M;
                for (int y = y0; y < y1; ++y) {
                    double s_re = 0, s_im = 0;
                    for (int k = 0; k < Mpre; ++k) {
                        const double c0 = szpre[3 * k], c1 = szpre[3 * k + 1], c2 = szpre[3 * k + 2];
                        // the host verified 1 + sigma.sigma' >= 0 in exact order; clamp
                        // contraction-rounded zeros (exact antipodes) to 0
                        const double base = 1 + sz[3 * y] * c0 + sz[3 * y + 1] * c1 + sz[3 * y + 2] * c2;
                        const double Bpq = pow(fmax(base, 0.0), e);
                        const double Ld = fac * ((c0 * lx) + (c1 * ly) + (c2 * lz));
                        if (tm.kind == 1) {
                            s_re += Bpq * fac_b * cos(mpq * Ld);
                            s_im += -Bpq * fac_b * sin(mpq * Ld);
                        } else {
                            s_re += Bpq * fac_c * cos(Ld);
                            s_im += -Bpq * fac_c * sin(Ld);
                        }
                    }
                    c_re += s_re;
                    c_im += s_im;
                }
            }
            acc_re += tm.scale * c_re;
            acc_im += tm.scale * c_im;
        }
    }
    tabs[int64_t(tab) * n + idx] = {acc_re, acc_im};
}

// ---------------------------------------------------------------------------
// k_fwd_z: AoSoA gather + DFT- along z.  unit = (slot, src, I).
// ---------------------------------------------------------------------------
template <int N>
__global__ __launch_bounds__(256) void k_fwd_z(AosoaPtrs ptr, Dof dof, cpx<double> *T0, int nsrc,
                                               int64_t nunits) {
    using G = Geo<N>;
    constexpr int E = G::E;
    const int lt = threadIdx.x % G::LPP;
    const int64_t unit = int64_t(blockIdx.x) * (blockDim.x / G::LPP) + threadIdx.x / G::LPP;
    if (unit >= nunits) return;   // whole unit exits together (no barriers here)
    const int I = int(unit % N);
    const int64_t ds = unit / N;
    const int s = int(ds % nsrc);
    const int slot = int(ds / nsrc);
    const int J = lt >> 1, half = lt & 1;
    const bool act = J < N;
    const double *f = ptr.f[s];
    const int64_t b = dof.base(slot);
    cpx<double> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int K = half * E + e;
        const int64_t vv = (int64_t(I) * N + J) * N + K;
        v[e] = {act ? f[b + vv * dof.soasz] : 0.0, 0.0};
    }
    fq::fft2_dif<N, -1>(v, half);
    cpx<double> *dst = T0 + (int64_t(slot) * nsrc + s) * G::VOL + int64_t(I) * N + J;
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) dst[int64_t(half + 2 * e) * G::N2] = v[e];
    }
}

// k_fwd_z4 (N = 32): the same for four consecutive batch slots per workgroup
// (one wave each, same source and row I).  The workgroup first reads the
// row's 1024 values of the four slots with the slots on adjacent lanes, so
// consecutive elements of one solution point meet in one AoSoA row segment,
// and stages them in LDS as [J][half][e][slot] (half-rows of 65 and rows of
// 130 doubles: lanes J and J + 16 share a bank pair, no other conflicts).
template <int N>
__global__ __launch_bounds__(256) void k_fwd_z4(AosoaPtrs ptr, Dof dof, cpx<double> *T0, int nsrc, int D) {
    using G = Geo<N>;
    static_assert(G::LPP == 64, "one row per wave");
    constexpr int E = G::E, U = 4, HS = U * E + 1, JS = 2 * HS;
    __shared__ double st[N * JS];
    const int tid = int(threadIdx.x);
    const int b = int(blockIdx.x);
    const int I = b % N, s = (b / N) % nsrc, sg = b / (N * nsrc);
    const double *f = ptr.f[s];
    {
        const int u2 = tid % U, slot2 = sg * U + u2;
        const int64_t b2 = slot2 < D ? dof.base(slot2) : 0;
#pragma unroll 4
        for (int i = 0; i < G::N2 / (256 / U); ++i) {
            const int pnt = tid / U + (256 / U) * i;            // (J, K) of row I
            const int J = pnt / N, K = pnt % N;
            const double x = slot2 < D ? f[b2 + ((int64_t(I) * N + J) * N + K) * dof.soasz] : 0.0;
            st[J * JS + (K / E) * HS + (K % E) * U + u2] = x;
        }
    }
    __syncthreads();
    const int u = tid / 64, lt = tid % 64;
    const int slot = sg * U + u;
    if (slot >= D) return;                           // no barrier follows
    const int J = lt >> 1, half = lt & 1;
    cpx<double> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = {st[J * JS + half * HS + e * U + u], 0.0};
    fq::fft2_dif<N, -1>(v, half);
    cpx<double> *dst = T0 + (int64_t(slot) * nsrc + s) * G::VOL + int64_t(I) * N + J;
#pragma unroll
    for (int e = 0; e < E; ++e)        // non-temporal: read once, by k_fwd_xy
        __builtin_nontemporal_store(nt_d2{v[e].x, v[e].y}, reinterpret_cast<nt_d2 *>(dst + int64_t(half + 2 * e) * G::N2));
}

// ---------------------------------------------------------------------------
// Shared plane helpers.  A plane unit has LPP lanes; `row` = lt >> 1 is the
// column J in the column phase and the row (x index) in the row phase.
// ---------------------------------------------------------------------------

// k_fwd_xy: plane (slot, src, K'):  DFT- along x then y, divide by n; with
// F32 non-null it also writes the FP32 copy the mixed-precision path reads.
// FJ64 / FJ32 (pair packing) receive the plane J = N/2 of the spectrum as
// [slot][src][I][K], so k_nyq reads that strided slice contiguously.
template <int N>
__global__ __launch_bounds__(kPlaneThreads) void k_fwd_xy(const cpx<double> *T0, cpx<double> *F,
                                                           cpx<float> *F32, cpx<double> *FJ64, cpx<float> *FJ32,
                                                           int64_t nunits) {
    using G = Geo<N>;
    constexpr int E = G::E;
    constexpr int UPB = kPlaneThreads / G::LPP;
    __shared__ cpx<double> lds[UPB][N * G::LDSROW];
    const int u = threadIdx.x / G::LPP;
    const int lt = threadIdx.x % G::LPP;
    const int64_t unit = int64_t(blockIdx.x) * UPB + u;
    const bool valid = unit < nunits;
    const int c = lt >> 1, half = lt & 1;
    const bool act = valid && c < N;
    const cpx<double> *src = T0 + unit * G::N2;
    cpx<double> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {                // T0 is read once: non-temporal
        cpx<double> t = {0, 0};
        if (act) {
            const nt_d2 x = __builtin_nontemporal_load(reinterpret_cast<const nt_d2 *>(src + (half * E + e) * N + c));
            t = {x.x, x.y};
        }
        v[e] = t;
    }
    fq::fft2_dif<N, -1>(v, half);                       // over I -> I' cyclic
    cpx<double> *L = lds[u];
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) L[(half + 2 * e) * G::LDSROW + c] = v[e];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = act ? L[c * G::LDSROW + half * E + e] : cpx<double>{0, 0};
    fq::fft2_dif<N, -1>(v, half);                       // over J -> J' cyclic, row I' = c
    const double nn = double(G::VOL);
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = {v[e].x / nn, v[e].y / nn};   // FTf /= vsize (vhs-gll.mako:120)
    __syncthreads();
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) L[c * G::LDSROW + half + 2 * e] = v[e];
    }
    __syncthreads();
    cpx<double> *dst = F + unit * G::N2;
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const cpx<double> x = L[(half * E + e) * G::LDSROW + c];
            dst[(half * E + e) * N + c] = x;
            if (F32) F32[unit * G::N2 + (half * E + e) * N + c] = cvt<float>(x);   // mixed precision copy
            if (c == N / 2) {
                const int64_t fj = (unit / N) * G::N2 + int64_t(half * E + e) * N + unit % N;
                if (FJ64) FJ64[fj] = x;
                if (FJ32) FJ32[fj] = cvt<float>(x);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_inv_z: unit = (slot, item, I).  Z_i = DFT+_z( e^{i th_z l_z} F_src ).
// ---------------------------------------------------------------------------
template <int N, typename T>
__global__ __launch_bounds__(256) void k_inv_z(const cpx<double> *F, cpx<T> *Z, const cpx<T> *ph,
                                               const int *item_src, int nitems, int nsrc,
                                               int64_t nunits) {
    using G = Geo<N>;
    constexpr int E = G::E;
    const int lt = threadIdx.x % G::LPP;
    const int64_t unit = int64_t(blockIdx.x) * (blockDim.x / G::LPP) + threadIdx.x / G::LPP;
    if (unit >= nunits) return;
    const int I = int(unit % N);
    const int64_t di = unit / N;
    const int it = int(di % nitems);
    const int slot = int(di / nitems);
    const int J = lt >> 1, half = lt & 1;
    const bool act = J < N;
    const int s = item_src[it];
    const cpx<double> *src = F + (int64_t(slot) * nsrc + s) * G::VOL + I * N + J;
    const cpx<T> *phz = ph + (int64_t(it) * 3 + 2) * N;
    cpx<T> v[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int K = half * E + e;
        cpx<T> x = act ? cvt<T>(src[int64_t(K) * G::N2]) : cpx<T>{0, 0};
        v[e] = fq::cmul(x, phz[K]);
    }
    fq::fft2_dif<N, 1>(v, half);
    cpx<T> *dst = Z + (int64_t(slot) * nitems + it) * G::VOL + I * N + J;
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) dst[int64_t(half + 2 * e) * G::N2] = v[e];
    }
}

// ---------------------------------------------------------------------------
// k_plane: unit = (slot, group, vz).  For each product of the group:
//   E_i = DFT+_{x,y}( e^{i(th_x l_x + th_y l_y)} Z_i(:,:,vz) ),  acc += coef E1 E2
// then P_g(:,:,vz) = DFT-_{x,y}(acc).
// ---------------------------------------------------------------------------
struct DevProd { int i1, i2; double coef; };

template <int N, typename T>
__device__ __forceinline__ void plane_inv(const cpx<T> *src, const cpx<T> *ph, cpx<T> *L, cpx<T> *v,
                                          int c, int half, bool act) {
    using G = Geo<N>;
    constexpr int E = G::E;
    // column phase: lane (J=c, half) over I (block)
    const cpx<T> py = act ? ph[N + c] : cpx<T>{1, 0};
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int I = half * E + e;
        cpx<T> x = act ? src[I * N + c] : cpx<T>{0, 0};
        v[e] = fq::cmul(x, fq::cmul(ph[I], py));
    }
    fq::fft2_dif<N, 1>(v, half);                 // -> vx = half + 2e
    __syncthreads();
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) L[(half + 2 * e) * G::LDSROW + c] = v[e];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = act ? L[c * G::LDSROW + half * E + e] : cpx<T>{0, 0};
    fq::fft2_dif<N, 1>(v, half);                 // row vx = c: -> vy = half + 2e
}

template <int N, typename T, bool MULTI>
__global__ __launch_bounds__(kPlaneThreads) void k_plane(const cpx<T> *Z, cpx<T> *P, const cpx<T> *ph,
                                               const DevProd *prods, const int *g_p0, const int *g_np,
                                               int nitems, int ngroups, int64_t nunits) {
    using G = Geo<N>;
    constexpr int E = G::E;
    constexpr int UPB = kPlaneThreads / G::LPP;
    __shared__ cpx<T> lds[UPB][N * G::LDSROW];
    const int u = threadIdx.x / G::LPP;
    const int lt = threadIdx.x % G::LPP;
    // the launch pads the grid to a multiple of 8, so xcd_block() is a
    // bijection of the grid; padded units fail `valid`
    const int64_t unit = xcd_block() * UPB + u;
    const bool valid = unit < nunits;
    const int vz = int(unit % N);
    const int64_t dg = unit / N;
    const int g = valid ? int(dg % ngroups) : 0;
    const int slot = int(dg / ngroups);
    const int c = lt >> 1, half = lt & 1;
    const bool act = valid && c < N;
    cpx<T> *L = lds[u];
    const int p0 = g_p0[g], np = g_np[g];
    const cpx<T> *zs = Z + int64_t(slot) * nitems * G::VOL + int64_t(vz) * G::N2;
    cpx<T> acc[E];
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = {0, 0};
    // plane_inv contains barriers: every lane of the block must run the same
    // number of product iterations, so loop to the block-wide maximum and
    // mask the iterations a unit does not own.
    __shared__ int s_np[UPB];
    if (lt == 0) s_np[u] = valid ? np : 0;
    __syncthreads();
    int npmax = 0;
    for (int k = 0; k < UPB; ++k) npmax = max(npmax, s_np[k]);
    for (int k = 0; k < npmax; ++k) {
        const bool on = act && k < np;
        DevProd pr = prods[p0 + (k < np ? k : 0)];
        cpx<T> e1[E], e2[E];
        plane_inv<N, T>(zs + int64_t(pr.i1) * G::VOL, ph + int64_t(pr.i1) * 3 * N, L, e1, c, half, on);
        plane_inv<N, T>(zs + int64_t(pr.i2) * G::VOL, ph + int64_t(pr.i2) * 3 * N, L, e2, c, half, on);
        const T cf = on ? T(pr.coef) : T(0);
        if (MULTI) {
#pragma unroll
            for (int e = 0; e < E; ++e) {
                cpx<T> pe = fq::cmul(e1[e], e2[e]);
                acc[e] = {acc[e].x + cf * pe.x, acc[e].y + cf * pe.y};
            }
        } else {
#pragma unroll
            for (int e = 0; e < E; ++e) acc[e] = fq::cscale(fq::cmul(e1[e], e2[e]), cf);
        }
    }
    // forward: acc is cyclic over vy (row vx = c)
    fq::fft2_dit<N, -1>(acc, half);               // -> J' = half*E + e
    __syncthreads();
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) L[c * G::LDSROW + half * E + e] = acc[e];
    }
    __syncthreads();
    // column phase: lane (J' = c, half), input over vx cyclic
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = act ? L[(half + 2 * e) * G::LDSROW + c] : cpx<T>{0, 0};
    fq::fft2_dit<N, -1>(acc, half);               // -> I' = half*E + e
    cpx<T> *dst = P + (int64_t(slot) * ngroups + g) * G::VOL + int64_t(vz) * G::N2;
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) dst[(half * E + e) * N + c] = acc[e];
    }
}

// ---------------------------------------------------------------------------
// k_prod: fused z-stage + plane stage.  One workgroup per (slot, group, G),
// G in [0,4): it owns the S = N/4 output planes vz = G + 4k.  For every
// product of the group the pruned z-DFT
//   Z_i(I,J,G+4k) = sum_a w_S^{ak} w_N^{aG} sum_b w_4^{bG} (F . phi_z)(a + S b)
// (w = e^{+2 pi i/.}) is computed from the L2-resident source volume
// straight into LDS, then wave k runs the x,y transforms of plane k, the
// product and the group sum; the group's P planes are written once.  This
// removes the round trip of every Z_i volume through HBM.  The two items of
// a product pass through the same LDS planes one after the other.
// ---------------------------------------------------------------------------
template <int N> struct ProdGeo {
    static constexpr int S = N / 4;
    static constexpr int E = N / 2;
    static constexpr int LPP = Geo<N>::LPP;
    static constexpr int THREADS = S * LPP;
    static constexpr int ROW = N + 1;
    static constexpr int PLANE = N * ROW;
};

// Pruned z-DFT of one item for the planes vz = G + 4k.  With K = a + S b
// (a < S, b < 4, N = 4S) the phase factorises, e^{i th l0(K)} = alpha_a beta_b
// (alpha_a = e^{i th a}, beta_b = e^{i th S (b - 4[b >= 2])}), so
//   Z(G + 4k) = FFT_S[ c_a sum_b d_b F(a + S b) ](k),
//   d_b = w_4^{bG} beta_b,  c_a = w_N^{aG} alpha_a   (uniform per item and G),
// and the x,y phases phi_x(I) phi_y(J) scale the S outputs of pencil (I,J).
// zc[(item*4 + G)*(4 + S) + .] holds {d_0..d_3, c_0..c_{S-1}}.
template <int N, typename T, int THREADS = ProdGeo<N>::THREADS>
__device__ __forceinline__ void zstage(const cpx<T> *__restrict__ src, const cpx<T> *__restrict__ zc,
                                       const cpx<T> *__restrict__ ph, cpx<T> *dst) {
    using PG = ProdGeo<N>;
    constexpr int S = PG::S;
    constexpr int ITER = (N * N + THREADS - 1) / THREADS;
    // d_0 = c_0 = 1 exactly (zero angle), so those products are skipped
    cpx<T> d[4], cc[S];
#pragma unroll
    for (int b = 1; b < 4; ++b) d[b] = zc[b];
#pragma unroll
    for (int a = 1; a < S; ++a) cc[a] = zc[4 + a];
    constexpr uint32_t PB = uint32_t(N * N * sizeof(cpx<T>));     // plane bytes
    const BufView sv(src, uint32_t(N) * PB);
#pragma unroll 1
    for (int it = 0; it < ITER; ++it) {
        const uint32_t p = threadIdx.x + uint32_t(it * THREADS);
        if (ITER * THREADS != N * N && p >= uint32_t(N * N)) break;
        const uint32_t I = p / N, J = p % N;
        const uint32_t vo = p * uint32_t(sizeof(cpx<T>));
        // the 4S loads go out in two batches of 2S, each issued before its first use
        cpx<T> t[S];
#pragma unroll
        for (int h2 = 0; h2 < 2; ++h2) {
            cpx<T> x[2][S];
#pragma unroll
            for (int bb = 0; bb < 2; ++bb) {
#pragma unroll
                for (int a = 0; a < S; ++a) x[bb][a] = sv.ld<T>(vo, uint32_t(a + S * (2 * h2 + bb)) * PB);
            }
#pragma unroll
            for (int a = 0; a < S; ++a) {
                if (h2 == 0) t[a] = x[0][a];
                else fq::cmac(t[a], x[0][a], d[2]);
                fq::cmac(t[a], x[1][a], d[2 * h2 + 1]);
            }
        }
#pragma unroll
        for (int a = 1; a < S; ++a) t[a] = fq::cmul(t[a], cc[a]);
        fq::LFFT<S, 1>::run(t);
        const cpx<T> pxy = fq::cmul(ph[I], ph[N + J]);
        cpx<T> *dd = dst + I * PG::ROW + J;
#pragma unroll
        for (int k = 0; k < S; ++k) dd[k * PG::PLANE] = fq::cmul(t[k], pxy);
    }
}

// Pruned z-DFT of the packed pair T_1 + i T_2 of one product (the
// pair-packed k_prod1, general pairs): both items' combinations as in
// zstage, with their x,y phases applied before one shared S-point transform.
// Each lane takes two neighbouring pencils (J, J + 1) per step with 16-byte
// loads, one source after the other.
template <int N, typename T, int THREADS>
__device__ __forceinline__ void zstage_pair(const cpx<T> *__restrict__ src1, const cpx<T> *__restrict__ src2,
                                            const cpx<T> *__restrict__ zc1, const cpx<T> *__restrict__ zc2,
                                            const cpx<T> *__restrict__ ph1, const cpx<T> *__restrict__ ph2,
                                            cpx<T> *dst) {
    using PG = ProdGeo<N>;
    constexpr int S = PG::S;
    constexpr int ITER = (N * N / 2 + THREADS - 1) / THREADS;
    cpx<T> d1[4], cc1[S], d2[4], cc2[S];
#pragma unroll
    for (int b = 1; b < 4; ++b) { d1[b] = zc1[b]; d2[b] = zc2[b]; }
#pragma unroll
    for (int a = 1; a < S; ++a) { cc1[a] = zc1[4 + a]; cc2[a] = zc2[4 + a]; }
    constexpr uint32_t PB = uint32_t(N * N * sizeof(cpx<T>));
    const BufView sv1(src1, uint32_t(N) * PB), sv2(src2, uint32_t(N) * PB);
#pragma unroll 1
    for (int it = 0; it < ITER; ++it) {
        const uint32_t p2 = threadIdx.x + uint32_t(it * THREADS);
        if ((N * N / 2) % THREADS != 0 && p2 >= uint32_t(N * N / 2)) break;
        const uint32_t I = p2 / (N / 2), J0 = 2 * (p2 % (N / 2));
        const uint32_t vo = (I * N + J0) * uint32_t(sizeof(cpx<T>));
        // t[h][item][a]
        cpx<T> t[2][2][S];
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const BufView &sv = m ? sv2 : sv1;
            const cpx<T> *dm = m ? d2 : d1;
            cpx2f x[4][S];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
#pragma unroll
                for (int a = 0; a < S; ++a) x[b][a] = ld2f(sv, vo, uint32_t(a + S * b) * PB);
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
#pragma unroll
                for (int a = 0; a < S; ++a) {
                    cpx<T> v = h ? x[0][a].b : x[0][a].a;
#pragma unroll
                    for (int b = 1; b < 4; ++b) fq::cmac(v, h ? x[b][a].b : x[b][a].a, dm[b]);
                    t[h][m][a] = v;
                }
            }
        }
        const cpx<T> px1 = ph1[I], px2 = ph2[I];
        const cpx2f py1 = *reinterpret_cast<const cpx2f *>(ph1 + N + J0);
        const cpx2f py2 = *reinterpret_cast<const cpx2f *>(ph2 + N + J0);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            // u = pxy1 cc1 t1 + i pxy2 cc2 t2, transformed once
            const cpx<T> pxy1 = fq::cmul(px1, h ? py1.b : py1.a);
            const cpx<T> q2 = fq::cmul(px2, h ? py2.b : py2.a);
            const cpx<T> pxy2 = {-q2.y, q2.x};
            cpx<T> u[S];
#pragma unroll
            for (int a = 0; a < S; ++a) {
                const cpx<T> v1 = a ? fq::cmul(t[h][0][a], cc1[a]) : t[h][0][a];
                const cpx<T> v2 = a ? fq::cmul(t[h][1][a], cc2[a]) : t[h][1][a];
                u[a] = fq::cmul(v1, pxy1);
                fq::cmac(u[a], v2, pxy2);
            }
            fq::LFFT<S, 1>::run(u);
            cpx<T> *dd = dst + I * PG::ROW + J0 + h;
#pragma unroll
            for (int k = 0; k < S; ++k) dd[k * PG::PLANE] = u[k];
        }
    }
}

// Pruned z-DFT of a conjugate pair: both items read one source and
// theta_2 = -theta_1 (every VHS product, and the same-species VSS pairs).
// With y_b = w_4^{bG} x_b, beta_b = p_b + i q_b, P = sum_b p_b y_b and
// Q = sum_b q_b y_b the two combinations are P + iQ and P - iQ, and with
// gamma = alpha_a pxy (item 1's phases)
//   gamma (P + iQ) + i conj(gamma) (P - iQ)
//       = (1 + i) [ Re((1 - i) gamma) P - Im((1 - i) gamma) Q ],
// so one pencil costs two real-coefficient sums and no per-item complex
// products.  zq[(prod*4 + G)*(3 + 2S) + .] holds {(p_b, q_b) for b = 1..3
// with the sign of w_4^{bG} folded in}, (1 - i) alpha_a and (1 + i) w_N^{aG}.  ODD (G odd): w_4^{bG} = +-i for odd b, applied as a
// rotation of x_b.  In FP32 each lane takes two neighbouring pencils
// (J, J + 1) per step with 16-byte loads: half the load instructions of one
// pencil per lane, which the vector memory pipeline, not the arithmetic,
// limits.
template <int S, typename T, bool ODD>
__device__ __forceinline__ void conj_pencil(const cpx<T> (&x)[4][S], cpx<T> pxy, const cpx<T> (&pq)[4],
                                            const cpx<T> (&am)[S], const cpx<T> (&cw)[S], cpx<T> (&u)[S]) {
#pragma unroll
    for (int a = 0; a < S; ++a) {
        cpx<T> P = x[0][a], Q;
#pragma unroll
        for (int b = 1; b < 4; ++b) {
            cpx<T> y = x[b][a];
            if (ODD && (b & 1)) y = {-y.y, y.x};
            P = {fq::fmadd(pq[b].x, y.x, P.x), fq::fmadd(pq[b].x, y.y, P.y)};
            if (b == 1) Q = {pq[b].y * y.x, pq[b].y * y.y};
            else Q = {fq::fmadd(pq[b].y, y.x, Q.x), fq::fmadd(pq[b].y, y.y, Q.y)};
        }
        const T A = fq::fmadd(am[a].x, pxy.x, -(am[a].y * pxy.y));
        const T B = -fq::fmadd(am[a].x, pxy.y, am[a].y * pxy.x);
        const cpx<T> z = {fq::fmadd(A, P.x, B * Q.x), fq::fmadd(A, P.y, B * Q.y)};
        u[a] = fq::cmul(z, cw[a]);
    }
    fq::LFFT<S, 1>::run(u);
}

template <int N, typename T, int THREADS, bool ODD>
__device__ __forceinline__ void zstage_conj(const cpx<T> *__restrict__ src, const cpx<T> *__restrict__ zq,
                                            const cpx<T> *__restrict__ ph1, cpx<T> *dst) {
    using PG = ProdGeo<N>;
    constexpr int S = PG::S;
    cpx<T> pq[4], am[S], cw[S];
#pragma unroll
    for (int b = 1; b < 4; ++b) pq[b] = zq[b - 1];
#pragma unroll
    for (int a = 0; a < S; ++a) { am[a] = zq[3 + a]; cw[a] = zq[3 + S + a]; }
    constexpr uint32_t PB = uint32_t(N * N * sizeof(cpx<T>));
    const BufView sv(src, uint32_t(N) * PB);
    if constexpr (sizeof(T) == 4) {
        // two neighbouring pencils (J, J + 1) per lane and step
        constexpr int ITER = (N * N / 2 + THREADS - 1) / THREADS;
#pragma unroll 1
        for (int it = 0; it < ITER; ++it) {
            const uint32_t p2 = threadIdx.x + uint32_t(it * THREADS);
            if ((N * N / 2) % THREADS != 0 && p2 >= uint32_t(N * N / 2)) break;
            const uint32_t I = p2 / (N / 2), J0 = 2 * (p2 % (N / 2));
            const uint32_t vo = (I * N + J0) * uint32_t(sizeof(cpx<T>));
            cpx2f x2[4][S];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
#pragma unroll
                for (int a = 0; a < S; ++a) x2[b][a] = ld2f(sv, vo, uint32_t(a + S * b) * PB);
            }
            const cpx<T> phx = ph1[I];
            const cpx2f phy = *reinterpret_cast<const cpx2f *>(ph1 + N + J0);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                cpx<T> x[4][S], u[S];
#pragma unroll
                for (int b = 0; b < 4; ++b) {
#pragma unroll
                    for (int a = 0; a < S; ++a) x[b][a] = h ? x2[b][a].b : x2[b][a].a;
                }
                conj_pencil<S, T, ODD>(x, fq::cmul(phx, h ? phy.b : phy.a), pq, am, cw, u);
                cpx<T> *dd = dst + I * PG::ROW + J0 + h;
#pragma unroll
                for (int k = 0; k < S; ++k) dd[k * PG::PLANE] = u[k];
            }
        }
    } else {
        // one pencil per lane and step (16-byte FP64 values)
        constexpr int ITER = (N * N + THREADS - 1) / THREADS;
#pragma unroll 1
        for (int it = 0; it < ITER; ++it) {
            const uint32_t p = threadIdx.x + uint32_t(it * THREADS);
            if ((N * N) % THREADS != 0 && p >= uint32_t(N * N)) break;
            const uint32_t I = p / N, J = p % N;
            const uint32_t vo = p * uint32_t(sizeof(cpx<T>));
            cpx<T> x[4][S], u[S];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
#pragma unroll
                for (int a = 0; a < S; ++a) x[b][a] = sv.ld<T>(vo, uint32_t(a + S * b) * PB);
            }
            conj_pencil<S, T, ODD>(x, fq::cmul(ph1[I], ph1[N + J]), pq, am, cw, u);
            cpx<T> *dd = dst + I * PG::ROW + J;
#pragma unroll
            for (int k = 0; k < S; ++k) dd[k * PG::PLANE] = u[k];
        }
    }
}

// Synchronisation around the in-place transpose of one LDS plane.  When a
// plane is exactly one wave (2N = 64 lanes) only that wave touches it: LDS
// executes one wave's accesses in issue order, so a wave-level barrier that
// keeps the compiler from moving LDS accesses across it is enough.  Planes
// that straddle waves need the workgroup barrier.
template <int N> __device__ __forceinline__ void plane_sync() {
    if constexpr (ProdGeo<N>::LPP == 64) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// E = DFT+_{x,y}(Z) of the LDS plane Lp (row stride ROW, phases already
// applied); the plane region is reused as the transpose buffer.
// Out: row vx = c, cyclic vy.
template <int N, typename T>
__device__ __forceinline__ void plane_inv_lds(cpx<T> *Lp, cpx<T> *v, int c, int half, bool act) {
    using PG = ProdGeo<N>;
    constexpr int E = PG::E;
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = act ? Lp[(half * E + e) * PG::ROW + c] : cpx<T>{0, 0};
    fq::fft2_dif<N, 1>(v, half);
    plane_sync<N>();
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) Lp[(half + 2 * e) * PG::ROW + c] = v[e];
    }
    plane_sync<N>();
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = act ? Lp[c * PG::ROW + half * E + e] : cpx<T>{0, 0};
    fq::fft2_dif<N, 1>(v, half);
}

template <int N, typename T, bool MULTI, bool PACK>
__global__ __launch_bounds__(ProdGeo<N>::THREADS, (sizeof(T) == 4 ? 4 : 2)) void k_prod(const cpx<T> *__restrict__ FT, cpx<T> *__restrict__ P, const cpx<T> *__restrict__ ph,
                                                              const cpx<T> *__restrict__ zc, const DevProd *__restrict__ prods,
                                                              const int *__restrict__ g_p0, const int *__restrict__ g_np,
                                                              const int *__restrict__ item_src, int ngroups, int nsrc,
                                                              const cpx<T> *__restrict__ NY, int nprods,
                                                              const cpx<T> *__restrict__ zq) {
    using PG = ProdGeo<N>;
    using G3 = Geo<N>;
    constexpr int E = PG::E;
    constexpr int ZC = 4 + PG::S;
    // LDS: the S planes, then (PACK, conjugate pairs only) the product's
    // Nyquist corrections as in k_prod1
    constexpr int NYL = PACK ? N * N + 2 * PG::S * N : 0;
    __shared__ cpx<T> lds[PG::S * PG::PLANE + NYL];
    // XCD-contiguous order when the grid allows it: the 4 * ngroups
    // workgroups of a DOF then share one L2 for its spectrum and corrections
    const int64_t unit = (gridDim.x & 7) == 0 ? xcd_block() : int64_t(blockIdx.x);
    const int G = int(unit % 4);
    const int g = int((unit / 4) % ngroups);
    const int slot = int(unit / (4 * int64_t(ngroups)));
    const int k = threadIdx.x / PG::LPP;
    const int lt = threadIdx.x % PG::LPP;
    const int c = lt >> 1, half = lt & 1;
    const bool act = c < N;
    const int vz = G + 4 * k;
    cpx<T> *L1 = lds + k * PG::PLANE;
    cpx<T> acc[E];
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = {0, 0};
    const int p0 = g_p0[g], np = g_np[g];
    for (int pi = p0; pi < p0 + np; ++pi) {
        const DevProd pr = prods[pi];
        const cpx<T> *src1 = FT + (int64_t(slot) * nsrc + item_src[pr.i1]) * G3::VOL;
        const cpx<T> *ph1 = ph + int64_t(pr.i1) * 3 * N;
        const T cf = T(pr.coef);
        if constexpr (PACK) {
            static_assert(PG::LPP == 64, "one plane per wave");
            cpx<T> ev[E];
            __syncthreads();                                   // LDS free
            const cpx<T> *zqp = zq + (int64_t(pi) * 4 + G) * (3 + 2 * PG::S);
            if (G & 1) zstage_conj<N, T, PG::THREADS, true>(src1, zqp, ph1, lds);
            else zstage_conj<N, T, PG::THREADS, false>(src1, zqp, ph1, lds);
            __syncthreads();
            constexpr int KPT = (N * N + PG::THREADS - 1) / PG::THREADS;
            constexpr int SN = PG::S * N;
            static_assert(SN <= PG::THREADS, "one J and one I element per thread");
            const cpx<T> *ny = NY + (int64_t(slot) * nprods + pi) * 3 * G3::N2;
            const int tid = int(threadIdx.x);
            cpx<T> *nyk = lds + PG::S * PG::PLANE, *nyj = nyk + N * N, *nyi = nyj + SN;
            // the corrections go to their own LDS area before the x,y
            // transforms: held in registers across them they would spill
#pragma unroll
            for (int r = 0; r < KPT; ++r) {
                const int idx = tid + r * PG::THREADS;
                if (N * N % PG::THREADS == 0 || idx < N * N) nyk[idx] = ny[idx];
            }
            if (tid < SN) {
                const int kk = tid / N, w = tid % N, z = G + 4 * kk;
                nyj[tid] = ny[G3::N2 + z * N + w];
                nyi[tid] = ny[2 * G3::N2 + w * N + z];
            }
            plane_inv_lds<N, T>(L1, ev, c, half, act);        // lane (c, half): vx = c, vy = half + 2e
            __syncthreads();
            const T sz = (vz & 1) ? T(-1) : T(1);
            const T sc = (c & 1) ? T(-1) : T(1);
            cpx<T> xj = act ? nyj[k * N + c] : cpx<T>{0, 0};
            if (half) xj = {-xj.x, -xj.y};                  // (-1)^vy, vy = half + 2e
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int vy = half + 2 * e;
                const cpx<T> xk = act ? nyk[vy * N + c] : cpx<T>{0, 0};
                const cpx<T> xi = nyi[k * N + vy];
                const T ux = fq::fmadd(sc, xi.x, fq::fmadd(sz, xk.x, xj.x));
                const T uy = fq::fmadd(sc, xi.y, fq::fmadd(sz, xk.y, xj.y));
                const T rx = ev[e].x - ux, ry = ev[e].y - uy;
                const T px = fq::fmadd(rx, ry, uy * ux);
                const T py = fq::fmadd(uy, ry, -(rx * ux));
                if constexpr (MULTI) acc[e] = {fq::fmadd(cf, px, acc[e].x), fq::fmadd(cf, py, acc[e].y)};
                else acc[e] = {px, py};                      // coefficient applied below
            }
        } else {
            const cpx<T> *src2 = FT + (int64_t(slot) * nsrc + item_src[pr.i2]) * G3::VOL;
            const cpx<T> *zc1 = zc + (int64_t(pr.i1) * 4 + G) * ZC, *zc2 = zc + (int64_t(pr.i2) * 4 + G) * ZC;
            const cpx<T> *ph2 = ph + int64_t(pr.i2) * 3 * N;
            cpx<T> e1[E], e2[E];
            __syncthreads();                                   // LDS free
            zstage<N, T>(src1, zc1, ph1, lds);
            __syncthreads();
            plane_inv_lds<N, T>(L1, e1, c, half, act);
            __syncthreads();
            zstage<N, T>(src2, zc2, ph2, lds);
            __syncthreads();
            plane_inv_lds<N, T>(L1, e2, c, half, act);
            if constexpr (MULTI) {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    const cpx<T> pe = fq::cmul(e1[e], e2[e]);
                    acc[e] = {fq::fmadd(cf, pe.x, acc[e].x), fq::fmadd(cf, pe.y, acc[e].y)};
                }
            } else {
                // one product per group: acc is not live across the item stages
#pragma unroll
                for (int e = 0; e < E; ++e) acc[e] = fq::cscale(fq::cmul(e1[e], e2[e]), cf);
            }
        }
    }
    if constexpr (PACK && !MULTI) {
        // the single product's coefficient (1 for VHS and VSS), applied once
        const T cf = np > 0 ? T(prods[p0].coef) : T(1);
        if (cf != T(1)) {
#pragma unroll
            for (int e = 0; e < E; ++e) acc[e] = fq::cscale(acc[e], cf);
        }
    }
    // forward transforms of the group sum: acc is row vx = c, cyclic over vy
    fq::fft2_dit<N, -1>(acc, half);                       // -> J' = half*E + e
    plane_sync<N>();
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) L1[c * PG::ROW + half * E + e] = acc[e];
    }
    plane_sync<N>();
#pragma unroll
    for (int e = 0; e < E; ++e) acc[e] = act ? L1[(half + 2 * e) * PG::ROW + c] : cpx<T>{0, 0};
    fq::fft2_dit<N, -1>(acc, half);                       // -> I' = half*E + e
    cpx<T> *dst = P + (int64_t(slot) * ngroups + g) * G3::VOL + int64_t(vz) * G3::N2;
    if (act) {
        // FP64: non-temporal stores, the volumes are read once by the
        // accumulation kernel (6% faster)
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if constexpr (sizeof(T) == 8)
                __builtin_nontemporal_store(nt_d2{acc[e].x, acc[e].y},
                                            reinterpret_cast<nt_d2 *>(dst + (half * E + e) * N + c));
            else
                dst[(half * E + e) * N + c] = acc[e];
        }
    }
}

// ---------------------------------------------------------------------------
// k_prod1: the product stage with one lane per pencil (FP32, N = 24, 32).
// The workgroup (slot, group, G) is the same as k_prod's, with the same
// pruned z stage, but each half-wave (32 lanes) holds one plane, one pencil
// per lane (lanes past N idle for N = 24): the x and y transforms are
// in-register FFTs with no lane exchange.  This trades occupancy (whole
// pencils need about twice the registers) for about a third fewer
// instructions per transform.
// ---------------------------------------------------------------------------
template <int N> struct Prod1Geo {
    static constexpr int S = N / 4;
    static_assert(S % 2 == 0 && N <= 32, "two planes per wave");
    static constexpr int THREADS = (S / 2) * 64;
};

template <int N, typename T>
__device__ __forceinline__ void plane_inv1(cpx<T> *Lp, cpx<T> *v, int c, bool act) {
    using PG = ProdGeo<N>;
    // column c over I -> vx (natural order)
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = act ? Lp[i * PG::ROW + c] : cpx<T>{0, 0};
    fq::LFFT<N, 1>::run(v);
    if (act) {
#pragma unroll
        for (int i = 0; i < N; ++i) Lp[i * PG::ROW + c] = v[i];
    }
    plane_sync<32>();                              // the plane lives in one wave
    // row vx = c over J -> vy
#pragma unroll
    for (int j = 0; j < N; ++j) v[j] = act ? Lp[c * PG::ROW + j] : cpx<T>{0, 0};
    fq::LFFT<N, 1>::run(v);
}

// Pair packing (PACK): every item volume T = phi(theta) . F_s is Hermitian,
// T(-l) = conj T(l) (F_s is the spectrum of a real f), except on the Nyquist
// set where a component of l is the unpaired index N/2.  So E = DFT+(T) =
// R + i S with R real and S = -i DFT+(A) from the anti-Hermitian part
// A(l) = T(l) (1 - e^{i N theta.nyq(l)}) / 2, nyq(l) the Nyquist components
// of l, and one complex transform of T_1 + i T_2 carries both items of a
// product: its real part is R_1 - S_2, its imaginary part S_1 + R_2.
// k_nyq evaluates S through the three Nyquist planes (see there); with
// X = DFT+ of the packed A over one plane (X = i a_1 - a_2, a_t real),
//   U(vx,vy,vz) = (-1)^vz X_K(vx,vy) + (-1)^vy X_J(vx,vz) + (-1)^vx X_I(vy,vz)
// gives S_1 = U.y and S_2 = -U.x, hence with R = E - U
//   E_1 = (R.x, U.y),  E_2 = (R.y, -U.x).
// This halves the z stage's transforms and the x,y transforms of the items.
//
// Group pairs (HP, 32^3 FP32, single-product groups): only the Hermitian part
// of the gain spectrum reaches Q (Q takes Re DFT+ of it), and with
// G_g = a_g + i b_g (a, b real) that part is
//   sum_c U_c DFT-(a_g) + V_c DFT-(b_g),
//   U_c = (W_c + W~_c)/2, V_c = i (W_c - W~_c)/2, W~_c(xi) = conj W_c(-xi),
// where V_c vanishes off the Nyquist set.  So the workgroup evaluates the
// two groups 2u and 2u + 1 and transforms a_2u + i a_2u+1 once; the
// accumulation separates the two Hermitian spectra (k_accum_h).  Of b_g it
// keeps the sums that give DFT-(b_g) on the three Nyquist planes
// (K' = N/2: sum over vz of (-1)^vz b; J' = N/2: over vy; I' = N/2: over vx),
// written to NB as per-workgroup partials for k_nyqb:
//   NB[slot][g] = { K: [G][vy][vx] (signed by (-1)^G), J: [vz][vx], I: [vz][vy] }.
constexpr int kNbPer32 = 6 * 32 * 32;             // floats of NB per (slot, group), N = 32

template <int N, typename T, bool MULTI, bool PACK, bool CONJ, bool HP = false>
__global__ __launch_bounds__(Prod1Geo<N>::THREADS, 2) void k_prod1(
    const cpx<T> *__restrict__ FT, cpx<T> *__restrict__ P, const cpx<T> *__restrict__ ph,
    const cpx<T> *__restrict__ zc, const DevProd *__restrict__ prods, const int *__restrict__ g_p0,
    const int *__restrict__ g_np, const int *__restrict__ item_src, int ngroups, int nsrc,
    const cpx<T> *__restrict__ NY, int nprods, const cpx<T> *__restrict__ zq, const int *__restrict__ pconj,
    float *__restrict__ NB) {
    static_assert(!HP || (PACK && !MULTI && N == 32 && sizeof(T) == 4), "group pairs: 32^3 FP32 packed kernel");
    using PG = ProdGeo<N>;
    using P1 = Prod1Geo<N>;
    using G3 = Geo<N>;
    constexpr int ZC = 4 + PG::S;
    // LDS: the S planes, then (PACK) the product's Nyquist corrections:
    // K piece [vy][vx], J piece [k][vx] and I piece [k][vy] for vz = G + 4k
    constexpr int NYL = PACK ? N * N + 2 * PG::S * N : 0;
    __shared__ cpx<T> lds[PG::S * PG::PLANE + NYL];
    // XCD-contiguous order when the grid allows it: the 4 * nvol
    // workgroups of a DOF then share one L2 for its spectrum and corrections
    const int64_t unit = (gridDim.x & 7) == 0 ? xcd_block() : int64_t(blockIdx.x);
    const int nvol = HP ? (ngroups + 1) / 2 : ngroups;    // volumes written per DOF
    const int G = int(unit % 4);
    const int gv = int((unit / 4) % nvol);
    const int slot = int(unit / (4 * int64_t(nvol)));
    const int k = int(threadIdx.x) / 32;           // plane of this half-wave
    const int c = int(threadIdx.x) % 32;           // pencil (column, then row) of this lane
    const bool act = c < N;
    const int vz = G + 4 * k;
    cpx<T> *L1 = lds + k * PG::PLANE;
    cpx<T> acc[N];
    T re0[HP ? N : 1];
    for (int m = 0; m < (HP ? 2 : 1); ++m) {
        const int g = HP ? 2 * gv + m : gv;
#pragma unroll
        for (int e = 0; e < N; ++e) acc[e] = {0, 0};
        if (HP && g >= ngroups) break;                 // odd group count: the pair's second member is zero
        const int p0 = g_p0[g], np = g_np[g];
        for (int pi = p0; pi < p0 + np; ++pi) {
            const DevProd pr = prods[pi];
            const cpx<T> *src1 = FT + (int64_t(slot) * nsrc + item_src[pr.i1]) * G3::VOL;
            const cpx<T> *src2 = FT + (int64_t(slot) * nsrc + item_src[pr.i2]) * G3::VOL;
            const cpx<T> *zc1 = zc + (int64_t(pr.i1) * 4 + G) * ZC, *zc2 = zc + (int64_t(pr.i2) * 4 + G) * ZC;
            const cpx<T> *ph1 = ph + int64_t(pr.i1) * 3 * N, *ph2 = ph + int64_t(pr.i2) * 3 * N;
            const T cf = T(pr.coef);
            if constexpr (PACK) {
                cpx<T> e[N];
                __syncthreads();                               // LDS free
                // conjugate products (one source, opposite phases) take the cheaper
                // z-stage: every product (CONJ) or those marked in pconj
                if (CONJ || (pconj && pconj[pi])) {
                    const cpx<T> *zqp = zq + (int64_t(pi) * 4 + G) * (3 + 2 * PG::S);
                    if (G & 1) zstage_conj<N, T, P1::THREADS, true>(src1, zqp, ph1, lds);
                    else zstage_conj<N, T, P1::THREADS, false>(src1, zqp, ph1, lds);
                } else {
                    zstage_pair<N, T, P1::THREADS>(src1, src2, zc1, zc2, ph1, ph2, lds);
                }
                __syncthreads();
                // the product's corrections: global loads in flight across the
                // x,y transforms, then staged in LDS for the whole workgroup
                constexpr int KPT = (N * N + P1::THREADS - 1) / P1::THREADS;
                constexpr int SN = PG::S * N;
                static_assert(SN <= P1::THREADS, "one J and one I element per thread");
                const cpx<T> *ny = NY + (int64_t(slot) * nprods + pi) * 3 * G3::N2;
                const int tid = int(threadIdx.x);
                cpx<T> rk[KPT], rj = {0, 0}, ri = {0, 0};
#pragma unroll
                for (int r = 0; r < KPT; ++r) {
                    const int idx = tid + r * P1::THREADS;
                    rk[r] = (N * N % P1::THREADS == 0 || idx < N * N) ? ny[idx] : cpx<T>{0, 0};
                }
                if (tid < SN) {
                    const int kk = tid / N, w = tid % N, z = G + 4 * kk;
                    rj = ny[G3::N2 + z * N + w];                       // (vx = w, vz = z)
                    ri = ny[2 * G3::N2 + w * N + z];                   // (vy = w, vz = z)
                }
                plane_inv1<N, T>(L1, e, c, act);               // lane c: row vx = c over vy
                cpx<T> *nyk = lds + PG::S * PG::PLANE, *nyj = nyk + N * N, *nyi = nyj + SN;
#pragma unroll
                for (int r = 0; r < KPT; ++r) {
                    const int idx = tid + r * P1::THREADS;
                    if (N * N % P1::THREADS == 0 || idx < N * N) nyk[idx] = rk[r];
                }
                if (tid < SN) { nyj[tid] = rj; nyi[tid] = ri; }
                __syncthreads();
                const T sz = (vz & 1) ? T(-1) : T(1);
                const T sc = (c & 1) ? T(-1) : T(1);
                const cpx<T> xj = act ? nyj[k * N + c] : cpx<T>{0, 0};
#pragma unroll
                for (int j = 0; j < N; ++j) {
                    const cpx<T> xk = act ? nyk[j * N + c] : cpx<T>{0, 0};
                    const cpx<T> xi = nyi[k * N + j];
                    const T jx = (j & 1) ? -xj.x : xj.x, jy = (j & 1) ? -xj.y : xj.y;
                    const T ux = fq::fmadd(sc, xi.x, fq::fmadd(sz, xk.x, jx));
                    const T uy = fq::fmadd(sc, xi.y, fq::fmadd(sz, xk.y, jy));
                    const T rx = e[j].x - ux, ry = e[j].y - uy;
                    // (rx + i uy)(ry - i ux)
                    const T px = fq::fmadd(rx, ry, uy * ux);
                    const T py = fq::fmadd(uy, ry, -(rx * ux));
                    if constexpr (MULTI) acc[j] = {fq::fmadd(cf, px, acc[j].x), fq::fmadd(cf, py, acc[j].y)};
                    else acc[j] = {px, py};                      // coefficient applied below
                }
            } else {
                cpx<T> e1[N], e2[N];
                __syncthreads();                               // LDS free
                zstage<N, T, P1::THREADS>(src1, zc1, ph1, lds);
                __syncthreads();
                plane_inv1<N, T>(L1, e1, c, act);
                __syncthreads();
                zstage<N, T, P1::THREADS>(src2, zc2, ph2, lds);
                __syncthreads();
                plane_inv1<N, T>(L1, e2, c, act);
                if constexpr (MULTI) {
#pragma unroll
                    for (int e = 0; e < N; ++e) {
                        const cpx<T> pe = fq::cmul(e1[e], e2[e]);
                        acc[e] = {fq::fmadd(cf, pe.x, acc[e].x), fq::fmadd(cf, pe.y, acc[e].y)};
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < N; ++e) acc[e] = fq::cscale(fq::cmul(e1[e], e2[e]), cf);
                }
            }
        }
        if constexpr (PACK && !MULTI) {
            // the single product's coefficient (1 for VHS and VSS), applied once
            const T cf = np > 0 ? T(prods[p0].coef) : T(1);
            if (cf != T(1)) {
#pragma unroll
                for (int e = 0; e < N; ++e) acc[e] = fq::cscale(acc[e], cf);
            }
        }
        if constexpr (HP) {
            if (NB) {
                // Nyquist-plane sums of b = Im G_g.  The plane region is free: its
                // last reads (plane_inv1) precede the barrier before the unpack.
                float *nb = NB + (int64_t(slot) * ngroups + g) * kNbPer32;
                T sj = 0;
#pragma unroll
                for (int j = 0; j < N; ++j) sj += (j & 1) ? -acc[j].y : acc[j].y;
                if (act) nb[4 * N * N + vz * N + c] = sj;                   // J: (vz, vx = c)
                // b planes as [k][vy][vx] with rows of BR = N + 4 floats: lane c's
                // stores for one vy are consecutive, a row is 16-byte aligned, and
                // 16 rows read by 16 lanes start on 16 different 4-bank groups
                constexpr int BR = N + 4;
                T *bl = reinterpret_cast<T *>(lds);                        // [S][N][BR]
                static_assert(PG::S * N * BR * sizeof(T) <= PG::S * PG::PLANE * sizeof(cpx<T>), "fits the planes");
                if (act) {
#pragma unroll
                    for (int j = 0; j < N; ++j) bl[(k * N + j) * BR + c] = acc[j].y;
                }
                __syncthreads();
                const int tid = int(threadIdx.x);
                const T sgn = (G & 1) ? T(-1) : T(1);                      // (-1)^vz on vz = G + 4k
                static_assert(N * N == 4 * P1::THREADS && P1::THREADS == PG::S * N, "thread map");
                {
                    // K: four consecutive vx of one vy per thread, summed over the S planes
                    const int y = tid / (N / 4), x0 = 4 * (tid % (N / 4));
                    float4 sk = {0, 0, 0, 0};
#pragma unroll
                    for (int kk = 0; kk < PG::S; ++kk) {
                        const float4 q = *reinterpret_cast<const float4 *>(bl + (kk * N + y) * BR + x0);
                        sk.x += q.x; sk.y += q.y; sk.z += q.z; sk.w += q.w;
                    }
                    *reinterpret_cast<float4 *>(nb + G * N * N + y * N + x0) =
                        float4{sgn * sk.x, sgn * sk.y, sgn * sk.z, sgn * sk.w};       // K: [G][vy][vx]
                }
                {
                    // I: the row (plane kk, vy) summed over vx with the sign (-1)^vx
                    const int kk = tid / N, y = tid % N;
                    T sI = 0;
#pragma unroll
                    for (int x4 = 0; x4 < N / 4; ++x4) {
                        const float4 q = *reinterpret_cast<const float4 *>(bl + (kk * N + y) * BR + 4 * x4);
                        sI += (q.x - q.y) + (q.z - q.w);
                    }
                    nb[5 * N * N + (G + 4 * kk) * N + y] = sI;               // I: (vz, vy)
                }
                // LDS reads done before the region is written again: by the
                // forward transforms after the pair's last group, or by the
                // next group's z stage, which is preceded by its own barrier
                if (m == 1 || 2 * gv + 1 >= ngroups) __syncthreads();
            }
            if (m == 0) {
#pragma unroll
                for (int j = 0; j < N; ++j) re0[j] = acc[j].x;
            }
        }
    }
    if constexpr (HP) {
        // a_2u + i a_2u+1 (acc is zero when the second member is absent)
#pragma unroll
        for (int j = 0; j < N; ++j) acc[j] = {re0[j], acc[j].x};
    }
    // forward transforms of the group sum: lane c holds row vx = c over vy
    fq::LFFT<N, -1>::run(acc);                            // -> J'
    plane_sync<32>();
    if (act) {
#pragma unroll
        for (int j = 0; j < N; ++j) L1[c * PG::ROW + j] = acc[j];
    }
    plane_sync<32>();
#pragma unroll
    for (int i = 0; i < N; ++i) acc[i] = act ? L1[i * PG::ROW + c] : cpx<T>{0, 0};   // column J' = c over vx
    fq::LFFT<N, -1>::run(acc);                            // -> I'
    cpx<T> *dst = P + (int64_t(slot) * nvol + gv) * G3::VOL + int64_t(vz) * G3::N2;
    // 16-byte stores: lanes c and c ^ 1 swap one value per row pair, the
    // even lane then writes row 2m and the odd lane row 2m + 1 of columns
    // (c & ~1, c | 1).  Non-temporal: the volumes are read once, by the
    // accumulation kernel after this kernel has finished (1% faster)
    const bool odd = c & 1;
#pragma unroll
    for (int m = 0; m < N / 2; ++m) {
        const cpx<T> recv = fq::swap_adj(odd ? acc[2 * m] : acc[2 * m + 1]);
        const cpx2f w = odd ? cpx2f{recv, acc[2 * m + 1]} : cpx2f{acc[2 * m], recv};
        if (act)
            __builtin_nontemporal_store(nt_f4{w.a.x, w.a.y, w.b.x, w.b.y},
                                        reinterpret_cast<nt_f4 *>(dst + (2 * m + (odd ? 1 : 0)) * N + (c & ~1)));
    }
}

// ---------------------------------------------------------------------------
// k_nyq: the Nyquist-plane corrections of the pair-packed product kernels
// (k_prod1 in FP32; the two-lane k_prod for conjugate pairs in FP64).
// One workgroup per (slot, product), one 32-lane group per plane piece:
//   piece 0 (K = N/2):                 rows a = I, lanes = J   -> NY0[vy][vx]
//   piece 1 (J = N/2, K != N/2):       rows a = I, lanes = K   -> NY1[vz][vx]
//   piece 2 (I = N/2, J, K != N/2):    rows a = K, lanes = J   -> NY2[vy][vz]
// Each lane gathers its column of A_1 + i A_2 (A_t = T_t (1 - e^{i N
// theta_t.nyq}) / 2 on the piece), and the piece's 2D DFT+ runs as a column
// transform, an LDS transpose and a row transform.  General pairs weight the
// two sources with the per-product table V_1 = phi_1 (1 - e^{..}) / 2,
// V_2 = i phi_2 (1 - e^{..}) / 2 (zero off the piece); conjugate pairs (SH)
// compute a real weight per point from item 1's phase table.
// ---------------------------------------------------------------------------
template <int N, typename T, bool SH>
__global__ __launch_bounds__(96, (sizeof(T) == 4 ? 4 : 2)) void k_nyq(const cpx<T> *__restrict__ FT, const cpx<T> *__restrict__ FJ,
                                               cpx<T> *__restrict__ NY,
                                               const cpx<T> *__restrict__ V, const cpx<T> *__restrict__ ph,
                                               const DevProd *__restrict__ prods,
                                               const int *__restrict__ item_src, int nprods, int nsrc) {
    using G3 = Geo<N>;
    constexpr int H = N / 2, ROW = N + 1;
    __shared__ cpx<T> lds[3 * N * ROW + (SH ? 2 * N : 0)];
    const int pi = int(blockIdx.x % unsigned(nprods));
    const int slot = int(blockIdx.x / unsigned(nprods));
    const int piece = int(threadIdx.x) / 32, c = int(threadIdx.x) % 32;
    const bool act = c < N;
    const DevProd pr = prods[pi];
    // element a of this lane's column: base + a * stride in the source volume
    // (pieces 0, 2) or in its J = N/2 slice [I][K] (piece 1)
    auto col = [&](int it) {
        const int64_t s = int64_t(slot) * nsrc + item_src[it];
        return piece == 1 ? FJ + s * G3::N2 : FT + s * G3::VOL;
    };
    const cpx<T> *F1 = col(pr.i1);
    const int base = piece == 0 ? H * G3::N2 + c : (piece == 1 ? c : H * N + c);
    const int stride = piece == 2 ? G3::N2 : N;
    cpx<T> v[N];
    if constexpr (SH) {
        // conjugate pair: with psi = phi_1 on the piece and h = (1 - e^{i N
        // theta_1.nyq}) / 2, A_1 + i A_2 = F (psi h + i conj(psi h))
        // = (1 + i) F Re((1 - i) psi h): a real weight per point, and the
        // factor (1 + i) is applied to the transform.  psi = phi_row(a) q
        // with q = phi_lane(c) phi_fixed(N/2) per lane; phi_row staged in LDS.
        const cpx<T> *p1 = ph + int64_t(pr.i1) * 3 * N;
        cpx<T> *prow = lds + 3 * N * ROW;                  // [0, N): phi_x, [N, 2N): phi_z
        if (threadIdx.x < 2 * N) prow[threadIdx.x] = p1[threadIdx.x < N ? threadIdx.x : threadIdx.x + N];
        // dims (x = 0, y = 1, z = 2) of the row, lane and fixed coordinate
        const int drow = piece == 2 ? 2 : 0, dlane = piece == 1 ? 2 : 1, dfix = piece == 0 ? 2 : (piece == 1 ? 1 : 0);
        auto nq = [&](int d) { const cpx<T> u = {p1[d * N + H].x, -p1[d * N + H].y}; return fq::cmul(u, u); };
        const cpx<T> q = fq::cmul(p1[dlane * N + (act ? c : 0)], p1[dfix * N + H]);
        cpx<T> w0 = nq(dfix);
        if (c == H) w0 = fq::cmul(w0, nq(dlane));
        const cpx<T> wH = fq::cmul(w0, nq(drow));
        auto qh = [&](cpx<T> w) {                       // (1 - i) q (1 - w) / 2
            const cpx<T> hq = fq::cmul(q, cpx<T>{T(0.5) * (T(1) - w.x), T(-0.5) * w.y});
            return cpx<T>{hq.x + hq.y, hq.y - hq.x};
        };
        const bool drop = !act || (piece == 1 && c == H) || (piece == 2 && c == H);
        const cpx<T> Q0 = drop ? cpx<T>{0, 0} : qh(w0);
        const cpx<T> QH = (drop || piece == 2) ? cpx<T>{0, 0} : qh(wH);
        __syncthreads();
        const cpx<T> *pr0 = prow + (drow == 2 ? N : 0);
#pragma unroll
        for (int a = 0; a < N; ++a) {
            const cpx<T> Q = a == H ? QH : Q0;
            const cpx<T> r = pr0[a];
            const T sw = fq::fmadd(r.x, Q.x, -(r.y * Q.y));
            const cpx<T> x = act ? F1[base + a * stride] : cpx<T>{0, 0};
            v[a] = {x.x * sw, x.y * sw};
        }
    } else {
        const cpx<T> *F2 = col(pr.i2);
        const cpx<T> *v1 = V + (int64_t(pi) * 6 + piece) * G3::N2 + c;     // [product][item][piece][a][lane]
        const cpx<T> *v2 = v1 + 3 * G3::N2;
#pragma unroll
        for (int a = 0; a < N; ++a) {
            if (!act) { v[a] = {0, 0}; continue; }
            const int off = base + a * stride;
            v[a] = fq::cmul(F1[off], v1[a * N]);
            fq::cmac(v[a], F2[off], v2[a * N]);
        }
    }
    fq::LFFT<N, 1>::run(v);                                // over a, lane c
    cpx<T> *L = lds + piece * N * ROW;
    if (act) {
#pragma unroll
        for (int a = 0; a < N; ++a) L[a * ROW + c] = v[a];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < N; ++b) v[b] = act ? L[c * ROW + b] : cpx<T>{0, 0};
    fq::LFFT<N, 1>::run(v);                                // over the lane coordinate
    if constexpr (SH) {
#pragma unroll
        for (int j = 0; j < N; ++j) v[j] = {v[j].x - v[j].y, v[j].x + v[j].y};   // (1 + i) X
    }
    cpx<T> *out = NY + ((int64_t(slot) * nprods + pi) * 3 + piece) * G3::N2;
    if constexpr (sizeof(T) == 4) {
        // 16-byte non-temporal stores after a swap with lane c ^ 1 (as in
        // k_prod1): the planes are read by the next kernel
        const bool odd = c & 1;
#pragma unroll
        for (int m = 0; m < N / 2; ++m) {
            const cpx<T> recv = fq::swap_adj(odd ? v[2 * m] : v[2 * m + 1]);
            const cpx2f w = odd ? cpx2f{recv, v[2 * m + 1]} : cpx2f{v[2 * m], recv};
            if (act)
                __builtin_nontemporal_store(nt_f4{w.a.x, w.a.y, w.b.x, w.b.y},
                                            reinterpret_cast<nt_f4 *>(out + (2 * m + (odd ? 1 : 0)) * N + (c & ~1)));
        }
    } else {
        if (act) {
#pragma unroll
            for (int j = 0; j < N; ++j) out[j * N + c] = v[j];
        }
    }
}


// Minimum waves per SIMD requested for k_accum: three for FP32 transforms
// (the 24-point radix-3 variant needs more registers and keeps two), two for FP64.
constexpr int accum_waves(int N, size_t tsize) { return tsize == 4 && N != 24 ? 3 : 2; }

// ---------------------------------------------------------------------------
// k_accum: unit = (slot, target, I').  H = sum_c W_c . DFT-_z(P_g(c)),
// L = sum_j Lt_j . F_src(j); Y = DFT+_z of H and L.  FP64 accumulation.
// ---------------------------------------------------------------------------
template <int N, typename T>
__global__ __launch_bounds__(256, accum_waves(N, sizeof(T))) void k_accum(const cpx<T> *P, const cpx<double> *F,
                                               const cpx<double> *W, const int *c_group,
                                               const int *t_c0, const int *t_cn, const int *l_src,
                                               const int *t_l0, const int *t_ln, int ncontribs,
                                               cpx<double> *Y, int ngroups, int nsrc, int ntgt, int D,
                                               int64_t nunits) {
    using G = Geo<N>;
    constexpr int E = G::E;
    const int lt = threadIdx.x % G::LPP;
    int64_t unit = int64_t(blockIdx.x) * (blockDim.x / G::LPP) + threadIdx.x / G::LPP;
    // With 64 lanes per unit a unit is one wave: its indices are wave-uniform
    // and live in SGPRs, and the pencil loads below use buffer views.
    constexpr bool UNI = G::LPP == 64;
    if constexpr (UNI) unit = __builtin_amdgcn_readfirstlane(int(unit));
    if (unit >= nunits) return;
    // slot fastest: neighbouring units read the same weight rows W[c][.][I]
    const int slot = int(unit % D);
    const int64_t ti = unit / D;
    const int t = int(ti % ntgt);
    const int I = int(ti / ntgt);
    const int J = lt >> 1, half = lt & 1;
    const bool act = J < N;
    const int64_t off = int64_t(I) * N + J;
    cpx<double> H[E];
#pragma unroll
    for (int e = 0; e < E; ++e) H[e] = {0, 0};
    const int c0 = t_c0[t], cn = t_cn[t];
    const cpx<T> *pbase = P + int64_t(slot) * ngroups * G::VOL + off;
    // lane byte offsets inside one volume (P pencil element half + 2e, W element half*E + e)
    const uint32_t voP = uint32_t(off + int64_t(half) * G::N2) * uint32_t(sizeof(cpx<T>));
    const uint32_t voW = uint32_t(off + int64_t(half) * E * G::N2) * uint32_t(sizeof(cpx<double>));
    auto load_pencil = [&](int ci, cpx<T> *dst) {
        if constexpr (UNI) {
            const BufView pv(P + (int64_t(slot) * ngroups + c_group[ci]) * G::VOL, uint32_t(G::VOL * sizeof(cpx<T>)));
#pragma unroll
            for (int e = 0; e < E; ++e)
                dst[e] = act ? pv.ld<T>(voP, uint32_t(2 * e * G::N2 * sizeof(cpx<T>))) : cpx<T>{0, 0};
        } else {
            const cpx<T> *src = pbase + int64_t(c_group[ci]) * G::VOL;
#pragma unroll
            for (int e = 0; e < E; ++e) dst[e] = act ? src[int64_t(half + 2 * e) * G::N2] : cpx<T>{0, 0};
        }
    };
    // FP32 pencils: software pipeline, the pencil of contribution ci+1 is in
    // flight while contribution ci is transformed and accumulated.  FP64
    // pencils are loaded in place (the register budget has no room for both).
    constexpr bool PF = sizeof(T) == 4;
    cpx<T> nxt[PF ? E : 1];
    if constexpr (PF) {
        if (cn > 0) load_pencil(c0, nxt);
    }
    for (int ci = c0; ci < c0 + cn; ++ci) {
        cpx<T> v[E];
        if constexpr (PF) {
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = nxt[e];
            if (ci + 1 < c0 + cn) load_pencil(ci + 1, nxt);
        } else {
            load_pencil(ci, v);
        }
        fq::fft2_dit<N, -1>(v, half);             // -> K' = half*E + e
        if constexpr (UNI) {
            const BufView wv(W + int64_t(ci) * G::VOL, uint32_t(G::VOL * sizeof(cpx<double>)));
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const cpx<double> we =
                    act ? wv.ld<double>(voW, uint32_t(e * G::N2 * sizeof(cpx<double>))) : cpx<double>{0, 0};
                fq::cmac(H[e], we, cvt<double>(v[e]));
            }
        } else {
            const cpx<double> *w = W + int64_t(ci) * G::VOL + off;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const cpx<double> we = act ? w[int64_t(half * E + e) * G::N2] : cpx<double>{0, 0};
                fq::cmac(H[e], we, cvt<double>(v[e]));
            }
        }
    }
    fq::fft2_dif<N, 1>(H, half);                  // -> vz = half + 2e
    cpx<double> *yg = Y + ((int64_t(slot) * ntgt + t) * 2 + 0) * G::VOL + off;
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) yg[int64_t(half + 2 * e) * G::N2] = H[e];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) H[e] = {0, 0};
    const int l0 = t_l0[t], ln = t_ln[t];
    for (int li = l0; li < l0 + ln; ++li) {
        const int s = l_src[li];
        const cpx<double> *fs = F + (int64_t(slot) * nsrc + s) * G::VOL + off;
        const cpx<double> *w = W + int64_t(ncontribs + li) * G::VOL + off;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int64_t k = int64_t(half * E + e) * G::N2;
            if (act) H[e] = H[e] + fq::cmul(w[k], fs[k]);
        }
    }
    fq::fft2_dif<N, 1>(H, half);
    cpx<double> *yl = Y + ((int64_t(slot) * ntgt + t) * 2 + 1) * G::VOL + off;
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) yl[int64_t(half + 2 * e) * G::N2] = H[e];
    }
}

// ---------------------------------------------------------------------------
// k_accum_lds: the same accumulation on the 32^3 grid (a pencil row fills
// one wave).  One workgroup = NS consecutive DOF slots of one
// (target, row I'), one wave each (NS = 4, two workgroups per CU, or NS = 8,
// one workgroup per CU and half the weight reads).  The weight row block W[c][.][I'][.]
// (N x N FP64 values) is the same for the four slots, so the workgroup loads
// it once per contribution into a double-buffered LDS area, one contribution
// ahead of its use, and the waves read it from LDS: the weight latency is
// hidden behind a whole contribution and weight reads from L2 drop NS-fold.
// Waves whose slot lies past the batch take part in the copies and barriers
// only.
// ---------------------------------------------------------------------------
template <int N, typename T, int NS = 4>
__global__ __launch_bounds__(64 * NS, 8 / NS) void k_accum_lds(
    const cpx<T> *__restrict__ P, const cpx<double> *__restrict__ F, const cpx<double> *__restrict__ W,
    const int *__restrict__ c_group, const int *__restrict__ t_c0, const int *__restrict__ t_cn,
    const int *__restrict__ l_src, const int *__restrict__ t_l0, const int *__restrict__ t_ln, int ncontribs,
    cpx<double> *__restrict__ Y, int ngroups, int nsrc, int ntgt, int D) {
    using G = Geo<N>;
    static_assert(G::LPP == 64, "one pencil row per wave");
    constexpr int E = G::E;
    constexpr int WR = N * N;                     // weight row block (K', J)
    constexpr int TH = 64 * NS;                   // NS waves, one DOF slot each
    constexpr int CP = (WR + TH - 1) / TH;        // 16-byte copies per thread
    __shared__ cpx<double> wl[2][WR];
    const int nsg = (D + NS - 1) / NS;
    const int b = blockIdx.x;
    const int sg = b % nsg;
    const int t = (b / nsg) % ntgt;
    const int I = b / (nsg * ntgt);
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / 64);
    const int slot = sg * NS + wave;
    const bool live = slot < D;                   // wave-uniform
    const int lt = threadIdx.x % 64;
    const int J = lt >> 1, half = lt & 1;
    const bool act = J < N;
    const int64_t off = int64_t(I) * N + J;
    const int c0 = t_c0[t], cn = t_cn[t];
    // weight copies: element idx = K' * N + J' of the row block (written out
    // in place below; a lambda capturing the array would put it in scratch)
    cpx<double> wp[CP];
    const cpx<double> *wrow = W + int64_t(I) * N;
    // FP64 pencils load through a buffer view of the (wave-uniform) group
    // volume (one lane offset register, plane offsets in SGPRs), which frees
    // the registers that keep the FP64 kernel at three waves per SIMD.  FP32
    // pencils load as 16-byte pairs: lane (J, half) reads columns J & ~1 and
    // J | 1 at the planes half + 2e with e = jo + 2f (jo = J & 1), and the
    // value of the other column goes to the lane of J ^ 1 (lane ^ 2) in
    // pair_swap, so a pencil costs E/2 load instructions instead of E.
    const cpx<T> *pslot = P + int64_t(live ? slot : 0) * ngroups * G::VOL;
    const uint32_t voP = uint32_t(off + int64_t(half) * G::N2) * uint32_t(sizeof(cpx<T>));
    const int jo = J & 1;
    auto load_pencil = [&](int ci, cpx<T> *dst) {
        if constexpr (sizeof(T) == 8) {
            // non-temporal: each pencil is read once (the weight rows stay in L2; 4% faster)
            const BufView pv(pslot + int64_t(c_group[ci]) * G::VOL, uint32_t(G::VOL * sizeof(cpx<T>)));
#pragma unroll
            for (int e = 0; e < E; ++e)
                dst[e] = (live && act) ? pv.ld_nt<T>(voP, uint32_t(2 * e * G::N2 * sizeof(cpx<T>))) : cpx<T>{0, 0};
        } else {
            const cpx<T> *src = pslot + int64_t(c_group[ci]) * G::VOL + int64_t(I) * N + (J & ~1);
#pragma unroll
            for (int f = 0; f < E / 2; ++f) {
                cpx2f w = {{0, 0}, {0, 0}};
                if (live && act) w = *reinterpret_cast<const cpx2f *>(src + int64_t(half + 2 * (jo + 2 * f)) * G::N2);
                dst[2 * f] = w.a;
                dst[2 * f + 1] = w.b;
            }
        }
    };
    // FP32 pairs as loaded -> this lane's pencil at planes half + 2e
    auto pair_swap = [&](cpx<T> *v) {
        if constexpr (sizeof(T) == 4) {
#pragma unroll
            for (int f = 0; f < E / 2; ++f) {
                const cpx<T> a = v[2 * f], b = v[2 * f + 1];
                const cpx<T> recv = fq::swap_q2(jo ? a : b);
                v[2 * f] = jo ? recv : a;
                v[2 * f + 1] = jo ? b : recv;
            }
        }
    };
    cpx<double> H[E];
#pragma unroll
    for (int e = 0; e < E; ++e) H[e] = {0, 0};
    // FP32 pencils are prefetched one contribution ahead; FP64 pencils are
    // loaded in place (no register room for both)
    constexpr bool PF = sizeof(T) == 4;
    cpx<T> nxt[PF ? E : 1];
    // compile-time copy loops (constant indices keep wp in registers)
    if (cn > 0) {
        fq::sfor<0, CP>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            const int idx = int(threadIdx.x) + TH * k;
            if (WR % TH == 0 || idx < WR) {
                const cpx<double> *q = wrow + int64_t(c0) * G::VOL + int64_t(idx / N) * G::N2 + idx % N;
                wp[k] = {q->x, q->y};           // field copies: no aggregate memcpy
            }
        });
        if constexpr (PF) load_pencil(c0, nxt);
    }
    for (int ci = c0, it = 0; ci < c0 + cn; ++ci, ++it) {
        const int buf = it & 1;
        // reads of buf ended before the previous barrier
        fq::sfor<0, CP>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            const int idx = int(threadIdx.x) + TH * k;
            if (WR % TH == 0 || idx < WR) {
                wl[buf][idx].x = wp[k].x;
                wl[buf][idx].y = wp[k].y;
            }
        });
        // next contribution's weights: the loads stay in flight across the
        // barrier and the transform below
        const int cn1 = ci + 1 < c0 + cn ? ci + 1 : ci;
        fq::sfor<0, CP>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            const int idx = int(threadIdx.x) + TH * k;
            if (WR % TH == 0 || idx < WR) {
                const cpx<double> *q = wrow + int64_t(cn1) * G::VOL + int64_t(idx / N) * G::N2 + idx % N;
                wp[k] = {q->x, q->y};
            }
        });
        __syncthreads();
        cpx<T> v[E];
        if constexpr (PF) {
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = nxt[e];
            if (ci + 1 < c0 + cn) load_pencil(ci + 1, nxt);
        } else {
            load_pencil(ci, v);
        }
        if (live) {
            pair_swap(v);
            fq::fft2_dit<N, -1>(v, half);             // -> K' = half*E + e
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const cpx<double> we = act ? wl[buf][(half * E + e) * N + J] : cpx<double>{0, 0};
                fq::cmac(H[e], we, cvt<double>(v[e]));
            }
        }
    }
    if (!live) return;                                // no barrier follows
    fq::fft2_dif<N, 1>(H, half);                      // -> vz = half + 2e
    cpx<double> *yg = Y + ((int64_t(slot) * ntgt + t) * 2 + 0) * G::VOL + off;
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) yg[int64_t(half + 2 * e) * G::N2] = H[e];
    }
#pragma unroll
    for (int e = 0; e < E; ++e) H[e] = {0, 0};
    const int l0 = t_l0[t], ln = t_ln[t];
    for (int li = l0; li < l0 + ln; ++li) {
        const int s = l_src[li];
        const cpx<double> *fs = F + (int64_t(slot) * nsrc + s) * G::VOL + off;
        const cpx<double> *w = W + int64_t(ncontribs + li) * G::VOL + off;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int64_t k = int64_t(half * E + e) * G::N2;
            if (act) H[e] = H[e] + fq::cmul(w[k], fs[k]);
        }
    }
    fq::fft2_dif<N, 1>(H, half);
    cpx<double> *yl = Y + ((int64_t(slot) * ntgt + t) * 2 + 1) * G::VOL + off;
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) yl[int64_t(half + 2 * e) * G::N2] = H[e];
    }
}

// ---------------------------------------------------------------------------
// Group-pair accumulation on the 32^3 grid (k_prod1 with HP).
//
// Volume u of a DOF holds Xu = DFT-_{x,y}(a_2u + i a_2u+1), a_g = Re G_g.
// With Z = DFT-_z Xu and xi- = -xi (mod N) the two Hermitian spectra
// A_g = DFT-(a_g) separate as
//   Z(xi) + conj Z(xi-) = 2 A_2u(xi),   Z(xi) - conj Z(xi-) = 2i A_2u+1(xi),
// and conj Z(xi-) along a z pencil is the z transform of the conjugated
// pencil at (-I', -J').  The Hermitian part of the gain spectrum,
//   H(xi) = sum_c U_c A_g(c) + V_c B_g(c)
// (U, V: k_prod1), is itself Hermitian, so only the rows I' = 0..N/2 are
// evaluated and row -I' is their conjugate.  B_g = DFT-(Im G_g) enters only
// on the Nyquist set, where V_c can be non-zero (k_nyqb).
//
// Tables (computed in FP64 from W; X stored in FP64, VN in FP32: its terms
// are the Nyquist-plane part of the gain, a small fraction of it):
//   X [entry][r][2][K'][J']: U_c/2 (m = 0, group 2u) and U_c'/(2i) (m = 1,
//     group 2u + 1) of an entry (volume u, target t) on row I' = r;
//   VN[contrib][kCPer32]:    V_c on the rows r <= N/2 of the three Nyquist
//     planes, each point once: VK[r][J'] (K' = N/2), VJ[r][K'] (J' = N/2,
//     K' != N/2), VI[J'][K'] (I' = N/2, J' and K' != N/2);
//   C [slot][target][kCPer32]: sum_c V_c B_g(c) in the same layout (k_nyqb).
// W is stored [K][I][J] with l = (l0[I], l0[J], l0[K]).
// ---------------------------------------------------------------------------
constexpr int kHRows32 = 17;                                 // rows I' = 0..N/2 (N = 32)
constexpr int kCPer32 = 2 * kHRows32 * 32 + 32 * 32;         // V / C entries per contribution / target

__global__ void k_htab(cpx<double> *X, const cpx<double> *W, const int *ent_c, int N, int R) {
    const int64_t n = int64_t(N) * N * N;
    const int per = R * 2 * N * N;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int e = blockIdx.y;
    if (idx >= per) return;
    const int J = idx % N, K = (idx / N) % N, m = (idx / (N * N)) % 2, r = idx / (2 * N * N);
    const int c = ent_c[2 * e + m];
    cpx<double> out = {0, 0};
    if (c >= 0) {
        const cpx<double> w = W[c * n + (int64_t(K) * N + r) * N + J];
        const cpx<double> wm = W[c * n + (int64_t((N - K) % N) * N + (N - r) % N) * N + (N - J) % N];
        const double ux = 0.5 * (w.x + wm.x), uy = 0.5 * (w.y - wm.y);     // U = (W + conj W(-xi)) / 2
        out = m == 0 ? cpx<double>{0.5 * ux, 0.5 * uy} : cpx<double>{0.5 * uy, -0.5 * ux};
    }
    X[int64_t(e) * per + idx] = out;
}

__global__ void k_vtab(cpx<float> *VN, const cpx<double> *W, int N, int R) {
    const int64_t n = int64_t(N) * N * N;
    const int H = N / 2;
    const int per = 2 * R * N + N * N;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (idx >= per) return;
    int I, J, K;
    bool keep;
    if (idx < R * N) {
        I = idx / N; J = idx % N; K = H; keep = true;
    } else if (idx < 2 * R * N) {
        const int i2 = idx - R * N;
        I = i2 / N; K = i2 % N; J = H; keep = K != H;
    } else {
        const int i3 = idx - 2 * R * N;
        J = i3 / N; K = i3 % N; I = H; keep = J != H && K != H;
    }
    cpx<double> out = {0, 0};
    if (keep) {
        const cpx<double> w = W[c * n + (int64_t(K) * N + I) * N + J];
        const cpx<double> wm = W[c * n + (int64_t((N - K) % N) * N + (N - I) % N) * N + (N - J) % N];
        const double dx = w.x - wm.x, dy = w.y + wm.y;                    // W - conj W(-xi)
        out = {-0.5 * dy, 0.5 * dx};                                       // V = i (W - conj W(-xi)) / 2
    }
    VN[int64_t(c) * per + idx] = {float(out.x), float(out.y)};
}

// ---------------------------------------------------------------------------
// k_nyqb: C = sum_c V_c B_g(c) per (slot, target) from the Nyquist sums NB
// of k_prod1.  On the three planes
//   B(I', J', N/2) = DFT-_{x,y} s_K,  B(I', N/2, K') = DFT-_{x,z} s_J,
//   B(N/2, J', K') = DFT-_{y,z} s_I,
// s_K summed over the four vz-group partials.  s_K and s_J are real, so one
// complex transform of Z = s_K + i s_J gives both: with Z~(p, q) = Z(-p, -q),
// B_K = (Z + conj Z~)/2 and B_J = (Z - conj Z~)/(2i).
// One workgroup per (slot, target), four waves, four contributions per
// round: wave w transforms contribution 4i + w, Z on lanes 0-31 and s_I on
// lanes 32-63 (one plane per half-wave, one pencil per lane read as
// coalesced row segments, in-register 32-point transforms and an LDS
// transpose), leaving both spectra in LDS as [p][q].
// Then every thread accumulates its own output points (k/256 of the 2112)
// over the round's four contributions, with coalesced weight reads, in FP64.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_nyqb(const float *__restrict__ NB, const cpx<float> *__restrict__ VN,
                                              const int *__restrict__ c_group, const int *__restrict__ t_c0,
                                              const int *__restrict__ t_cn, cpx<double> *__restrict__ C,
                                              int ngroups, int ntgt) {
    constexpr int N = 32, R = kHRows32, ROW = N + 1, NPT = kCPer32, PPT = (NPT + 255) / 256;
    __shared__ cpx<float> lds[4][2][N * ROW];          // [wave][Z, s_I spectrum]
    const int slot = int(blockIdx.x) / ntgt, t = int(blockIdx.x) % ntgt;
    const int tid = int(threadIdx.x);
    const int wave = __builtin_amdgcn_readfirstlane(tid / 64);
    const int hw = (tid / 32) & 1;
    const int c = tid % 32;
    cpx<float> *L = lds[wave][hw];
    const int c0 = t_c0[t], cn = t_cn[t];
    cpx<double> acc[PPT];
#pragma unroll
    for (int j = 0; j < PPT; ++j) acc[j] = {0, 0};
    for (int base = 0; base < cn; base += 4) {
        if (base + wave < cn) {                            // wave-uniform
            const float *nb = NB + (int64_t(slot) * ngroups + c_group[c0 + base + wave]) * kNbPer32;
            // lane c takes index c of the first axis and w runs over the second,
            // so every load reads one contiguous row segment across the lanes:
            // Z(vx = c, w) = s_K(vx, vy = w) + i s_J(vx, vz = w), or s_I(vy = c, vz = w)
            cpx<float> z[N];
            if (hw == 0) {
#pragma unroll
                for (int w = 0; w < N; ++w) {
                    // non-temporal: the sums are read once (10% faster)
                    const float sk = (__builtin_nontemporal_load(nb + w * N + c) +
                                      __builtin_nontemporal_load(nb + N * N + w * N + c)) +
                                     (__builtin_nontemporal_load(nb + 2 * N * N + w * N + c) +
                                      __builtin_nontemporal_load(nb + 3 * N * N + w * N + c));
                    z[w] = {sk, __builtin_nontemporal_load(nb + 4 * N * N + w * N + c)};
                }
            } else {
#pragma unroll
                for (int w = 0; w < N; ++w) z[w] = {__builtin_nontemporal_load(nb + 5 * N * N + w * N + c), 0.f};
            }
            fq::LFFT<N, -1>::run(z);                       // along w: z[q]
            plane_sync<32>();
#pragma unroll
            for (int q = 0; q < N; ++q) L[c * ROW + q] = z[q];
            plane_sync<32>();
#pragma unroll
            for (int i = 0; i < N; ++i) z[i] = L[i * ROW + c];
            fq::LFFT<N, -1>::run(z);                       // along the first axis: z[p] = spectrum(p, q = c)
            plane_sync<32>();
#pragma unroll
            for (int p = 0; p < N; ++p) L[p * ROW + c] = z[p];     // [p][q]
        }
        __syncthreads();
        for (int w2 = 0; w2 < 4 && base + w2 < cn; ++w2) {
            const cpx<float> *v = VN + int64_t(c0 + base + w2) * NPT;
            const cpx<float> *Z0 = lds[w2][0], *Z1 = lds[w2][1];
#pragma unroll
            for (int j = 0; j < PPT; ++j) {
                const int pt = tid + 256 * j;
                if (pt < NPT) {
                    cpx<double> bv;
                    if (pt < 2 * R * N) {
                        const int i2 = pt < R * N ? pt : pt - R * N;
                        const int p = i2 / N, q = i2 % N;
                        const cpx<float> z = Z0[p * ROW + q], zp = Z0[((N - p) % N) * ROW + (N - q) % N];
                        bv = pt < R * N ? cpx<double>{0.5 * (double(z.x) + zp.x), 0.5 * (double(z.y) - zp.y)}
                                        : cpx<double>{0.5 * (double(z.y) + zp.y), -0.5 * (double(z.x) - zp.x)};
                    } else {
                        const int i3 = pt - 2 * R * N;
                        bv = cvt<double>(Z1[(i3 / N) * ROW + i3 % N]);
                    }
                    fq::cmac(acc[j], cvt<double>(v[pt]), bv);
                }
            }
        }
        __syncthreads();                                   // LDS reads done before the next round
    }
    cpx<double> *cc = C + (int64_t(slot) * ntgt + t) * NPT;
#pragma unroll
    for (int j = 0; j < PPT; ++j) {
        const int pt = tid + 256 * j;
        if (pt < NPT) cc[pt] = acc[j];
    }
}

// ---------------------------------------------------------------------------
// k_accum_h: the group-pair accumulation.  One workgroup = NS consecutive
// DOF slots of one (target, row pair r, -r), r = 0..N/2, one wave each; lane
// (J, half) holds the z pencil (r, J) of the volume and the conjugated pencil
// (-r, -J).  Per entry (volume u, target t):
//   s = DFT-_z(v1 + v2) = 2 A_2u,  d = DFT-_z(v1 - v2) = 2i A_2u+1,
//   H += X[e][r][0] s + X[e][r][1] d                               (FP64)
// with the entry's weight block X[e][r] (2 N^2 FP64 values, the same for the
// NS slots) staged in double-buffered LDS one entry ahead, as k_accum_lds
// does.  Then C (k_nyqb) is added on the Nyquist points of row r.  The loss
// spectrum is made Hermitian too, Hl = (L(xi) + conj L(xi-))/2 with
// L = sum Lt . F: like the gain it only enters through the real part of its
// inverse transform.  Both are transformed back along z and written as one
// volume Y = DFT+_z(H) + i DFT+_z(Hl) (k_out1); row -r holds
// conj DFT+_z(H) + i conj DFT+_z(Hl).
// ---------------------------------------------------------------------------
template <int NS>
__global__ __launch_bounds__(64 * NS, 8 / NS) void k_accum_h(
    const cpx<float> *__restrict__ P, const cpx<double> *__restrict__ F, const cpx<double> *__restrict__ X,
    const int *__restrict__ e_vol, const int *__restrict__ t_e0, const int *__restrict__ t_en,
    const cpx<double> *__restrict__ C, const cpx<double> *__restrict__ W, const int *__restrict__ l_src,
    const int *__restrict__ t_l0, const int *__restrict__ t_ln, int ncontribs, cpx<double> *__restrict__ Y,
    int nvol, int nsrc, int ntgt, int D) {
    constexpr int N = 32, R = kHRows32;
    using G = Geo<N>;
    constexpr int E = G::E;
    constexpr int WR = 2 * N * N;                 // weight block of one (entry, row)
    constexpr int TH = 64 * NS;
    constexpr int CP = WR / TH;                   // 16-byte copies per thread
    static_assert(WR % TH == 0, "whole copies");
    __shared__ cpx<double> wl[2][WR];
    const int nsg = (D + NS - 1) / NS;
    const int b = blockIdx.x;
    const int sg = b % nsg;
    const int t = (b / nsg) % ntgt;
    const int r = b / (nsg * ntgt);
    const int wave = __builtin_amdgcn_readfirstlane(int(threadIdx.x) / 64);
    const int slot = sg * NS + wave;
    const bool live = slot < D;                   // wave-uniform
    const int lt = threadIdx.x % 64;
    const int J = lt >> 1, half = lt & 1;
    const int I1 = r, I2 = (N - r) % N, Jp = (N - J) % N;
    const int jo = J & 1;
    const int e0 = t_e0[t], en = t_en[t];
    cpx<double> wp[CP];
    const cpx<double> *xr = X + int64_t(r) * WR;
    const cpx<float> *pslot = P + int64_t(live ? slot : 0) * nvol * G::VOL;
    // v1: 16-byte column pairs at the planes half + 2(jo + 2f), completed by
    // the exchange with lane ^ 2 below (as in k_accum_lds); v2: the partner
    // column -J of row -r at the planes half + 2e.  Non-temporal loads: the
    // volumes stream through once, and the weight rows X, which the
    // workgroups of one row share, keep their place in L2 (3% faster)
    auto load = [&](int ei, cpx<float> *v1, cpx<float> *v2) {
        const cpx<float> *vol = pslot + int64_t(e_vol[ei]) * G::VOL;
        const cpx<float> *s1 = vol + int64_t(I1) * N + (J & ~1);
        const cpx<float> *s2 = vol + int64_t(I2) * N + Jp;
#pragma unroll
        for (int f = 0; f < E / 2; ++f) {
            cpx2f w = {{0, 0}, {0, 0}};
            if (live) {
                const nt_f4 t4 = __builtin_nontemporal_load(
                    reinterpret_cast<const nt_f4 *>(s1 + int64_t(half + 2 * (jo + 2 * f)) * G::N2));
                w = {{t4.x, t4.y}, {t4.z, t4.w}};
            }
            v1[2 * f] = w.a;
            v1[2 * f + 1] = w.b;
        }
#pragma unroll
        for (int e = 0; e < E; ++e) {
            cpx<float> t = {0, 0};
            if (live) {
                const nt_f2 t2 = __builtin_nontemporal_load(reinterpret_cast<const nt_f2 *>(s2 + int64_t(half + 2 * e) * G::N2));
                t = {t2.x, t2.y};
            }
            v2[e] = t;
        }
    };
    cpx<double> H[E];
#pragma unroll
    for (int e = 0; e < E; ++e) H[e] = {0, 0};
    cpx<float> n1[E], n2[E];
    if (en > 0) {
        fq::sfor<0, CP>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            const cpx<double> *q = xr + int64_t(e0) * R * WR + int(threadIdx.x) + TH * k;
            wp[k] = {q->x, q->y};                 // field copies: no aggregate memcpy
        });
        load(e0, n1, n2);
    }
    for (int ei = e0, it = 0; ei < e0 + en; ++ei, ++it) {
        const int buf = it & 1;
        fq::sfor<0, CP>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            wl[buf][int(threadIdx.x) + TH * k].x = wp[k].x;
            wl[buf][int(threadIdx.x) + TH * k].y = wp[k].y;
        });
        const int en1 = ei + 1 < e0 + en ? ei + 1 : ei;
        fq::sfor<0, CP>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            const cpx<double> *q = xr + int64_t(en1) * R * WR + int(threadIdx.x) + TH * k;
            wp[k] = {q->x, q->y};
        });
        __syncthreads();
        cpx<float> v1[E], v2[E];
#pragma unroll
        for (int e = 0; e < E; ++e) { v1[e] = n1[e]; v2[e] = n2[e]; }
        if (ei + 1 < e0 + en) load(ei + 1, n1, n2);
        if (live) {
#pragma unroll
            for (int f = 0; f < E / 2; ++f) {
                const cpx<float> a = v1[2 * f], bb = v1[2 * f + 1];
                const cpx<float> recv = fq::swap_q2(jo ? a : bb);
                v1[2 * f] = jo ? recv : a;
                v1[2 * f + 1] = jo ? bb : recv;
            }
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const cpx<float> a = v1[e], q = v2[e];
                v1[e] = {a.x + q.x, a.y - q.y};            // v1 + conj v2
                v2[e] = {a.x - q.x, a.y + q.y};            // v1 - conj v2
            }
            fq::fft2_dit<N, -1>(v1, half);                 // -> K' = half*E + e
            fq::fft2_dit<N, -1>(v2, half);
            const cpx<double> *w0 = wl[buf];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int K = half * E + e;
                fq::cmac(H[e], w0[K * N + J], cvt<double>(v1[e]));
                fq::cmac(H[e], w0[N * N + K * N + J], cvt<double>(v2[e]));
            }
        }
    }
    if (!live) return;                                // no barrier follows
    if (C) {
        const cpx<double> *cc = C + (int64_t(slot) * ntgt + t) * kCPer32;
        if (half == 1) H[0] = H[0] + cc[r * N + J];                      // K' = N/2
        if (J == N / 2) {
#pragma unroll
            for (int e = 0; e < E; ++e) H[e] = H[e] + cc[R * N + r * N + half * E + e];
        }
        if (r == N / 2) {
#pragma unroll
            for (int e = 0; e < E; ++e) H[e] = H[e] + cc[2 * R * N + J * N + half * E + e];
        }
    }
    // Hermitian loss on row r: the point xi- = (-r, -J, -K') of each K'
    cpx<double> Hl[E];
#pragma unroll
    for (int e = 0; e < E; ++e) Hl[e] = {0, 0};
    const int l0 = t_l0[t], ln = t_ln[t];
    const int64_t off1 = int64_t(I1) * N + J, off2 = int64_t(I2) * N + Jp;
    for (int li = l0; li < l0 + ln; ++li) {
        const int s = l_src[li];
        const cpx<double> *fs = F + (int64_t(slot) * nsrc + s) * G::VOL;
        const cpx<double> *w = W + int64_t(ncontribs + li) * G::VOL;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int K = half * E + e, Km = (N - K) % N;
            const cpx<double> a = fq::cmul(w[int64_t(K) * G::N2 + off1], fs[int64_t(K) * G::N2 + off1]);
            const cpx<double> m = fq::cmul(w[int64_t(Km) * G::N2 + off2], fs[int64_t(Km) * G::N2 + off2]);
            Hl[e] = {Hl[e].x + 0.5 * (a.x + m.x), Hl[e].y + 0.5 * (a.y - m.y)};
        }
    }
    fq::fft2_dif<N, 1>(H, half);                      // -> vz = half + 2e
    fq::fft2_dif<N, 1>(Hl, half);
    // non-temporal stores, as the volume loads (8% faster accumulation)
    cpx<double> *y = Y + (int64_t(slot) * ntgt + t) * G::VOL;
#pragma unroll
    for (int e = 0; e < E; ++e)
        __builtin_nontemporal_store(nt_d2{H[e].x - Hl[e].y, H[e].y + Hl[e].x},
                                    reinterpret_cast<nt_d2 *>(y + int64_t(half + 2 * e) * G::N2 + off1));
    if (I2 != I1) {
#pragma unroll
        for (int e = 0; e < E; ++e)
            __builtin_nontemporal_store(nt_d2{H[e].x + Hl[e].y, Hl[e].x - H[e].y},
                                        reinterpret_cast<nt_d2 *>(y + int64_t(half + 2 * e) * G::N2 + off2));
    }
}

// ---------------------------------------------------------------------------
// k_out: unit = (slot, target, vz).  gain = DFT+_{x,y}(Yg), loss = DFT+_{x,y}(Yl)
// Q[v] = Re gain - f[v] Re loss, scattered into the AoSoA matrix
// (output kernel, vhs-gll.mako:246-266; vss-gll.mako:275-300).
// ---------------------------------------------------------------------------
template <int N>
__device__ __forceinline__ void plane_inv64(const cpx<double> *src, cpx<double> *L, cpx<double> *v, int c,
                                            int half, bool act) {
    using G = Geo<N>;
    constexpr int E = G::E;
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = act ? src[(half * E + e) * N + c] : cpx<double>{0, 0};
    fq::fft2_dif<N, 1>(v, half);
    __syncthreads();
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) L[(half + 2 * e) * G::LDSROW + c] = v[e];
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < E; ++e) v[e] = act ? L[c * G::LDSROW + half * E + e] : cpx<double>{0, 0};
    fq::fft2_dif<N, 1>(v, half);
}

template <int N>
__global__ __launch_bounds__(kPlaneThreads, 2) void k_out(const cpx<double> *Y, AosoaPtrs ptr, Dof dof, int ntgt,
                                                int64_t nunits) {
    using G = Geo<N>;
    constexpr int E = G::E;
    constexpr int UPB = kPlaneThreads / G::LPP;
    __shared__ cpx<double> lds[UPB][N * G::LDSROW];
    const int u = threadIdx.x / G::LPP;
    const int lt = threadIdx.x % G::LPP;
    const int64_t unit = int64_t(blockIdx.x) * UPB + u;
    const bool valid = unit < nunits;
    const int vz = int(unit % N);
    const int64_t dt = unit / N;
    const int t = int(dt % ntgt);
    const int slot = int(dt / ntgt);
    const int c = lt >> 1, half = lt & 1;
    const bool act = valid && c < N;
    cpx<double> *L = lds[u];
    const cpx<double> *yg = Y + ((int64_t(slot) * ntgt + t) * 2 + 0) * G::VOL + int64_t(vz) * G::N2;
    // only the real parts of the gain and loss transforms are used
    double gr[E];
    {
        cpx<double> gv[E];
        plane_inv64<N>(yg, L, gv, c, half, act);
#pragma unroll
        for (int e = 0; e < E; ++e) gr[e] = gv[e].x;
    }
    cpx<double> lv[E];
    plane_inv64<N>(yg + G::VOL, L, lv, c, half, act);
    if (act) {
        const double *f = ptr.f[t];
        double *q = ptr.q[t];
        const int64_t b = dof.base(slot);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int vy = half + 2 * e;
            const int64_t vv = (int64_t(c) * N + vy) * N + vz;
            const int64_t a = b + vv * dof.soasz;
            q[a] = gr[e] - f[a] * lv[e].x;
        }
    }
}

// ---------------------------------------------------------------------------
// k_out1: the output stage of the group-pair path.  Y = Yg + i Yl with the
// gain and loss spectra Hermitian, so DFT+_{x,y}(Y) = gain + i loss with
// both real: one complex transform per plane, Q = Re - f Im.  A workgroup
// transforms one plane of four consecutive batch slots (one wave each), stages
// the results in LDS as [point][slot], and reads f and writes Q with the four
// slots on adjacent lanes, so that consecutive elements of one solution point
// meet in one AoSoA row segment.  With MOM it also writes each slot's plane
// sums of Q {1, c, |c|^2} (FP64) to mom[slot][t][vz][5] for k_project_b.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

template <int N, bool MOM>
__global__ __launch_bounds__(4 * Geo<N>::LPP, 2) void k_out1(const cpx<double> *Y, AosoaPtrs ptr, Dof dof, int ntgt,
                                                   int D, double *mom, double L) {
    using G = Geo<N>;
    static_assert(G::LPP == 64, "one plane per wave");
    constexpr int E = G::E, U = 4;                 // batch slots per workgroup, one wave each
    constexpr int TH = U * G::LPP;
    constexpr int PPT = G::N2 / (TH / U);          // grid points per thread in the write phase
    __shared__ cpx<double> lds[U][N * G::LDSROW];
    static_assert(U * G::N2 <= U * N * G::LDSROW, "staging fits the transpose buffers");
    const int tid = int(threadIdx.x);
    const int u = tid / G::LPP, lt = tid % G::LPP;
    const int b = int(blockIdx.x);
    const int vz = b % N, t = (b / N) % ntgt, sg = b / (N * ntgt);
    const int slot = sg * U + u;
    const int c = lt >> 1, half = lt & 1;
    const bool act = slot < D && c < N;
    cpx<double> v[E];
    plane_inv64<N>(Y + (int64_t(slot < D ? slot : 0) * ntgt + t) * G::VOL + int64_t(vz) * G::N2, lds[u], v, c,
                   half, act);
    // stage (gain, loss) as [point][slot], point = vx * N + vy
    __syncthreads();
    cpx<double> *st = &lds[0][0];
    if (act) {
#pragma unroll
        for (int e = 0; e < E; ++e) st[(c * N + half + 2 * e) * U + u] = v[e];
    }
    __syncthreads();
    // thread (slot u2, points p): the U slots of one point are adjacent AoSoA
    // elements when they are consecutive elements of one solution point
    const int u2 = tid % U, slot2 = sg * U + u2;
    double m[5] = {0, 0, 0, 0, 0};
    if (slot2 < D) {
        const double *f = ptr.f[t];
        double *q = ptr.q[t];
        const int64_t b2 = dof.base(slot2);
        const double h = 2.0 * L / N, c0 = -L + L / N, cz = c0 + h * vz;
#pragma unroll 4
        for (int i = 0; i < PPT; ++i) {
            const int pnt = tid / U + (TH / U) * i;
            const int vx = pnt / N, vy = pnt % N;
            const int64_t a = b2 + ((int64_t(vx) * N + vy) * N + vz) * dof.soasz;
            const cpx<double> g = st[pnt * U + u2];
            const double qv = g.x - f[a] * g.y;
            q[a] = qv;
            if constexpr (MOM) {
                const double cx = c0 + h * vx, cy = c0 + h * vy;
                m[0] += qv;
                m[1] += qv * cx;
                m[2] += qv * cy;
                m[3] += qv * cz;
                m[4] += qv * (cx * cx + cy * cy + cz * cz);
            }
        }
    }
    if constexpr (MOM) {
        // sums of the threads of one slot: lanes u2 + 4j of each wave, then the U waves
#pragma unroll
        for (int k = 0; k < 5; ++k) {
#pragma unroll
            for (int o = U; o < 64; o <<= 1) m[k] += __shfl_xor(m[k], o);
        }
        __syncthreads();                             // staging reads done
        double *red = reinterpret_cast<double *>(&lds[0][0]);       // [wave][u2][5]
        if (lt < U) {
#pragma unroll
            for (int k = 0; k < 5; ++k) red[(u * U + lt) * 5 + k] = m[k];
        }
        __syncthreads();
        if (tid < U * 5) {
            const int uu = tid / 5, k = tid % 5, sl = sg * U + uu;
            if (sl < D) {
                double acc = 0;
#pragma unroll
                for (int w = 0; w < U; ++w) acc += red[(w * U + uu) * 5 + k];
                mom[((int64_t(sl) * ntgt + t) * N + vz) * 5 + k] = acc;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Conservation projection (FP64).  With S species of mass ratios m_s the
// collision invariants are the mass of every species and the total momentum
// and energy, sum_s m_s {v, |v|^2} Q_s (momentum and energy move between
// species; invariants of frfs/plugins/dgfsmomwriterstd.py:142-214).  The
// constraint rows a = 0..S+3 are
//   phi_s(s', v) = [s = s'],  phi_{S+d}(s', v) = m_s' v_d,  phi_{S+3}(s', v) = m_s' |v|^2,
// and Q <- Q - sum_a lambda_a phi_a with lambda = G^{-1} (A Q),
// G = A A^T: the least-squares correction that zeroes every invariant.  For
// one species this is the projection onto the complement of {1, v, |v|^2}.
// One block per slot, all species of the slot together.
// ---------------------------------------------------------------------------
constexpr int kProjMaxRows = 4 + 4;            // up to 4 species

__device__ __forceinline__ void phis(int64_t v, int N, double L, double *ph) {
    const int I = int(v / (int64_t(N) * N)), J = int((v / N) % N), K = int(v % N);
    const double h = 2.0 * L / N;
    const double cx = -L + L / N + h * I, cy = -L + L / N + h * J, cz = -L + L / N + h * K;
    ph[0] = 1; ph[1] = cx; ph[2] = cy; ph[3] = cz; ph[4] = cx * cx + cy * cy + cz * cz;
}

// One workgroup per DOF.  Workgroups are dealt round-robin to the eight XCDs,
// so workgroup w takes slot (w % 8) * per + w / 8: each XCD works through a
// contiguous DOF range, and the DOFs whose values share an AoSoA cache line
// (neighbouring elements) meet in the same L2.
__global__ __launch_bounds__(256) void k_project(AosoaPtrs ptr, Dof dof, int ns, int N, double L,
                                                 const double *ginv, const double *mass, int D) {
    const int per = (D + 7) / 8;
    const int slot = int(blockIdx.x & 7) * per + int(blockIdx.x >> 3);
    if (slot >= D) return;
    const int64_t b = dof.base(slot);
    const int64_t n = int64_t(N) * N * N;
    const int K = ns + 4;
    __shared__ double red[5][256];
    __shared__ double aq[kProjMaxRows];
    if (threadIdx.x < kProjMaxRows) aq[threadIdx.x] = 0;
    __syncthreads();
    for (int s = 0; s < ns; ++s) {
        const double *q = ptr.q[s];
        double m[5] = {0, 0, 0, 0, 0};
        for (int64_t v = threadIdx.x; v < n; v += blockDim.x) {
            double ph[5];
            phis(v, N, L, ph);
            const double qv = q[b + v * dof.soasz];
#pragma unroll
            for (int k = 0; k < 5; ++k) m[k] += qv * ph[k];
        }
#pragma unroll
        for (int k = 0; k < 5; ++k) red[k][threadIdx.x] = m[k];
        __syncthreads();
        for (int st = blockDim.x / 2; st > 0; st >>= 1) {
            if (threadIdx.x < st)
#pragma unroll
                for (int k = 0; k < 5; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            aq[s] = red[0][0];
            for (int k = 1; k < 5; ++k) aq[ns + k - 1] += mass[s] * red[k][0];
        }
        __syncthreads();
    }
    double lam[kProjMaxRows];
    for (int i = 0; i < K; ++i) {
        double acc = 0;
        for (int k = 0; k < K; ++k) acc += ginv[i * K + k] * aq[k];
        lam[i] = acc;
    }
    for (int s = 0; s < ns; ++s) {
        double *q = ptr.q[s];
        const double ms = mass[s];
        for (int64_t v = threadIdx.x; v < n; v += blockDim.x) {
            double ph[5];
            phis(v, N, L, ph);
            const double corr = lam[s] + ms * (lam[ns] * ph[1] + lam[ns + 1] * ph[2] + lam[ns + 2] * ph[3] +
                                               lam[ns + 3] * ph[4]);
            q[b + v * dof.soasz] -= corr;
        }
    }
}

// The same projection with the moments of Q taken from the per-plane sums
// that k_out1 wrote (mom[slot][s][vz][5]), so Q is read once and written once,
// over blocks of 32 consecutive batch slots and chunks of kProjChunk grid
// points: lane e of each 32-lane group takes slot s0 + e, so with one solution
// point per element and 32-aligned DOFs the group's accesses are one
// contiguous AoSoA row segment (32 elements x 8 bytes) per grid point.
constexpr int kProjChunk = 1024;

__global__ __launch_bounds__(256) void k_project_b(AosoaPtrs ptr, Dof dof, int ns, int N, double L,
                                                   const double *ginv, const double *mass, const double *mom, int D) {
    const int64_t n = int64_t(N) * N * N;
    const int nchunk = int(n / kProjChunk);
    const int blk = int(blockIdx.x) / nchunk, ch = int(blockIdx.x) % nchunk;
    const int s0 = 32 * blk;
    const int K = ns + 4;
    __shared__ double lam_s[32][kProjMaxRows];
    const int tid = int(threadIdx.x);
    if (tid < 32) {
        const int slot = s0 + tid;
        double aq[kProjMaxRows];
#pragma unroll
        for (int i = 0; i < kProjMaxRows; ++i) aq[i] = 0;
        if (slot < D) {
            for (int s = 0; s < ns; ++s) {
                double m[5] = {0, 0, 0, 0, 0};
                const double *ms = mom + (int64_t(slot) * ns + s) * N * 5;
                for (int z = 0; z < N; ++z) {
#pragma unroll
                    for (int k = 0; k < 5; ++k) m[k] += ms[z * 5 + k];
                }
                aq[s] = m[0];
                for (int k = 1; k < 5; ++k) aq[ns + k - 1] += mass[s] * m[k];
            }
        }
        for (int i = 0; i < K; ++i) {
            double acc = 0;
            for (int k = 0; k < K; ++k) acc += ginv[i * K + k] * aq[k];
            lam_s[tid][i] = acc;
        }
    }
    __syncthreads();
    const int e = tid % 32, vg = tid / 32;
    const int slot = s0 + e;
    if (slot >= D) return;                        // no barrier follows
    const int64_t b = dof.base(slot);
    const double *lam = lam_s[e];
    for (int s = 0; s < ns; ++s) {
        double *q = ptr.q[s];
        const double ms = mass[s];
        for (int i = vg; i < kProjChunk; i += 8) {
            const int64_t v = int64_t(ch) * kProjChunk + i;
            double ph[5];
            phis(v, N, L, ph);
            const double corr = lam[s] + ms * (lam[ns] * ph[1] + lam[ns + 1] * ph[2] + lam[ns + 2] * ph[3] +
                                               lam[ns + 3] * ph[4]);
            q[b + v * dof.soasz] -= corr;
        }
    }
}

// ---------------------------------------------------------------------------
// Grids without native kernels (any even N other than 8, 16, 24 and 32).
// The batched 3D transforms run in the ROCm FFT library (hipFFT plans, the
// counterpart of the reference's cuFFT plans, frfs/solvers/dgfs/scattering.py:
// 141-150 and frfs/solvers/dgfs/cufft.py:231-281), and the pointwise stages of
// the plan in the kernels below:
//   F = DFT-(f)/n;  per product A = phi_1 F, B = phi_2 F, both DFT+;
//   G_g = sum coef A B, DFT-;  Y_t[0] += W_c G_g;  Y_t[1] += W_l F;
//   Y DFT+;  Q_t = Re Y_t[0] - f_t Re Y_t[1].
// Volumes are [slot][.][K][I][J] (velocity (I, J, K), wavenumbers
// (l0[I], l0[J], l0[K])) as everywhere in this file.  A batch of D DOFs sits
// in scratch sized for `chunk` slots; the library plans are made once for
// `chunk`, and the slots D..chunk-1 they also transform are independent of
// the others and never read back.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void gen_kij(int64_t idx, int N, int &K, int &I, int &J) {
    const int64_t n2 = int64_t(N) * N;
    K = int(idx / n2);
    I = int((idx / N) % N);
    J = int(idx % N);
}

// F[slot][s] = f_s / n (real) for the batch's slots, zero for the others.
__global__ __launch_bounds__(256) void k_gen_gather(AosoaPtrs ptr, Dof dof, cpx<double> *F, int nsrc, int D,
                                                    int N, double inv_n, int64_t total) {
    const int64_t n = int64_t(N) * N * N;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t idx = i % n, ds = i / n;
        const int s = int(ds % nsrc), slot = int(ds / nsrc);
        double v = 0.0;
        if (slot < D) {
            int K, I, J;
            gen_kij(idx, N, K, I, J);
            v = ptr.f[s][dof.base(slot) + ((int64_t(I) * N + J) * N + K) * dof.soasz] * inv_n;
        }
        F[i] = {v, 0.0};
    }
}

// A = phi(theta_i1) F_s1 and B = phi(theta_i2) F_s2 (B = A + bofs) with the
// separable phases ph[(i * 3 + d) * N + k] of the axes d = x (I), y (J), z (K).
template <typename T>
__global__ __launch_bounds__(256) void k_gen_phase(const cpx<double> *F, const cpx<T> *ph, int i1, int s1,
                                                   int i2, int s2, int nsrc, cpx<T> *A, int64_t bofs, int N,
                                                   int64_t total) {
    const int64_t n = int64_t(N) * N * N;
    const cpx<T> *p1 = ph + int64_t(i1) * 3 * N, *p2 = ph + int64_t(i2) * 3 * N;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t idx = i % n, slot = i / n;
        int K, I, J;
        gen_kij(idx, N, K, I, J);
        const cpx<T> f1 = cvt<T>(F[(slot * nsrc + s1) * n + idx]), f2 = cvt<T>(F[(slot * nsrc + s2) * n + idx]);
        A[i] = fq::cmul(f1, fq::cmul(fq::cmul(p1[I], p1[N + J]), p1[2 * N + K]));
        A[bofs + i] = fq::cmul(f2, fq::cmul(fq::cmul(p2[I], p2[N + J]), p2[2 * N + K]));
    }
}

// G = coef A B (first product of the group) or G += coef A B.
template <typename T>
__global__ __launch_bounds__(256) void k_gen_prod(const cpx<T> *A, int64_t bofs, cpx<T> *G, T coef, int first,
                                                  int64_t total) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += int64_t(gridDim.x) * blockDim.x) {
        const cpx<T> p = fq::cscale(fq::cmul(A[i], A[bofs + i]), coef);
        G[i] = first ? p : G[i] + p;
    }
}

// Y[slot * ystride + idx] += W[idx] X[slot * xstride + idx]  (FP64)
template <typename T>
__global__ __launch_bounds__(256) void k_gen_wmac(const cpx<T> *X, int64_t xstride, const cpx<double> *W,
                                                  cpx<double> *Y, int64_t ystride, int N, int64_t total) {
    const int64_t n = int64_t(N) * N * N;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t idx = i % n, slot = i / n;
        cpx<double> &y = Y[slot * ystride + idx];
        y = y + fq::cmul(W[idx], cvt<double>(X[slot * xstride + idx]));
    }
}

// zero fill (kept out of hipMemsetAsync so that a captured graph replays it)
__global__ __launch_bounds__(256) void k_gen_zero(cpx<double> *Y, int64_t total) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += int64_t(gridDim.x) * blockDim.x)
        Y[i] = {0.0, 0.0};
}

// Q_t = Re DFT+(H_t) - f_t Re DFT+(L_t) from Y[slot][t][0 = gain, 1 = loss].
__global__ __launch_bounds__(256) void k_gen_out(const cpx<double> *Y, AosoaPtrs ptr, Dof dof, int ntgt, int N,
                                                 int64_t total) {
    const int64_t n = int64_t(N) * N * N;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += int64_t(gridDim.x) * blockDim.x) {
        const int64_t idx = i % n, dt = i / n;
        const int t = int(dt % ntgt), slot = int(dt / ntgt);
        int K, I, J;
        gen_kij(idx, N, K, I, J);
        const int64_t a = dof.base(slot) + ((int64_t(I) * N + J) * N + K) * dof.soasz;
        const cpx<double> *y = Y + (int64_t(slot) * ntgt + t) * 2 * n + idx;
        ptr.q[t][a] = y[0].x - ptr.f[t][a] * y[n].x;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
template <typename X> int upload(X **dst, const std::vector<X> &v) {
    *dst = nullptr;
    if (v.empty()) return FQ_SUCCESS;
    HIPCHK(hipMalloc(dst, v.size() * sizeof(X)));
    HIPCHK(hipMemcpy(*dst, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice));
    return FQ_SUCCESS;
}

// grids with native transform kernels; other even N up to kGenMaxN take the
// library-transform path
bool n_supported(int N) { return N == 8 || N == 16 || N == 24 || N == 32; }
constexpr int kGenMaxN = 128;

}  // namespace

struct fq_plan {
    fq::Plan P;
};

// Device scratch of one internal batch: the forward spectra F (FP64) and
// their FP32 copy, Z (staged inverse transforms, also T0 and Y), the
// product volumes Pg and the Nyquist-plane data NY/FJ of the pair-packed
// product kernels.  Every kernel writes what it later reads, so the memory
// needs no clearing between calls.
struct Scratch {
    cpx<double> *F = nullptr;
    void *Z = nullptr;
    void *Pg = nullptr;
    cpx<float> *F32 = nullptr;
    void *NY = nullptr;
    void *FJ = nullptr;
    float *NB = nullptr;          // group pairs: Nyquist sums of Im G_g (k_prod1 -> k_nyqb)
    cpx<double> *C = nullptr;     // group pairs: Nyquist corrections (k_nyqb -> k_accum_h)
    double *MOM = nullptr;        // group pairs with projection: per-plane sums of Q (k_out1 -> k_project_b)
};

struct fq_handle {
    fq::Plan P;
    int precision = 0;
    int flags = 0;
    int device = 0;
    int64_t chunk = 0;
    // device tables
    void *ph = nullptr;          // [items][3][N] cpx<T>
    int *item_src = nullptr;
    DevProd *prods = nullptr;
    int *g_p0 = nullptr, *g_np = nullptr;
    int *c_group = nullptr, *t_c0 = nullptr, *t_cn = nullptr;
    int *l_src = nullptr, *t_l0 = nullptr, *t_ln = nullptr;
    cpx<double> *W = nullptr;    // [contribs + loss][N^3]
    double *ginv = nullptr;      // [(ns+4)^2] inverse Gram matrix of the conservation rows
    double *pmass = nullptr;     // [ns] species mass ratios used by the projection
    void *zc = nullptr;          // [items][4][4 + N/4] k_prod z-stage coefficients in T
    bool fused = true;           // k_prod path (else k_inv_z + k_plane)
    bool two_lane = false;       // FP32 24^3/32^3: two-lane k_prod instead of one-lane k_prod1
    bool pack = false;           // k_prod1 with pair packing (k_nyq corrections in NY)
    void *NY = nullptr;          // [chunk][products][3][N^2] cpx<T> Nyquist-plane corrections
    void *FJ = nullptr;          // [chunk][sources][N^2] cpx<T> spectrum slice J = N/2 for k_nyq
    void *nyv = nullptr;         // [products][2][3][N][N] cpx<T> k_nyq weights
    bool pack_conj = false;      // every product: one source, conjugate phases (zstage_conj)
    int *pconj = nullptr;        // general FP32 plans: per product, 1 = conjugate pair (zstage_conj)
    void *zq = nullptr;          // [products][4][3 + 2 N/4] cpx<T> zstage_conj coefficients
    cpx<float> *F32 = nullptr;   // FP32 copy of F (mixed, fused)
    // group pairs (32^3 FP32, single-product groups): k_prod1 transforms two
    // groups' real parts as one volume, k_accum_h accumulates the Hermitian
    // part of the gain spectrum on rows 0..N/2, k_nyqb the Nyquist terms
    bool hpair = false;
    bool hp_nyq = false;         // the plan's V tables are not all zero (VSS)
    int nvol = 0;                // product volumes per DOF: ngroups, or ceil(ngroups / 2)
    cpx<double> *X = nullptr;     // [entries][17][2][N][N]
    cpx<float> *VN = nullptr;    // [contribs][kCPer32]
    int *e_vol = nullptr, *t_e0 = nullptr, *t_en = nullptr;
    float *NB = nullptr;         // scratch [chunk][groups][kNbPer32]
    cpx<double> *Cn = nullptr;   // scratch [chunk][targets][kCPer32]
    double *MOM = nullptr;       // scratch [chunk][targets][N][5]
    // grids without native kernels: hipFFT plans for `chunk` slots (F,
    // the items A and B, the group volume G, the gain and loss Y) sharing one
    // work area, and the contributions of each group
    bool generic = false;
    hipfftHandle gp_src = nullptr, gp_items = nullptr, gp_group = nullptr, gp_out = nullptr;
    void *gen_work = nullptr;
    std::vector<std::vector<int>> g_contribs;
    // scratch
    cpx<double> *F = nullptr;
    void *Z = nullptr;           // also T0 and Y
    void *Pg = nullptr;
    size_t scratch = 0;
    // FQ_FLAG_OVERLAP: two scratch sets; odd batches on a second stream
    int nsets = 1;
    hipStream_t st2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // per-DOF bytes of each scratch region (for carving the static array of
    // fq_eval_noalloc) and that array's device address on this handle's device
    size_t per_dof_F = 0, per_dof_Z = 0, per_dof_P = 0, per_dof_F32 = 0, per_dof_NY = 0, per_dof_FJ = 0;
    size_t per_dof_NB = 0, per_dof_C = 0, per_dof_M = 0;
    unsigned char *noalloc = nullptr;
    int cus = 256;               // compute units of the device
    // optional per-kernel HIP-event timing
    bool prof = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_rec;
};

static const char *kKernelNames[FQ_NKERNELS] = {"k_fwd_z", "k_fwd_xy", "k_inv_z", "k_plane", "k_accum",
                                                "k_out",   "k_project", "k_prod",  "k_nyq",   "k_nyqb"};

namespace {

template <typename T> int upload_phases(fq_handle *h) {
    const fq::Plan &P = h->P;
    const int N = P.N;
    std::vector<cpx<T>> ph(P.items.size() * 3 * N);
    for (size_t i = 0; i < P.items.size(); ++i)
        for (int d = 0; d < 3; ++d)
            for (int k = 0; k < N; ++k) {
                const double l = double(k < N / 2 ? k : k - N);
                const double a = P.items[i].th[d] * l;
                ph[(i * 3 + d) * N + k] = {T(std::cos(a)), T(std::sin(a))};
            }
    cpx<T> *d = nullptr;
    int rc = upload(&d, ph);
    h->ph = d;
    // z-stage coefficients of k_prod: for item i and vz-group G (N = 4S),
    // d_b = e^{2 pi i bG/4} e^{i th S b'} (b' = b, or b - 4 for b >= 2) and
    // c_a = e^{2 pi i aG/N} e^{i th a}
    const int S = N / 4, ZC = 4 + S;
    std::vector<cpx<T>> zc(P.items.size() * 4 * ZC);
    for (size_t i = 0; i < P.items.size(); ++i) {
        const double th = P.items[i].th[2];
        for (int G = 0; G < 4; ++G) {
            cpx<T> *z = zc.data() + (i * 4 + G) * ZC;
            for (int b = 0; b < 4; ++b) {
                const double bp = b < 2 ? b : b - 4;
                const double ang = 2 * M_PI * ((b * G) % 4) / 4.0 + th * S * bp;
                z[b] = {T(std::cos(ang)), T(std::sin(ang))};
            }
            for (int a = 0; a < S; ++a) {
                const double ang = 2 * M_PI * double((a * G) % N) / N + th * a;
                z[4 + a] = {T(std::cos(ang)), T(std::sin(ang))};
            }
        }
    }
    cpx<T> *dz = nullptr;
    if (rc == FQ_SUCCESS) rc = upload(&dz, zc);
    h->zc = dz;
    return rc;
}

// Weights of k_nyq for general pairs of the pair-packed product kernel (for
// conjugate pairs k_nyq computes them from the phase tables): for product p and
// item t (t = 0: i1, t = 1: i2) on plane piece q at (row a, lane c),
//   V = i^t e^{i theta.l0(I,J,K)} (1 - e^{i N theta.nyq(I,J,K)}) / 2,
// with (I,J,K) the piece's coordinates (k_nyq) and zero where the piece
// excludes the point; nyq = the components at the unpaired index N/2.
template <typename T> int upload_nyq(fq_handle *h) {
    const fq::Plan &P = h->P;
    const int N = P.N, H = N / 2;
    const size_t N2 = size_t(N) * N;
    std::vector<int> pc(P.prods.size());
    bool conj = true, anyconj = false;
    for (size_t p = 0; p < P.prods.size(); ++p) {
        const auto &pr = P.prods[p];
        bool c = P.items[pr.i1].src == P.items[pr.i2].src;
        for (int d = 0; d < 3; ++d) c = c && P.items[pr.i2].th[d] == -P.items[pr.i1].th[d];
        pc[p] = c ? 1 : 0;
        conj = conj && c;
        anyconj = anyconj || c;
    }
    h->pack_conj = conj;
    // the FP64 (two-lane) product kernel packs conjugate pairs only
    if (sizeof(T) == 8 && !conj) {
        h->pack = false;
        return FQ_SUCCESS;
    }
    int rc = FQ_SUCCESS;
    if (!conj) {
    // general pairs: k_nyq reads these weights (conjugate pairs compute theirs)
    std::vector<cpx<T>> v(P.prods.size() * 6 * N2);
    auto l0 = [&](int k) { return double(k < H ? k : k - N); };
    for (size_t p = 0; p < P.prods.size(); ++p) {
        const int items[2] = {P.prods[p].i1, P.prods[p].i2};
        for (int t = 0; t < 2; ++t) {
            const double *th = P.items[items[t]].th;
            for (int q = 0; q < 3; ++q)
                for (int a = 0; a < N; ++a)
                    for (int c = 0; c < N; ++c) {
                        const int I = q == 2 ? H : a, J = q == 1 ? H : c, K = q == 0 ? H : (q == 1 ? c : a);
                        const bool keep = q == 0 || (q == 1 ? c != H : (a != H && c != H));
                        double re = 0, im = 0;
                        if (keep) {
                            const double ang = th[0] * l0(I) + th[1] * l0(J) + th[2] * l0(K);
                            const double nq = N * ((I == H ? th[0] : 0.) + (J == H ? th[1] : 0.) + (K == H ? th[2] : 0.));
                            // e^{i ang} (1 - e^{i nq}) / 2
                            const double hr = 0.5 * (1 - std::cos(nq)), hi = -0.5 * std::sin(nq);
                            re = std::cos(ang) * hr - std::sin(ang) * hi;
                            im = std::cos(ang) * hi + std::sin(ang) * hr;
                            if (t == 1) { const double r = re; re = -im; im = r; }
                        }
                        v[((p * 2 + t) * 3 + q) * N2 + size_t(a) * N + c] = {T(re), T(im)};
                    }
        }
    }
    cpx<T> *dv = nullptr;
    rc = upload(&dv, v);
    h->nyv = dv;
    }
    if (rc != FQ_SUCCESS || !anyconj) return rc;
    if (!conj) {
        // general FP32 plan with some conjugate products: mark them for zstage_conj
        int *dp = nullptr;
        rc = upload(&dp, pc);
        h->pconj = dp;
        if (rc != FQ_SUCCESS) return rc;
    }
    // zstage_conj coefficients of item 1 of every product, per vz-group G:
    // beta_b = e^{i th S b'} (b' = b, or b - 4 for b >= 2) times the sign of
    // w_4^{bG} (its factor i, for odd bG, is applied in the kernel);
    // alpha_a = e^{i th a}; the twiddle w_N^{aG}
    const int S = N / 4, ZQ = 3 + 2 * S;
    std::vector<cpx<T>> zq(P.prods.size() * 4 * ZQ);
    for (size_t p = 0; p < P.prods.size(); ++p) {
        const double th = P.items[P.prods[p].i1].th[2];
        for (int G = 0; G < 4; ++G) {
            cpx<T> *z = zq.data() + (p * 4 + G) * ZQ;
            for (int b = 1; b < 4; ++b) {
                const int q4 = (b * G) % 4;               // w_4^{bG} = i^q4
                const double sg = (q4 == 0 || q4 == 1) ? 1. : -1.;
                const double bp = b < 2 ? b : b - 4;
                z[b - 1] = {T(sg * std::cos(th * S * bp)), T(sg * std::sin(th * S * bp))};
            }
            for (int a = 0; a < S; ++a) {
                const double ca = std::cos(th * a), sa = std::sin(th * a);
                z[3 + a] = {T(ca + sa), T(sa - ca)};                    // (1 - i) alpha_a
                const double tw = 2 * M_PI * double((a * G) % N) / N;
                const double cr = std::cos(tw), ci = std::sin(tw);
                z[3 + S + a] = {T(cr - ci), T(cr + ci)};                // (1 + i) w_N^{aG}
            }
        }
    }
    cpx<T> *dz = nullptr;
    rc = upload(&dz, zq);
    h->zq = dz;
    return rc;
}

int build_tables(fq_handle *h) {
    const fq::Plan &P = h->P;
    const int N = P.N;
    const int64_t n = int64_t(N) * N * N;
    const int ntab = P.n_tables();
    std::vector<DevTerm> terms;
    std::vector<int> tbeg(ntab + 1, 0);
    for (auto &t : P.terms) tbeg[t.table + 1]++;
    for (int i = 0; i < ntab; ++i) tbeg[i + 1] += tbeg[i];
    terms.resize(P.terms.size());
    std::vector<int> fill(ntab, 0);
    for (auto &t : P.terms) terms[tbeg[t.table] + fill[t.table]++] = {t.kind, t.x, t.y, t.pair, t.A, t.B, t.scale};
    DevTerm *d_terms = nullptr;
    int *d_tbeg = nullptr;
    double *d_qz = nullptr, *d_qw = nullptr, *d_sz = nullptr, *d_szpre = nullptr, *d_m = nullptr,
           *d_eta = nullptr, *d_gam = nullptr;
    int rc = FQ_SUCCESS;
    auto chk = [&](int r) { if (rc == FQ_SUCCESS) rc = r; };
    chk(upload(&d_terms, terms));
    chk(upload(&d_tbeg, tbeg));
    chk(upload(&d_qz, P.qz));
    chk(upload(&d_qw, P.qw));
    chk(upload(&d_sz, P.sz));
    std::vector<double> one = {1.0};
    chk(upload(&d_szpre, P.szpre.empty() ? one : P.szpre));
    chk(upload(&d_m, P.masses.empty() ? one : P.masses));
    chk(upload(&d_eta, P.eta.empty() ? one : P.eta));
    chk(upload(&d_gam, P.gam.empty() ? one : P.gam));
    if (rc == FQ_SUCCESS) {
        hipError_t e = hipMalloc(&h->W, size_t(ntab) * n * sizeof(cpx<double>));
        if (e != hipSuccess) rc = fail(FQ_ERR_ALLOC, std::string("hipMalloc tables: ") + hipGetErrorString(e));
    }
    if (rc == FQ_SUCCESS && ntab > 0) {
        dim3 grid(unsigned((n + 255) / 256), unsigned(ntab));
        hipLaunchKernelGGL(k_tables, grid, dim3(256), 0, 0, h->W, d_terms, d_tbeg, N, P.L, P.sw, P.swpre, P.M,
                           P.Mpre, P.Nrho, P.nspcs, d_qz, d_qw, d_sz, d_szpre, d_m, d_eta, d_gam);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = fail(FQ_ERR_HIP, std::string("k_tables: ") + hipGetErrorString(e));
    }
    for (void *p : {(void *)d_terms, (void *)d_tbeg, (void *)d_qz, (void *)d_qw, (void *)d_sz, (void *)d_szpre,
                    (void *)d_m, (void *)d_eta, (void *)d_gam})
        if (p) (void)hipFree(p);
    return rc;
}

int build_projection(fq_handle *h) {
    // Base Gram matrix of {1, v, |v|^2} over the velocity grid (host FP64),
    // assembled into the Gram matrix of the species constraint rows and inverted.
    const fq::Plan &P = h->P;
    const int N = P.N;
    const double L = P.L;
    const int ns = P.ntgt;
    if (ns > 4) return fail(FQ_ERR_UNSUPPORTED, "conservation projection supports up to 4 species");
    std::vector<double> m(ns, 1.0);
    for (int s = 0; s < ns && s < int(P.masses.size()); ++s) m[s] = P.masses[s];
    double Gb[5][5] = {{0}};
    const double h_ = 2.0 * L / N;
    for (int I = 0; I < N; ++I)
        for (int J = 0; J < N; ++J)
            for (int K = 0; K < N; ++K) {
                const double cx = -L + L / N + h_ * I, cy = -L + L / N + h_ * J, cz = -L + L / N + h_ * K;
                const double ph[5] = {1, cx, cy, cz, cx * cx + cy * cy + cz * cz};
                for (int a = 0; a < 5; ++a)
                    for (int b = 0; b < 5; ++b) Gb[a][b] += ph[a] * ph[b];
            }
    const int K = ns + 4;
    double sm2 = 0;
    for (int s = 0; s < ns; ++s) sm2 += m[s] * m[s];
    std::vector<double> G(size_t(K) * K, 0.0);
    for (int r = 0; r < ns; ++r) {
        G[r * K + r] = Gb[0][0];
        for (int k = 1; k < 5; ++k) G[r * K + ns + k - 1] = G[(ns + k - 1) * K + r] = m[r] * Gb[0][k];
    }
    for (int a = 1; a < 5; ++a)
        for (int b = 1; b < 5; ++b) G[(ns + a - 1) * K + ns + b - 1] = sm2 * Gb[a][b];
    // Gauss-Jordan inverse with partial pivoting
    std::vector<double> A(size_t(K) * 2 * K, 0.0);
    for (int a = 0; a < K; ++a) {
        for (int b = 0; b < K; ++b) A[a * 2 * K + b] = G[a * K + b];
        A[a * 2 * K + K + a] = 1.0;
    }
    for (int c = 0; c < K; ++c) {
        int piv = c;
        for (int r = c + 1; r < K; ++r)
            if (std::fabs(A[r * 2 * K + c]) > std::fabs(A[piv * 2 * K + c])) piv = r;
        for (int b = 0; b < 2 * K; ++b) std::swap(A[c * 2 * K + b], A[piv * 2 * K + b]);
        const double d = A[c * 2 * K + c];
        if (d == 0.0) return fail(FQ_ERR_INVALID, "singular conservation Gram matrix");
        for (int b = 0; b < 2 * K; ++b) A[c * 2 * K + b] /= d;
        for (int r = 0; r < K; ++r)
            if (r != c) {
                const double f = A[r * 2 * K + c];
                for (int b = 0; b < 2 * K; ++b) A[r * 2 * K + b] -= f * A[c * 2 * K + b];
            }
    }
    std::vector<double> gi(size_t(K) * K);
    for (int a = 0; a < K; ++a)
        for (int b = 0; b < K; ++b) gi[a * K + b] = A[a * 2 * K + K + b];
    int rc = upload(&h->ginv, gi);
    if (rc == FQ_SUCCESS) rc = upload(&h->pmass, m);
    return rc;
}

// Group-pair tables: the entries (volume u, target t) with the contributions
// of groups 2u and 2u + 1 to t, their weights X and the Nyquist weights VN of
// every contribution, both derived from W on the device.
int build_hpair(fq_handle *h) {
    const fq::Plan &P = h->P;
    const int N = P.N, R = kHRows32;
    std::map<std::pair<int, int>, int> cix;           // (group, target) -> contribution
    for (size_t ci = 0; ci < P.contribs.size(); ++ci) cix[{P.contribs[ci].group, P.contribs[ci].tgt}] = int(ci);
    std::vector<int> ent_c, e_vol, te0(P.ntgt, 0), ten(P.ntgt, 0);
    for (int t = 0; t < P.ntgt; ++t) {
        te0[t] = int(e_vol.size());
        for (int u = 0; u < h->nvol; ++u) {
            int cc[2] = {-1, -1};
            for (int m = 0; m < 2; ++m) {
                auto it = cix.find({2 * u + m, t});
                if (it != cix.end()) cc[m] = it->second;
            }
            if (cc[0] < 0 && cc[1] < 0) continue;
            ent_c.push_back(cc[0]);
            ent_c.push_back(cc[1]);
            e_vol.push_back(u);
        }
        ten[t] = int(e_vol.size()) - te0[t];
    }
    int rc = FQ_SUCCESS;
    auto chk = [&](int r) { if (rc == FQ_SUCCESS) rc = r; };
    int *d_ent = nullptr;
    chk(upload(&d_ent, ent_c));
    chk(upload(&h->e_vol, e_vol));
    chk(upload(&h->t_e0, te0));
    chk(upload(&h->t_en, ten));
    const int nent = int(e_vol.size());
    const size_t xper = size_t(R) * 2 * N * N;
    if (rc == FQ_SUCCESS && nent > 0) {
        hipError_t e = hipMalloc(&h->X, nent * xper * sizeof(cpx<double>));
        if (e != hipSuccess) rc = fail(FQ_ERR_ALLOC, std::string("hipMalloc group-pair tables: ") + hipGetErrorString(e));
        if (rc == FQ_SUCCESS) {
            hipLaunchKernelGGL(k_htab, dim3(unsigned((xper + 255) / 256), unsigned(nent)), dim3(256), 0, 0, h->X,
                               h->W, d_ent, N, R);
            e = hipGetLastError();
            if (e != hipSuccess) rc = fail(FQ_ERR_HIP, std::string("k_htab: ") + hipGetErrorString(e));
        }
    }
    const int nc = int(P.contribs.size());
    if (rc == FQ_SUCCESS && h->hp_nyq && nc > 0) {
        hipError_t e = hipMalloc(&h->VN, size_t(nc) * kCPer32 * sizeof(cpx<float>));
        if (e != hipSuccess) rc = fail(FQ_ERR_ALLOC, std::string("hipMalloc Nyquist tables: ") + hipGetErrorString(e));
        if (rc == FQ_SUCCESS) {
            hipLaunchKernelGGL(k_vtab, dim3(unsigned((kCPer32 + 255) / 256), unsigned(nc)), dim3(256), 0, 0, h->VN,
                               h->W, N, R);
            e = hipGetLastError();
            if (e != hipSuccess) rc = fail(FQ_ERR_HIP, std::string("k_vtab: ") + hipGetErrorString(e));
        }
    }
    if (rc == FQ_SUCCESS) {
        hipError_t e = hipDeviceSynchronize();
        if (e != hipSuccess) rc = fail(FQ_ERR_HIP, std::string("group-pair tables: ") + hipGetErrorString(e));
    }
    if (d_ent) (void)hipFree(d_ent);
    return rc;
}

template <int N, typename T> int run_chunk(fq_handle *h, const Scratch &sc, const AosoaPtrs &ptr, const Dof &dof,
                                           int D, bool prof, hipStream_t st) {
    using G = Geo<N>;
    const fq::Plan &P = h->P;
    const int nsrc = P.nsrc, ntgt = P.ntgt, nitems = int(P.items.size()), ngroups = P.ngroups;
    constexpr int upb = kPencilThreads / G::LPP;
    constexpr int upp = kPlaneThreads / G::LPP;
    cpx<double> *T0 = reinterpret_cast<cpx<double> *>(sc.Z);
    cpx<T> *Z = reinterpret_cast<cpx<T> *>(sc.Z);
    cpx<T> *Pg = reinterpret_cast<cpx<T> *>(sc.Pg);
    cpx<double> *Y = reinterpret_cast<cpx<double> *>(sc.Z);
    const cpx<T> *ph = reinterpret_cast<const cpx<T> *>(h->ph);
    // HIP-event brackets around each launch when profiling is on
    auto evget = [&]() {
        hipEvent_t e;
        if (!h->ev_pool.empty()) { e = h->ev_pool.back(); h->ev_pool.pop_back(); }
        else (void)hipEventCreate(&e);
        return e;
    };
    std::pair<hipEvent_t, hipEvent_t> cur{};
    auto pre = [&]() {
        if (prof) { cur = {evget(), evget()}; (void)hipEventRecord(cur.first, st); }
    };
    auto post = [&](int k) {
        if (prof) { (void)hipEventRecord(cur.second, st); h->ev_rec.push_back({k, cur}); }
    };
    auto blocks = [&](int64_t units) { return unsigned((units + upb - 1) / upb); };
    auto pblocks = [&](int64_t units) { return unsigned((units + upp - 1) / upp); };
    auto xcd8 = [](unsigned nb) { return (nb + 7) / 8 * 8; };    // grids of xcd_block() kernels
    const dim3 bt(kPencilThreads), pt(kPlaneThreads);
    int64_t u;
    u = int64_t(D) * nsrc * N;
    pre();
    if constexpr (N == 32)
        hipLaunchKernelGGL(k_fwd_z4<N>, dim3(unsigned((D + 3) / 4 * int64_t(nsrc) * N)), bt, 0, st, ptr, dof, T0,
                           nsrc, D);
    else
        hipLaunchKernelGGL(k_fwd_z<N>, dim3(blocks(u)), bt, 0, st, ptr, dof, T0, nsrc, u);
    post(0);
    // the fused FP32 path reads an FP32 copy of F, written by k_fwd_xy itself
    const bool want32 = h->fused && std::is_same<T, float>::value;
    pre();
    cpx<double> *FJ64 = h->pack && sizeof(T) == 8 ? reinterpret_cast<cpx<double> *>(sc.FJ) : nullptr;
    cpx<float> *FJ32 = h->pack && sizeof(T) == 4 ? reinterpret_cast<cpx<float> *>(sc.FJ) : nullptr;
    hipLaunchKernelGGL(k_fwd_xy<N>, dim3(pblocks(u)), pt, 0, st, T0, sc.F, want32 ? sc.F32 : nullptr, FJ64, FJ32, u);
    post(1);
    if (h->fused) {
        const cpx<T> *FT = reinterpret_cast<const cpx<T> *>(want32 ? static_cast<void *>(sc.F32)
                                                                   : static_cast<void *>(sc.F));
        const cpx<T> *zc = reinterpret_cast<const cpx<T> *>(h->zc);
        const dim3 grid(unsigned(int64_t(D) * ngroups * 4));
        bool multi = false;
        for (int g = 0; g < ngroups; ++g) multi |= P.group_np[g] > 1;
        bool one_lane = false;
        if constexpr ((N == 24 || N == 32) && sizeof(T) == 4) {
            // one lane per pencil, with pair packing unless switched off
            one_lane = !h->two_lane;
            if (one_lane) {
                const dim3 blk1(Prod1Geo<N>::THREADS);
                const int nprods = int(P.prods.size());
                cpx<T> *NY = reinterpret_cast<cpx<T> *>(sc.NY);
                const bool cj = h->pack_conj;
                if (h->pack) {
                    auto knyq = cj ? k_nyq<N, T, true> : k_nyq<N, T, false>;
                    pre();
                    hipLaunchKernelGGL(knyq, dim3(unsigned(int64_t(D) * nprods)), dim3(96), 0, st, FT,
                                       reinterpret_cast<const cpx<T> *>(sc.FJ), NY,
                                       reinterpret_cast<const cpx<T> *>(h->nyv), ph, h->prods, h->item_src, nprods,
                                       nsrc);
                    post(8);
                }
                using KP = decltype(&k_prod1<N, T, true, true, false>);
                KP kern;
                dim3 pgrid = grid;
                if (!h->pack) kern = multi ? k_prod1<N, T, true, false, false> : k_prod1<N, T, false, false, false>;
                else if (cj) kern = multi ? k_prod1<N, T, true, true, true> : k_prod1<N, T, false, true, true>;
                else kern = multi ? k_prod1<N, T, true, true, false> : k_prod1<N, T, false, true, false>;
                if constexpr (N == 32) {
                    if (h->hpair) {
                        kern = cj ? k_prod1<N, T, false, true, true, true> : k_prod1<N, T, false, true, false, true>;
                        pgrid = dim3(unsigned(int64_t(D) * h->nvol * 4));
                    }
                }
                const cpx<T> *zq = reinterpret_cast<const cpx<T> *>(h->zq);
                pre();
                hipLaunchKernelGGL(kern, pgrid, blk1, 0, st, FT, Pg, ph, zc, h->prods, h->g_p0, h->g_np,
                                   h->item_src, ngroups, nsrc, NY, nprods, zq, h->pconj,
                                   h->hpair && h->hp_nyq ? sc.NB : nullptr);
                post(7);
                if (h->hpair && h->hp_nyq) {
                    pre();
                    hipLaunchKernelGGL(k_nyqb, dim3(unsigned(int64_t(D) * ntgt)), dim3(256), 0, st, sc.NB, h->VN,
                                       h->c_group, h->t_c0, h->t_cn, sc.C, ngroups, ntgt);
                    post(9);
                }
            }
        }
        if (!one_lane) {
            // the accumulating variant also for single-product groups in FP64:
            // with its accumulator live across the item stages the compiler
            // keeps the FP64 kernel within 256 VGPRs without spills
            const bool m2 = multi || sizeof(T) == 8;
            const int nprods = int(P.prods.size());
            cpx<T> *NY = reinterpret_cast<cpx<T> *>(sc.NY);
            const cpx<T> *zq = reinterpret_cast<const cpx<T> *>(h->zq);
            using KP = decltype(&k_prod<N, T, true, false>);
            KP kern = m2 ? k_prod<N, T, true, false> : k_prod<N, T, false, false>;
            if constexpr ((N == 24 || N == 32) && sizeof(T) == 8) {
                if (h->pack) {
                    pre();
                    hipLaunchKernelGGL((k_nyq<N, T, true>), dim3(unsigned(int64_t(D) * nprods)), dim3(96), 0, st, FT,
                                       reinterpret_cast<const cpx<T> *>(sc.FJ), NY, nullptr, ph, h->prods,
                                       h->item_src, nprods, nsrc);
                    post(8);
                    // packed: one item stage per product, so the single-product
                    // variant stays within 256 VGPRs
                    kern = multi ? k_prod<N, T, true, true> : k_prod<N, T, false, true>;
                }
            }
            pre();
            hipLaunchKernelGGL(kern, grid, dim3(ProdGeo<N>::THREADS), 0, st, FT, Pg, ph, zc, h->prods, h->g_p0,
                               h->g_np, h->item_src, ngroups, nsrc, NY, nprods, zq);
            post(7);
        }
    } else {
    u = int64_t(D) * nitems * N;
    pre();
    hipLaunchKernelGGL((k_inv_z<N, T>), dim3(blocks(u)), bt, 0, st, sc.F, Z, ph, h->item_src, nitems, nsrc, u);
    post(2);
    u = int64_t(D) * ngroups * N;
    bool multi = false;
    for (int g = 0; g < ngroups; ++g) multi |= P.group_np[g] > 1;
    pre();
    if (multi)
        hipLaunchKernelGGL((k_plane<N, T, true>), dim3(xcd8(pblocks(u))), pt, 0, st, Z, Pg, ph, h->prods, h->g_p0,
                           h->g_np, nitems, ngroups, u);
    else
        hipLaunchKernelGGL((k_plane<N, T, false>), dim3(xcd8(pblocks(u))), pt, 0, st, Z, Pg, ph, h->prods, h->g_p0,
                           h->g_np, nitems, ngroups, u);
    post(3);
    }
    u = int64_t(D) * ntgt * N;
    pre();
    if (h->hpair) {
        if constexpr (N == 32 && sizeof(T) == 4) {
            // group pairs: rows r = 0..N/2, eight (or with FQ_FLAG_ACCUM_QUAD four)
            // DOF slots per workgroup
            const bool eight = !(h->flags & FQ_FLAG_ACCUM_QUAD);
            auto kacc = eight ? k_accum_h<8> : k_accum_h<4>;
            const int NS = eight ? 8 : 4;
            hipLaunchKernelGGL(kacc, dim3(unsigned(int64_t((D + NS - 1) / NS) * ntgt * kHRows32)), dim3(64 * NS), 0,
                               st, reinterpret_cast<const cpx<float> *>(sc.Pg), sc.F, h->X, h->e_vol, h->t_e0,
                               h->t_en, h->hp_nyq ? sc.C : nullptr, h->W, h->l_src, h->t_l0, h->t_ln,
                               int(P.contribs.size()), Y, h->nvol, nsrc, ntgt, D);
        }
    } else if constexpr (N == 32) {
        // FP32 pencils: eight DOF slots per workgroup (one per CU) unless
        // FQ_FLAG_ACCUM_QUAD asks for four (two per CU); FP64: four (three
        // workgroups per CU, register-limited)
        const bool eight = sizeof(T) == 4 && !(h->flags & FQ_FLAG_ACCUM_QUAD);
        auto kacc = eight ? k_accum_lds<N, T, 8> : k_accum_lds<N, T, 4>;
        const int NS = eight ? 8 : 4;
        hipLaunchKernelGGL(kacc, dim3(unsigned(int64_t((D + NS - 1) / NS) * ntgt * N)), dim3(64 * NS), 0, st,
                           Pg, sc.F, h->W, h->c_group, h->t_c0, h->t_cn, h->l_src, h->t_l0, h->t_ln,
                           int(P.contribs.size()), Y, ngroups, nsrc, ntgt, D);
    } else {
        hipLaunchKernelGGL((k_accum<N, T>), dim3(blocks(u)), bt, 0, st, Pg, sc.F, h->W, h->c_group, h->t_c0,
                           h->t_cn, h->l_src, h->t_l0, h->t_ln, int(P.contribs.size()), Y, ngroups, nsrc, ntgt, D, u);
    }
    post(4);
    pre();
    bool out1 = false;
    if constexpr (N == 32) {
        // group pairs: gain + i loss in one volume, one transform per plane
        if (h->hpair) {
            const dim3 g1(unsigned((D + 3) / 4 * int64_t(ntgt) * N)), b1(4 * Geo<N>::LPP);
            if (h->flags & FQ_FLAG_PROJECT)
                hipLaunchKernelGGL((k_out1<N, true>), g1, b1, 0, st, Y, ptr, dof, ntgt, D, sc.MOM, P.L);
            else
                hipLaunchKernelGGL((k_out1<N, false>), g1, b1, 0, st, Y, ptr, dof, ntgt, D, nullptr, P.L);
            out1 = true;
        }
    }
    if (!out1) hipLaunchKernelGGL(k_out<N>, dim3(pblocks(u)), pt, 0, st, Y, ptr, dof, ntgt, u);
    post(5);
    if (h->flags & FQ_FLAG_PROJECT) {
        // per batch, right after k_out: the batch's Q is still in L2
        pre();
        if (h->hpair)
            hipLaunchKernelGGL(k_project_b, dim3(unsigned((D + 31) / 32 * (int64_t(N) * N * N / kProjChunk))),
                               dim3(256), 0, st, ptr, dof, ntgt, N, P.L, h->ginv, h->pmass, sc.MOM, int(D));
        else
        hipLaunchKernelGGL(k_project, dim3(unsigned(8 * ((D + 7) / 8))), dim3(256), 0, st, ptr, dof, ntgt, N, P.L,
                           h->ginv, h->pmass, int(D));
        post(6);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FQ_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return FQ_SUCCESS;
}

// One batch of D DOFs on a grid without native kernels (see k_gen_gather):
// library transforms of `chunk` slots and grid-stride pointwise kernels.
#define FFTCHK(x)                                                                          \
    do {                                                                                   \
        hipfftResult r_ = (x);                                                             \
        if (r_ != HIPFFT_SUCCESS) return fail(FQ_ERR_HIP, "hipFFT error " + std::to_string(int(r_)) + " in " #x); \
    } while (0)

template <typename T> hipfftResult gen_exec(hipfftHandle p, cpx<T> *x, int dir) {
    if constexpr (sizeof(T) == 8)
        return hipfftExecZ2Z(p, reinterpret_cast<hipfftDoubleComplex *>(x), reinterpret_cast<hipfftDoubleComplex *>(x),
                             dir);
    else
        return hipfftExecC2C(p, reinterpret_cast<hipfftComplex *>(x), reinterpret_cast<hipfftComplex *>(x), dir);
}

template <typename T>
int run_chunk_gen(fq_handle *h, const Scratch &sc, const AosoaPtrs &ptr, const Dof &dof, int D, bool prof,
                  hipStream_t st) {
    const fq::Plan &P = h->P;
    const int N = P.N, nsrc = P.nsrc, ntgt = P.ntgt;
    const int64_t n = int64_t(N) * N * N, chunk = h->chunk;
    cpx<double> *F = sc.F;
    cpx<double> *Y = reinterpret_cast<cpx<double> *>(sc.Z);
    cpx<T> *A = reinterpret_cast<cpx<T> *>(sc.Pg);                  // A, then B at A + chunk n, then G
    cpx<T> *Gv = A + 2 * chunk * n;
    const cpx<T> *ph = reinterpret_cast<const cpx<T> *>(h->ph);
    auto grid = [&](int64_t total) { return dim3(unsigned(std::min<int64_t>((total + 255) / 256, 16384))); };
    const dim3 bt(256);
    std::pair<hipEvent_t, hipEvent_t> cur{};
    auto pre = [&]() {
        if (!prof) return;
        auto get = [&]() {
            hipEvent_t e;
            if (!h->ev_pool.empty()) { e = h->ev_pool.back(); h->ev_pool.pop_back(); }
            else (void)hipEventCreate(&e);
            return e;
        };
        cur = {get(), get()};
        (void)hipEventRecord(cur.first, st);
    };
    auto post = [&](int k) {
        if (prof) { (void)hipEventRecord(cur.second, st); h->ev_rec.push_back({k, cur}); }
    };
    for (hipfftHandle p : {h->gp_src, h->gp_items, h->gp_group, h->gp_out}) FFTCHK(hipfftSetStream(p, st));
    // F = DFT-(f)/n
    pre();
    hipLaunchKernelGGL(k_gen_gather, grid(chunk * nsrc * n), bt, 0, st, ptr, dof, F, nsrc, D, N, 1.0 / double(n),
                       chunk * nsrc * n);
    FFTCHK(gen_exec<double>(h->gp_src, F, HIPFFT_FORWARD));
    post(0);
    // the loss spectra Y_t[1] = sum_l W_l F_src(l); the gain Y_t[0] starts at zero
    pre();
    hipLaunchKernelGGL(k_gen_zero, grid(int64_t(D) * ntgt * 2 * n), bt, 0, st, Y, int64_t(D) * ntgt * 2 * n);
    for (size_t l = 0; l < P.loss.size(); ++l)
        hipLaunchKernelGGL(k_gen_wmac<double>, grid(D * n), bt, 0, st, F + int64_t(P.loss[l].src) * n, nsrc * n,
                           h->W + (int64_t(P.contribs.size()) + int64_t(l)) * n, Y + (int64_t(P.loss[l].tgt) * 2 + 1) * n,
                           int64_t(ntgt) * 2 * n, N, D * n);
    post(4);
    for (int g = 0; g < P.ngroups; ++g) {
        pre();
        for (int k = 0; k < P.group_np[g]; ++k) {
            const fq::PProd &pr = P.prods[P.group_p0[g] + k];
            hipLaunchKernelGGL(k_gen_phase<T>, grid(D * n), bt, 0, st, F, ph, pr.i1, P.items[pr.i1].src, pr.i2,
                               P.items[pr.i2].src, nsrc, A, chunk * n, N, D * n);
            FFTCHK(gen_exec<T>(h->gp_items, A, HIPFFT_BACKWARD));
            hipLaunchKernelGGL(k_gen_prod<T>, grid(D * n), bt, 0, st, A, chunk * n, Gv, T(pr.coef), k == 0 ? 1 : 0,
                               D * n);
        }
        FFTCHK(gen_exec<T>(h->gp_group, Gv, HIPFFT_FORWARD));
        post(7);
        pre();
        for (int c : h->g_contribs[g])
            hipLaunchKernelGGL(k_gen_wmac<T>, grid(D * n), bt, 0, st, Gv, n, h->W + int64_t(c) * n,
                               Y + int64_t(P.contribs[c].tgt) * 2 * n, int64_t(ntgt) * 2 * n, N, D * n);
        post(4);
    }
    // Q_t = Re DFT+(H_t) - f_t Re DFT+(L_t)
    pre();
    FFTCHK(gen_exec<double>(h->gp_out, Y, HIPFFT_BACKWARD));
    hipLaunchKernelGGL(k_gen_out, grid(int64_t(D) * ntgt * n), bt, 0, st, Y, ptr, dof, ntgt, N,
                       int64_t(D) * ntgt * n);
    post(5);
    if (h->flags & FQ_FLAG_PROJECT) {
        pre();
        hipLaunchKernelGGL(k_project, dim3(unsigned(8 * ((D + 7) / 8))), dim3(256), 0, st, ptr, dof, ntgt, N, P.L,
                           h->ginv, h->pmass, int(D));
        post(6);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(FQ_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return FQ_SUCCESS;
}

// hipFFT plans of a generic-grid handle for its `chunk` slots, with one work
// area of the largest size any of them asks for.
int make_gen_plans(fq_handle *h) {
    const fq::Plan &P = h->P;
    const int N = P.N;
    const int n = N * N * N;
    int dims[3] = {N, N, N};
    const hipfftType tt = h->precision == FQ_PREC_MIXED ? HIPFFT_C2C : HIPFFT_Z2Z;
    const int64_t c = h->chunk;
    struct { hipfftHandle *p; hipfftType t; int64_t batch; } pl[4] = {
        {&h->gp_src, HIPFFT_Z2Z, c * P.nsrc}, {&h->gp_items, tt, 2 * c}, {&h->gp_group, tt, c},
        {&h->gp_out, HIPFFT_Z2Z, c * P.ntgt * 2}};
    size_t wmax = 0;
    for (auto &q : pl) {
        if (q.batch > INT32_MAX / 2) return fail(FQ_ERR_UNSUPPORTED, "batch too large for the FFT library plans");
        size_t ws = 0;
        FFTCHK(hipfftCreate(q.p));
        FFTCHK(hipfftSetAutoAllocation(*q.p, 0));
        FFTCHK(hipfftMakePlanMany(*q.p, 3, dims, nullptr, 1, n, nullptr, 1, n, q.t, int(q.batch), &ws));
        wmax = std::max(wmax, ws);
    }
    if (wmax > 0) {
        hipError_t e = hipMalloc(&h->gen_work, wmax);
        if (e != hipSuccess) return fail(FQ_ERR_ALLOC, "hipMalloc FFT work area failed (reduce max_chunk)");
        for (auto &q : pl) FFTCHK(hipfftSetWorkArea(*q.p, h->gen_work));
        h->scratch += wmax;                   // reported with the batch scratch
    }
    h->g_contribs.assign(P.ngroups, {});
    for (size_t ci = 0; ci < P.contribs.size(); ++ci) h->g_contribs[P.contribs[ci].group].push_back(int(ci));
    return FQ_SUCCESS;
}

template <int N>
int dispatch_prec(fq_handle *h, const Scratch &sc, const AosoaPtrs &ptr, const Dof &dof, int D, bool prof,
                  hipStream_t st) {
    return h->precision == FQ_PREC_MIXED ? run_chunk<N, float>(h, sc, ptr, dof, D, prof, st)
                                         : run_chunk<N, double>(h, sc, ptr, dof, D, prof, st);
}

// Batches of at most `chunk` DOFs on the scratch sets `sc[0..nsets)`; prof =
// per-kernel HIP-event brackets (which create events, so the no-alloc entry
// point passes false).  With two sets, batch i runs on set i % 2: even
// batches on the caller's stream, odd ones on the handle's second stream,
// forked from and joined back to the caller's stream by events, so one
// batch's memory-bound stages can share the device with the next batch's
// product kernel.  Batches i and i + 2 share a set and a stream.
int eval_impl(fq_handle *h, const Scratch *sc, int nsets, int64_t chunk, bool prof, const AosoaPtrs &ptr,
              int32_t nupts, int64_t ldim, int32_t neles, int64_t first_dof, int64_t ndof, void *stream) {
    if (!h) return fail(FQ_ERR_INVALID, "null handle");
    if (nupts <= 0 || neles <= 0 || ndof < 0 || first_dof < 0)
        return fail(FQ_ERR_INVALID, "invalid DOF range or shape");
    if (first_dof + ndof > int64_t(nupts) * neles) return fail(FQ_ERR_INVALID, "DOF range exceeds nupts*neles");
    const int64_t n = int64_t(h->P.N) * h->P.N * h->P.N;
    const int64_t nparr = (neles + 31) / 32 * 32;
    if (ldim < n * nparr) return fail(FQ_ERR_INVALID, "ldim smaller than N^3 * padded neles");
    if (ndof == 0) return FQ_SUCCESS;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev != h->device) return fail(FQ_ERR_INVALID, "handle was created on another device");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (chunk <= 0) return fail(FQ_ERR_INVALID, "no scratch for this call");
    // equal batches of at most chunk DOFs (no short tail batch) ...
    int64_t nbatch = (ndof + chunk - 1) / chunk;
    int64_t per = (ndof + nbatch - 1) / nbatch;
    // ... except on the 32^3 grid, where a batch is a multiple of the DOFs
    // whose k_accum_lds workgroups (NS DOFs x N rows each) fill whole rounds
    // of the device's workgroup slots when the scratch holds at least that
    // many: the last batch is then the short one
    if (h->P.N == 32 && h->fused && !h->generic && nbatch > 1) {
        // FP32: 8 slots per workgroup and one workgroup per CU, or 4 slots and
        // two; FP64: 4 slots, three workgroups per CU
        // group pairs: 8 slots x 17 row pairs per workgroup, one workgroup per CU
        const int64_t align = h->hpair ? int64_t(h->cus) * 8 / kHRows32 / 8 * 8
                                       : int64_t(h->cus) * (h->precision == FQ_PREC_MIXED ? 8 : 12) / 32;
        if (align > 0 && chunk >= align) {
            const int64_t pmax = chunk / align * align;
            nbatch = (ndof + pmax - 1) / pmax;
            per = ((ndof + nbatch - 1) / nbatch + align - 1) / align * align;
        }
    }
    const bool two = nsets > 1 && h->st2 && (ndof + per - 1) / per > 1;
    if (two) {
        HIPCHK(hipEventRecord(h->ev_fork, st));
        HIPCHK(hipStreamWaitEvent(h->st2, h->ev_fork, 0));
    }
    int64_t i = 0;
    for (int64_t c0 = 0; c0 < ndof; c0 += per, ++i) {
        const int D = int(std::min<int64_t>(per, ndof - c0));
        Dof dof{first_dof + c0, nupts, 32, ldim, n};
        const Scratch &s = sc[two ? i % 2 : 0];
        const hipStream_t bs = two && (i % 2) ? h->st2 : st;
        int rc;
        if (h->generic) {
            rc = h->precision == FQ_PREC_MIXED ? run_chunk_gen<float>(h, s, ptr, dof, D, prof, bs)
                                               : run_chunk_gen<double>(h, s, ptr, dof, D, prof, bs);
            if (rc != FQ_SUCCESS) return rc;
            continue;
        }
        switch (h->P.N) {
        case 8: rc = dispatch_prec<8>(h, s, ptr, dof, D, prof, bs); break;
        case 16: rc = dispatch_prec<16>(h, s, ptr, dof, D, prof, bs); break;
        case 24: rc = dispatch_prec<24>(h, s, ptr, dof, D, prof, bs); break;
        case 32: rc = dispatch_prec<32>(h, s, ptr, dof, D, prof, bs); break;
        default: rc = fail(FQ_ERR_UNSUPPORTED, "grid size not built");
        }
        if (rc != FQ_SUCCESS) return rc;
    }
    if (two) {
        HIPCHK(hipEventRecord(h->ev_join, h->st2));
        HIPCHK(hipStreamWaitEvent(st, h->ev_join, 0));
    }
    return FQ_SUCCESS;
}

// The evaluation entry points on the handle's own scratch.
int eval_own(fq_handle *h, const AosoaPtrs &ptr, int32_t nupts, int64_t ldim, int32_t neles, int64_t first_dof,
             int64_t ndof, void *stream) {
    if (h->flags & FQ_FLAG_NO_SCRATCH)
        return fail(FQ_ERR_INVALID, "handle created with FQ_FLAG_NO_SCRATCH has no scratch; use fq_eval_noalloc");
    // set s of each region starts s * (per-DOF bytes * chunk) into it
    Scratch sc[2];
    for (int k = 0; k < h->nsets; ++k) {
        const size_t c = size_t(h->chunk) * k;
        auto at = [&](void *base, size_t per) -> void * {
            return base ? static_cast<void *>(static_cast<unsigned char *>(base) + per * c) : nullptr;
        };
        sc[k].F = static_cast<cpx<double> *>(at(h->F, h->per_dof_F));
        sc[k].Z = at(h->Z, h->per_dof_Z);
        sc[k].Pg = at(h->Pg, h->per_dof_P);
        sc[k].F32 = static_cast<cpx<float> *>(at(h->F32, h->per_dof_F32));
        sc[k].NY = at(h->NY, h->per_dof_NY);
        sc[k].FJ = at(h->FJ, h->per_dof_FJ);
        sc[k].NB = static_cast<float *>(at(h->NB, h->per_dof_NB));
        sc[k].C = static_cast<cpx<double> *>(at(h->Cn, h->per_dof_C));
        sc[k].MOM = static_cast<double *>(at(h->MOM, h->per_dof_M));
    }
    return eval_impl(h, sc, h->nsets, h->chunk, h->prof, ptr, nupts, ldim, neles, first_dof, ndof, stream);
}

std::mutex g_noalloc_mu;
std::map<int, unsigned char *> g_noalloc_ptr;     // device -> address of `scratch`

// Address of the static scratch on the current device `dev` (nullptr on failure).
unsigned char *noalloc_resolve(int dev) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    auto it = g_noalloc_ptr.find(dev);
    if (it != g_noalloc_ptr.end()) return it->second;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return nullptr;
    }
    g_noalloc_ptr[dev] = static_cast<unsigned char *>(p);
    return static_cast<unsigned char *>(p);
}

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Bytes of the static scratch that a batch of c DOFs occupies.
size_t noalloc_bytes(const fq_handle *h, int64_t c) {
    const size_t cc = size_t(c);
    return align256(h->per_dof_F * cc) + align256(h->per_dof_Z * cc) + align256(h->per_dof_P * cc) +
           align256(h->per_dof_F32 * cc) + align256(h->per_dof_NY * cc) + align256(h->per_dof_FJ * cc) +
           align256(h->per_dof_NB * cc) + align256(h->per_dof_C * cc) + align256(h->per_dof_M * cc);
}

// Carves the regions of a batch of c DOFs out of the static scratch at base.
Scratch noalloc_carve(const fq_handle *h, unsigned char *base, int64_t c) {
    const size_t cc = size_t(c);
    Scratch sc;
    size_t o = 0;
    auto take = [&](size_t per) -> void * {
        if (per == 0) return nullptr;
        void *p = base + o;
        o += align256(per * cc);
        return p;
    };
    sc.F = static_cast<cpx<double> *>(take(h->per_dof_F));
    sc.Z = take(h->per_dof_Z);
    sc.Pg = take(h->per_dof_P);
    sc.F32 = static_cast<cpx<float> *>(take(h->per_dof_F32));
    sc.NY = take(h->per_dof_NY);
    sc.FJ = take(h->per_dof_FJ);
    sc.NB = static_cast<float *>(take(h->per_dof_NB));
    sc.C = static_cast<cpx<double> *>(take(h->per_dof_C));
    sc.MOM = static_cast<double *>(take(h->per_dof_M));
    return sc;
}

}  // namespace

int fq_internal_fail(int code, const std::string &msg) { return fail(code, msg); }

extern "C" {

const char *fq_last_error(void) { return g_err.c_str(); }

int fq_plan_create(const fq_params *prm, fq_plan **out) {
    if (!prm || !out) return fail(FQ_ERR_INVALID, "null argument");
    *out = nullptr;
    auto *p = new fq_plan();
    std::string err;
    if (!fq::plan_build(*prm, p->P, err)) {
        delete p;
        return fail(FQ_ERR_INVALID, err);
    }
    *out = p;
    return FQ_SUCCESS;
}

int fq_plan_destroy(fq_plan *p) {
    delete p;
    return FQ_SUCCESS;
}

int fq_plan_get_info(const fq_plan *p, fq_plan_info *info) {
    if (!p || !info) return fail(FQ_ERR_INVALID, "null argument");
    fq::plan_info(p->P, *info);
    return FQ_SUCCESS;
}

int fq_plan_export(const fq_plan *pl, int32_t *item_src, double *item_theta, int32_t *prod_i1, int32_t *prod_i2,
                   double *prod_coef, int32_t *prod_group, int32_t *contrib_group, int32_t *contrib_target,
                   int32_t *loss_src, int32_t *loss_target) {
    if (!pl) return fail(FQ_ERR_INVALID, "null plan");
    const fq::Plan &P = pl->P;
    for (size_t i = 0; i < P.items.size(); ++i) {
        if (item_src) item_src[i] = P.items[i].src;
        if (item_theta)
            for (int d = 0; d < 3; ++d) item_theta[3 * i + d] = P.items[i].th[d];
    }
    for (size_t k = 0; k < P.prods.size(); ++k) {
        if (prod_i1) prod_i1[k] = P.prods[k].i1;
        if (prod_i2) prod_i2[k] = P.prods[k].i2;
        if (prod_coef) prod_coef[k] = P.prods[k].coef;
        if (prod_group) prod_group[k] = P.prods[k].group;
    }
    for (size_t c = 0; c < P.contribs.size(); ++c) {
        if (contrib_group) contrib_group[c] = P.contribs[c].group;
        if (contrib_target) contrib_target[c] = P.contribs[c].tgt;
    }
    for (size_t l = 0; l < P.loss.size(); ++l) {
        if (loss_src) loss_src[l] = P.loss[l].src;
        if (loss_target) loss_target[l] = P.loss[l].tgt;
    }
    return FQ_SUCCESS;
}

int fq_create(const fq_params *prm, fq_handle **out) {
    if (!prm || !out) return fail(FQ_ERR_INVALID, "null argument");
    *out = nullptr;
    if (prm->precision != FQ_PREC_FP64 && prm->precision != FQ_PREC_MIXED)
        return fail(FQ_ERR_INVALID, "unknown precision");
    const bool generic = !n_supported(prm->Nv) || (prm->flags & FQ_FLAG_LIBRARY_FFT);
    if (generic && (prm->Nv < 4 || prm->Nv > kGenMaxN || prm->Nv % 2))
        return fail(FQ_ERR_UNSUPPORTED, "Nv must be even, from 4 to " + std::to_string(kGenMaxN));
    if (generic && (prm->flags & FQ_FLAG_NO_SCRATCH))
        return fail(FQ_ERR_UNSUPPORTED, "FQ_FLAG_NO_SCRATCH needs the native kernels (Nv 8, 16, 24, 32, "
                                         "without FQ_FLAG_LIBRARY_FFT)");
    if (prm->soasz != 32) return fail(FQ_ERR_UNSUPPORTED, "soasz must be 32");
    auto *h = new fq_handle();
    std::string err;
    if (!fq::plan_build(*prm, h->P, err)) {
        delete h;
        return fail(FQ_ERR_INVALID, err);
    }
    h->precision = prm->precision;
    h->flags = prm->flags;
    h->generic = generic;
    // the schedule flags select native kernels; the library path has one schedule
    h->fused = generic || !(prm->flags & FQ_FLAG_UNFUSED);
    h->two_lane = !generic && (prm->flags & FQ_FLAG_TWO_LANE) != 0;
    h->pack = !generic && h->fused && !(h->two_lane && h->precision == FQ_PREC_MIXED) &&
              (prm->Nv == 24 || prm->Nv == 32) && !(prm->flags & FQ_FLAG_NO_PACK);
    h->nvol = generic ? 3 : h->P.ngroups;         // library path: the items A, B and the group G
    if (h->pack && !h->two_lane && prm->Nv == 32 && h->precision == FQ_PREC_MIXED && !(prm->flags & FQ_FLAG_NO_HPAIR)) {
        bool single = true;
        for (int g = 0; g < h->P.ngroups; ++g) single = single && h->P.group_np[g] == 1;
        h->hpair = single;
        // VHS weights are real and even (sinc of |l|), so V vanishes there
        h->hp_nyq = single && h->P.model == FQ_MODEL_VSS;
        if (single) h->nvol = (h->P.ngroups + 1) / 2;
    }
    if (hipGetDevice(&h->device) != hipSuccess) {
        delete h;
        return fail(FQ_ERR_HIP, "no HIP device");
    }
    if (hipDeviceGetAttribute(&h->cus, hipDeviceAttributeMultiprocessorCount, h->device) != hipSuccess ||
        h->cus <= 0)
        h->cus = 256;
    const fq::Plan &P = h->P;
    const int N = P.N;
    const int64_t n = int64_t(N) * N * N;
    int rc = h->precision == FQ_PREC_MIXED ? upload_phases<float>(h) : upload_phases<double>(h);
    std::vector<int> isrc;
    for (auto &it : P.items) isrc.push_back(it.src);
    std::vector<DevProd> dp;
    for (auto &pr : P.prods) dp.push_back({pr.i1, pr.i2, pr.coef});
    std::vector<int> cg, tc0(P.ntgt, 0), tcn(P.ntgt, 0), ls, tl0(P.ntgt, 0), tln(P.ntgt, 0);
    for (auto &c : P.contribs) { cg.push_back(c.group); tcn[c.tgt]++; }
    for (int t = 1; t < P.ntgt; ++t) tc0[t] = tc0[t - 1] + tcn[t - 1];
    for (auto &l : P.loss) { ls.push_back(l.src); tln[l.tgt]++; }
    for (int t = 1; t < P.ntgt; ++t) tl0[t] = tl0[t - 1] + tln[t - 1];
    auto chk = [&](int r) { if (rc == FQ_SUCCESS) rc = r; };
    chk(upload(&h->item_src, isrc));
    chk(upload(&h->prods, dp));
    chk(upload(&h->g_p0, P.group_p0));
    chk(upload(&h->g_np, P.group_np));
    chk(upload(&h->c_group, cg));
    chk(upload(&h->t_c0, tc0));
    chk(upload(&h->t_cn, tcn));
    chk(upload(&h->l_src, ls));
    chk(upload(&h->t_l0, tl0));
    chk(upload(&h->t_ln, tln));
    if (rc == FQ_SUCCESS) rc = build_tables(h);
    if (rc == FQ_SUCCESS) rc = build_projection(h);
    if (rc == FQ_SUCCESS && h->pack)
        rc = h->precision == FQ_PREC_MIXED ? upload_nyq<float>(h) : upload_nyq<double>(h);
    if (rc == FQ_SUCCESS && h->hpair) rc = build_hpair(h);
    // scratch: F (FP64, nsrc) + max(Z (items, T), T0 (nsrc, FP64), Y (2 ntgt, FP64)) + P (groups, T)
    const size_t tsz = h->precision == FQ_PREC_MIXED ? sizeof(cpx<float>) : sizeof(cpx<double>);
    const size_t per_dof_F = size_t(P.nsrc) * n * sizeof(cpx<double>);
    // Z holds the staged inverse transforms (three-kernel path only), T0 and Y
    const size_t per_dof_Z = std::max({h->fused ? size_t(0) : P.items.size() * n * tsz,
                                       size_t(P.nsrc) * n * sizeof(cpx<double>),
                                       size_t(2 * P.ntgt) * n * sizeof(cpx<double>)});
    const size_t per_dof_P = size_t(h->nvol) * n * tsz;
    const size_t per_dof_F32 = h->precision == FQ_PREC_MIXED ? size_t(P.nsrc) * n * sizeof(cpx<float>) : 0;
    // Nyquist-plane corrections of the pair-packed product kernel and the
    // spectrum slices they are computed from
    const size_t per_dof_NY = h->pack ? P.prods.size() * 3 * size_t(N) * N * tsz : 0;
    const size_t per_dof_FJ = h->pack ? size_t(P.nsrc) * N * N * tsz : 0;
    // group pairs: the Nyquist sums of Im G_g and the per-target corrections
    const size_t per_dof_NB = h->hp_nyq ? size_t(P.ngroups) * kNbPer32 * sizeof(float) : 0;
    const size_t per_dof_C = h->hp_nyq ? size_t(P.ntgt) * kCPer32 * sizeof(cpx<double>) : 0;
    const size_t per_dof_M = h->hpair && (h->flags & FQ_FLAG_PROJECT) ? size_t(P.ntgt) * N * 5 * sizeof(double) : 0;
    const size_t per_dof = per_dof_F + per_dof_Z + per_dof_P + per_dof_F32 + per_dof_NY + per_dof_FJ +
                           per_dof_NB + per_dof_C + per_dof_M;
    h->per_dof_F = per_dof_F; h->per_dof_Z = per_dof_Z; h->per_dof_P = per_dof_P;
    h->per_dof_F32 = per_dof_F32; h->per_dof_NY = per_dof_NY; h->per_dof_FJ = per_dof_FJ;
    h->per_dof_NB = per_dof_NB; h->per_dof_C = per_dof_C; h->per_dof_M = per_dof_M;
    // the static scratch of fq_eval_noalloc, resolved with this device current
    h->noalloc = noalloc_resolve(h->device);
    if (rc == FQ_SUCCESS && (h->flags & FQ_FLAG_NO_SCRATCH) && !h->noalloc)
        rc = fail(FQ_ERR_HIP, "hipGetSymbolAddress of the static scratch failed");
    // two scratch sets (FQ_FLAG_OVERLAP) share the budget; max_chunk bounds each set's batch
    // (the library path's plans share one work area: one set, one stream)
    h->nsets = (h->flags & FQ_FLAG_OVERLAP) && !(h->flags & FQ_FLAG_NO_SCRATCH) && !generic ? 2 : 1;
    const size_t ns = size_t(h->nsets);
    int64_t chunk = prm->max_chunk;
    if (chunk <= 0) {
        const size_t budget = size_t(8) << 30;    // 8 GiB default scratch budget (of 288 GB HBM)
        // the library path also holds an FFT work area of up to one Y per DOF
        chunk = std::max<int64_t>(1, int64_t(budget / ((per_dof + (generic ? per_dof_Z : 0)) * ns)));
        chunk = std::min<int64_t>(chunk, generic ? 512 : 4096);
    }
    h->chunk = chunk;
    if (rc == FQ_SUCCESS && (h->flags & FQ_FLAG_NO_SCRATCH)) {
        h->chunk = 0;                      // evaluation only through fq_eval_noalloc
    } else if (rc == FQ_SUCCESS) {
        const size_t cs = size_t(chunk) * ns;
        hipError_t e1 = hipMalloc(&h->F, per_dof_F * cs);
        hipError_t e2 = hipMalloc(&h->Z, per_dof_Z * cs);
        hipError_t e3 = hipMalloc(&h->Pg, per_dof_P * cs);
        hipError_t e4 = hipSuccess;
        if (h->precision == FQ_PREC_MIXED) e4 = hipMalloc(&h->F32, per_dof_F32 * cs);
        hipError_t e5 = hipSuccess;
        if (h->pack) {
            e5 = hipMalloc(&h->NY, per_dof_NY * cs);
            if (e5 == hipSuccess) e5 = hipMalloc(&h->FJ, per_dof_FJ * cs);
        }
        if (e5 == hipSuccess && h->hp_nyq) {
            e5 = hipMalloc(&h->NB, per_dof_NB * cs);
            if (e5 == hipSuccess) e5 = hipMalloc(&h->Cn, per_dof_C * cs);
        }
        if (e5 == hipSuccess && per_dof_M) e5 = hipMalloc(&h->MOM, per_dof_M * cs);
        if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess || e5 != hipSuccess)
            rc = fail(FQ_ERR_ALLOC, "hipMalloc scratch failed (reduce max_chunk)");
        h->scratch = per_dof * cs;
        if (rc == FQ_SUCCESS && generic) {
            // unused batch slots of A, B and G start as zeros, not stale bits
            if (hipMemset(h->Pg, 0, per_dof_P * cs) != hipSuccess) rc = fail(FQ_ERR_HIP, "hipMemset scratch");
            if (rc == FQ_SUCCESS) rc = make_gen_plans(h);
        }
        if (rc == FQ_SUCCESS && h->nsets == 2) {
            // the second stream keeps the default priority
            if (hipStreamCreateWithFlags(&h->st2, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming) != hipSuccess)
                rc = fail(FQ_ERR_HIP, "second stream or events for FQ_FLAG_OVERLAP");
        }
    }
    if (rc != FQ_SUCCESS) {
        std::string keep = g_err;
        fq_destroy(h);
        g_err = keep;
        return rc;
    }
    *out = h;
    return FQ_SUCCESS;
}

int fq_profile_enable(fq_handle *h, int on) {
    if (!h) return fail(FQ_ERR_INVALID, "null handle");
    for (auto &r : h->ev_rec) { h->ev_pool.push_back(r.second.first); h->ev_pool.push_back(r.second.second); }
    h->ev_rec.clear();
    h->prof = on != 0;
    return FQ_SUCCESS;
}

int fq_profile_read(fq_handle *h, double *ms, int64_t *launches) {
    if (!h || !ms || !launches) return fail(FQ_ERR_INVALID, "null argument");
    for (int k = 0; k < FQ_NKERNELS; ++k) { ms[k] = 0; launches[k] = 0; }
    for (auto &r : h->ev_rec) {
        HIPCHK(hipEventSynchronize(r.second.second));
        float t = 0;
        HIPCHK(hipEventElapsedTime(&t, r.second.first, r.second.second));
        ms[r.first] += t;
        launches[r.first] += 1;
        h->ev_pool.push_back(r.second.first);
        h->ev_pool.push_back(r.second.second);
    }
    h->ev_rec.clear();
    return FQ_SUCCESS;
}

const char *fq_kernel_name(int k) { return (k >= 0 && k < FQ_NKERNELS) ? kKernelNames[k] : ""; }

int fq_chunk_dofs(const fq_handle *h, int64_t *chunk) {
    if (!h || !chunk) return fail(FQ_ERR_INVALID, "null argument");
    *chunk = h->chunk;
    return FQ_SUCCESS;
}

int fq_get_paths(const fq_handle *h, int32_t *paths) {
    if (!h || !paths) return fail(FQ_ERR_INVALID, "null argument");
    const int N = h->P.N;
    const bool one_lane = h->fused && (N == 24 || N == 32) && h->precision == FQ_PREC_MIXED && !h->two_lane;
    *paths = (h->fused ? FQ_PATH_FUSED : 0) | (one_lane ? FQ_PATH_ONE_LANE : 0) |
             (h->pack ? FQ_PATH_PAIR_PACK : 0) | (h->hpair ? FQ_PATH_GROUP_PAIRS : 0) |
             (h->nsets > 1 ? FQ_PATH_OVERLAP : 0);
    if (h->generic) *paths = FQ_PATH_LIBRARY_FFT;
    return FQ_SUCCESS;
}

int fq_destroy(fq_handle *h) {
    if (!h) return FQ_SUCCESS;
    // A handle may be released (e.g. by a garbage collector) while another
    // stream is being captured into a graph; relaxed capture mode keeps these
    // frees from invalidating that capture.
    hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
    (void)hipThreadExchangeStreamCaptureMode(&mode);
    for (auto &r : h->ev_rec) { (void)hipEventDestroy(r.second.first); (void)hipEventDestroy(r.second.second); }
    for (auto e : h->ev_pool) (void)hipEventDestroy(e);
    if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
    if (h->ev_join) (void)hipEventDestroy(h->ev_join);
    if (h->st2) (void)hipStreamDestroy(h->st2);
    for (hipfftHandle p : {h->gp_src, h->gp_items, h->gp_group, h->gp_out})
        if (p) (void)hipfftDestroy(p);
    if (h->gen_work) (void)hipFree(h->gen_work);
    for (void *p : {h->ph, (void *)h->item_src, (void *)h->prods, (void *)h->g_p0, (void *)h->g_np,
                    (void *)h->c_group, (void *)h->t_c0, (void *)h->t_cn, (void *)h->l_src, (void *)h->t_l0,
                    (void *)h->t_ln, (void *)h->W, (void *)h->ginv, (void *)h->pmass, (void *)h->F, h->Z,
                    h->Pg, h->zc, (void *)h->F32, h->NY, h->FJ, (void *)h->nyv, (void *)h->zq,
                    (void *)h->pconj, (void *)h->X, (void *)h->VN, (void *)h->e_vol, (void *)h->t_e0,
                    (void *)h->t_en, (void *)h->NB, (void *)h->Cn, (void *)h->MOM})
        if (p) (void)hipFree(p);
    (void)hipThreadExchangeStreamCaptureMode(&mode);
    delete h;
    return FQ_SUCCESS;
}

int fq_get_plan_info(const fq_handle *h, fq_plan_info *info) {
    if (!h || !info) return fail(FQ_ERR_INVALID, "null argument");
    fq::plan_info(h->P, *info);
    return FQ_SUCCESS;
}

int fq_scratch_bytes(const fq_handle *h, int64_t *bytes) {
    if (!h || !bytes) return fail(FQ_ERR_INVALID, "null argument");
    *bytes = int64_t(h->scratch);
    return FQ_SUCCESS;
}

int fq_eval(fq_handle *h, const double *f, double *q, int32_t nupts, int64_t ldim, int32_t neles,
            int64_t first_dof, int64_t ndof, void *stream) {
    if (!h) return fail(FQ_ERR_INVALID, "null handle");
    if (h->P.nsrc != 1 || h->P.ntgt != 1)
        return fail(FQ_ERR_INVALID, "fq_eval needs a single-species model; use fq_eval_multi/fq_eval_bilinear");
    if (!f || !q) return fail(FQ_ERR_INVALID, "null f or q");
    AosoaPtrs ptr{};
    ptr.f[0] = f;
    ptr.q[0] = q;
    return eval_own(h, ptr, nupts, ldim, neles, first_dof, ndof, stream);
}

int fq_eval_noalloc(fq_handle *h, const double *f, double *q, int32_t nupts, int64_t ldim, int32_t neles,
                    int64_t first_dof, int64_t ndof, void *stream) {
    if (!h) return fail(FQ_ERR_INVALID, "null handle");
    if (h->P.nsrc != 1 || h->P.ntgt != 1)
        return fail(FQ_ERR_INVALID, "fq_eval_noalloc needs a single-species model");
    if (!f || !q) return fail(FQ_ERR_INVALID, "null f or q");
    if (!h->noalloc) return fail(FQ_ERR_HIP, "static scratch not resolved on this handle's device");
    if (h->generic)
        return fail(FQ_ERR_UNSUPPORTED, "fq_eval_noalloc needs a grid with native kernels (8, 16, 24, 32)");
    // the largest batch whose regions fit the static scratch
    const size_t one = noalloc_bytes(h, 1);
    if (one > size_t(FQ_NOALLOC_SCRATCH_BYTES))
        return fail(FQ_ERR_UNSUPPORTED, "one DOF of this plan needs " + std::to_string(one) +
                                            " bytes of scratch, more than the " +
                                            std::to_string(int64_t(FQ_NOALLOC_SCRATCH_BYTES)) +
                                            " bytes of fq_eval_noalloc's static scratch");
    int64_t c = std::min<int64_t>(4096, int64_t(size_t(FQ_NOALLOC_SCRATCH_BYTES) / one));
    while (c > 1 && noalloc_bytes(h, c) > size_t(FQ_NOALLOC_SCRATCH_BYTES)) --c;
    AosoaPtrs ptr{};
    ptr.f[0] = f;
    ptr.q[0] = q;
    const Scratch sc = noalloc_carve(h, h->noalloc, c);
    return eval_impl(h, &sc, 1, c, false, ptr, nupts, ldim, neles, first_dof, ndof,
                     stream);
}

int fq_eval_multi(fq_handle *h, const double *const *f, double *const *q, int32_t nupts, int64_t ldim,
                  int32_t neles, int64_t first_dof, int64_t ndof, void *stream) {
    if (!h || !f || !q) return fail(FQ_ERR_INVALID, "null argument");
    if (h->P.model != FQ_MODEL_VSS) return fail(FQ_ERR_INVALID, "fq_eval_multi needs a vss-gll model");
    AosoaPtrs ptr{};
    for (int s = 0; s < h->P.nspcs; ++s) {
        if (!f[s] || !q[s]) return fail(FQ_ERR_INVALID, "null species pointer");
        ptr.f[s] = f[s];
        ptr.q[s] = q[s];
    }
    return eval_own(h, ptr, nupts, ldim, neles, first_dof, ndof, stream);
}

int fq_eval_bilinear(fq_handle *h, const double *f, const double *g, double *q, int32_t nupts, int64_t ldim,
                     int32_t neles, int64_t first_dof, int64_t ndof, void *stream) {
    if (!h || !f || !g || !q) return fail(FQ_ERR_INVALID, "null argument");
    if (h->P.model != FQ_MODEL_VHS_NODAL) return fail(FQ_ERR_INVALID, "fq_eval_bilinear needs vhs-gll-nodal");
    AosoaPtrs ptr{};
    ptr.f[0] = f;
    ptr.f[1] = g;
    ptr.q[0] = q;
    return eval_own(h, ptr, nupts, ldim, neles, first_dof, ndof, stream);
}

}  // extern "C"
