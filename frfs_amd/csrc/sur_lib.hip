// Fourier-layer surrogate of the collision operator: forward, backward,
// loss, Adam and the AoSoA gather / conservation-check / scatter kernels.
//
// The network (N^3 grid with N = 16 or 32, width 32, 8^3 retained modes,
// L layers):
//   v_0      = P g + p0                                 (lift, per point)
//   X_l(k)   = (1/n) sum_x v_l(x) e^{-2 pi i k.x/N}    (k_sur_dft, per channel)
//   Y_l(o,k) = sum_i X_l(i,k) R_l(i,o,k)               (k_sur_mix_lds, per mode)
//   s_l(x)   = Re sum_k Y_l(k) e^{+2 pi i k.x/N}        (k_sur_idft, per channel)
//   z_l      = Wm_l v_l + b_l + s_l,  v_{l+1} = gelu(z_l) (last layer: z_l)
//   out      = Qw . v_L + q0                            (projection)
// Backward passes reuse the same transforms: the adjoint of Re(IDFT) is the
// unscaled forward DFT and the adjoint of (1/n) DFT is (1/n) Re(IDFT).
// Weight gradients are per-workgroup sums over 512 grid points, added with
// vector-memory atomics into a caller-zeroed gradient vector.
//
// Layouts: activations are [batch][channel][x] FP32 with x = (I*N+J)*N+K;
// mode arrays are [batch][channel][k] complex with k = (px*8+py)*8+pz and
// wavenumber index kidx(p) per axis; theta holds P, p0, per layer
// (R (i,o,k) complex interleaved, Wm (o,i), b), Qw, q0.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>

#include "../../include/frfs_sur.h"

int fq_internal_fail(int code, const std::string &msg);   // fq_lib.hip

namespace {

constexpr int SW = 32, SM = 8, SM3 = SM * SM * SM;
constexpr int64_t SR = 2LL * SW * SW * SM3;          // floats of one layer's R
constexpr int64_t SLAYER = SR + SW * SW + SW;        // floats of one layer

// Grid constants of the N^3 grid (N = 16 or 32).
template <int N> struct SG {
    static_assert(N == 16 || N == 32, "grids of 16^3 and 32^3 points");
    static constexpr int N2 = N * N, N3 = N * N * N;
    static constexpr int LOG3 = N == 16 ? 12 : 15;  // log2(N^3)
};

template <int N> __device__ __forceinline__ int kidx(int p) { return p < SM / 2 ? p : p + N - SM; }

// cos(2 pi j / 32), j = 0..31.
__device__ constexpr float kCos32[32] = {
    1.0f, 0.9807852804032304f, 0.9238795325112867f, 0.8314696123025452f,
    0.7071067811865476f, 0.5555702330196023f, 0.38268343236508984f, 0.19509032201612833f,
    0.0f, -0.1950903220161282f, -0.3826834323650897f, -0.555570233019602f,
    -0.7071067811865475f, -0.8314696123025453f, -0.9238795325112867f, -0.9807852804032304f,
    -1.0f, -0.9807852804032304f, -0.9238795325112868f, -0.8314696123025455f,
    -0.7071067811865477f, -0.5555702330196022f, -0.38268343236509034f, -0.19509032201612866f,
    0.0f, 0.1950903220161283f, 0.38268343236509f, 0.5555702330196018f,
    0.7071067811865474f, 0.8314696123025452f, 0.9238795325112865f, 0.9807852804032303f};
// cos(2 pi j / N) and sin(2 pi j / N) = cos(2 pi (j - N/4) / N), N = 16 or
// 32: the table entry (j mod N) (32 / N).  Called with compile-time j
// (unrolled loops) the values fold into the instructions; otherwise they are
// uniform loads of the table.
template <int N> __device__ __forceinline__ float cosN(int j) { return kCos32[(j * (32 / N)) & 31]; }
template <int N> __device__ __forceinline__ float sinN(int j) { return cosN<N>(j - N / 4); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
// Activation storage: FP32, or FP16 / BF16 for inference (arithmetic stays
// FP32; conversions round to nearest even).
template <typename TA> __device__ __forceinline__ float ldA(const TA *p, int64_t i) { return float(p[i]); }
template <typename TA> __device__ __forceinline__ void stA(TA *p, int64_t i, float v) { p[i] = TA(v); }
// two neighbouring values (even index), one 4- or 8-byte access
template <typename TA> struct alignas(2 * sizeof(TA)) Pair { TA a, b; };
template <typename TA> __device__ __forceinline__ float2 ld2(const TA *p, int64_t i) {
    const Pair<TA> h = *reinterpret_cast<const Pair<TA> *>(p + i);
    return make_float2(float(h.a), float(h.b));
}
template <typename TA> __device__ __forceinline__ void st2(TA *p, int64_t i, float2 v) {
    Pair<TA> h;
    h.a = TA(v.x);
    h.b = TA(v.y);
    *reinterpret_cast<Pair<TA> *>(p + i) = h;
}
// eight 16-bit values, one 16-byte store
template <typename TA> struct alignas(16) Pack8 { TA v[8]; };

__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * __expf(-0.5f * z * z);
}

// X[slice] = scale * sum_x v[slice](x) e^{-i k.x} over the retained modes.
// 16^3 grid: one workgroup per (batch, channel) slice; separable z, y, x stages in LDS,
// each thread reading its inputs once and producing all the modes of its
// line with compile-time twiddles.
template <typename TA>
__global__ __launch_bounds__(256) void k_sur_dft16(const TA *__restrict__ v, float scale, float2 *__restrict__ X) {
    __shared__ float vin[SG<16>::N2 * 17];
    __shared__ float2 A[SG<16>::N2 * 9];            // [I][J][pz]
    __shared__ float2 Bs[16 * SM * SM];      // [I][py][pz]
    const int t = threadIdx.x;
    const int64_t slice = blockIdx.x;
    const TA *src = v + slice * SG<16>::N3;
#pragma unroll
    for (int r = 0; r < SG<16>::N3 / 256; ++r) {
        const int e = t + 256 * r;
        vin[(e >> 4) * 17 + (e & 15)] = ldA(src, e);
    }
    __syncthreads();
    {
        float a[16];
#pragma unroll
        for (int K = 0; K < 16; ++K) a[K] = vin[t * 17 + K];
#pragma unroll
        for (int p = 0; p < SM; ++p) {
            const int k = kidx<16>(p);
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int K = 0; K < 16; ++K) {
                re += a[K] * cosN<16>(k * K);
                im -= a[K] * sinN<16>(k * K);
            }
            A[t * 9 + p] = make_float2(re, im);
        }
    }
    __syncthreads();
    // y: thread t < 128 owns (I, pz) = (t / 8, t % 8): its 16 values once, the
    // 8 y modes with compile-time twiddles
    if (t < 16 * SM) {
        const int I = t >> 3, pz = t & 7;
        float2 a[16];
#pragma unroll
        for (int J = 0; J < 16; ++J) a[J] = A[(I * 16 + J) * 9 + pz];
#pragma unroll
        for (int py = 0; py < SM; ++py) {
            const int k = kidx<16>(py);
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int J = 0; J < 16; ++J) {
                const float c = cosN<16>(k * J), sn = sinN<16>(k * J);
                re += a[J].x * c + a[J].y * sn;
                im += a[J].y * c - a[J].x * sn;
            }
            Bs[(I * SM + py) * SM + pz] = make_float2(re, im);
        }
    }
    __syncthreads();
    // x: thread t < 64 owns (py, pz) = (t / 8, t % 8): 16 values, the 8 x modes
    if (t < SM * SM) {
        float2 a[16];
#pragma unroll
        for (int I = 0; I < 16; ++I) a[I] = Bs[I * SM * SM + t];
#pragma unroll
        for (int px = 0; px < SM; ++px) {
            const int k = kidx<16>(px);
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int I = 0; I < 16; ++I) {
                const float c = cosN<16>(k * I), sn = sinN<16>(k * I);
                re += a[I].x * c + a[I].y * sn;
                im += a[I].y * c - a[I].x * sn;
            }
            X[slice * SM3 + px * SM * SM + t] = make_float2(scale * re, scale * im);
        }
    }
}

// out[slice](x) = (accumulate ? out : 0) + scale * Re sum_k Y[slice](k) e^{+i k.x}
// (16^3 grid).
template <typename TA>
__global__ __launch_bounds__(256) void k_sur_idft16(const float2 *__restrict__ Y, float scale, int accumulate,
                                                  TA *__restrict__ out) {
    __shared__ float2 Ys[SM3];               // [px][py][pz]
    __shared__ float2 T1[16 * SM * SM];      // [I][py][pz]
    __shared__ float2 T2[SG<16>::N2 * 9];           // [I][J][pz]
    const int t = threadIdx.x;
    const int64_t slice = blockIdx.x;
    Ys[t] = Y[slice * SM3 + t];
    Ys[t + 256] = Y[slice * SM3 + t + 256];
    __syncthreads();
    // x: thread t < 64 owns (py, pz) = (t / 8, t % 8): its 8 modes once, the 16
    // grid values I with compile-time twiddles
    if (t < SM * SM) {
        float2 a[SM];
#pragma unroll
        for (int px = 0; px < SM; ++px) a[px] = Ys[px * SM * SM + t];
#pragma unroll
        for (int I = 0; I < 16; ++I) {
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int px = 0; px < SM; ++px) {
                const float c = cosN<16>(kidx<16>(px) * I), sn = sinN<16>(kidx<16>(px) * I);
                re += a[px].x * c - a[px].y * sn;
                im += a[px].x * sn + a[px].y * c;
            }
            T1[I * SM * SM + t] = make_float2(re, im);
        }
    }
    __syncthreads();
    // y: thread t < 128 owns (I, pz) = (t / 8, t % 8): 8 values, the 16 rows J
    if (t < 16 * SM) {
        const int I = t >> 3, pz = t & 7;
        float2 a[SM];
#pragma unroll
        for (int py = 0; py < SM; ++py) a[py] = T1[(I * SM + py) * SM + pz];
#pragma unroll
        for (int J = 0; J < 16; ++J) {
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int py = 0; py < SM; ++py) {
                const float c = cosN<16>(kidx<16>(py) * J), sn = sinN<16>(kidx<16>(py) * J);
                re += a[py].x * c - a[py].y * sn;
                im += a[py].x * sn + a[py].y * c;
            }
            T2[(I * 16 + J) * 9 + pz] = make_float2(re, im);
        }
    }
    __syncthreads();
    // z: thread t owns grid row (I, J) = t: its 8 modes from LDS once, the 16
    // outputs with compile-time twiddles, written as four 16-byte stores
    float2 a[SM];
#pragma unroll
    for (int pz = 0; pz < SM; ++pz) a[pz] = T2[t * 9 + pz];
    float o[16];
#pragma unroll
    for (int K = 0; K < 16; ++K) {
        float sum = 0.f;
#pragma unroll
        for (int pz = 0; pz < SM; ++pz) sum += a[pz].x * cosN<16>(kidx<16>(pz) * K) - a[pz].y * sinN<16>(kidx<16>(pz) * K);
        o[K] = scale * sum;
    }
    if constexpr (std::is_same<TA, float>::value) {
        float4 *dst = reinterpret_cast<float4 *>(out + slice * SG<16>::N3 + t * 16);
#pragma unroll
        for (int q = 0; q < 16 / 4; ++q) {
            float4 w = make_float4(o[4 * q], o[4 * q + 1], o[4 * q + 2], o[4 * q + 3]);
            if (accumulate) {
                const float4 p0 = dst[q];
                w.x += p0.x;
                w.y += p0.y;
                w.z += p0.z;
                w.w += p0.w;
            }
            dst[q] = w;
        }
    } else {                                  // 16-bit inference output (never accumulated)
        static_assert(sizeof(TA) == 2, "16-bit storage");
        Pack8<TA> *dst = reinterpret_cast<Pack8<TA> *>(out + slice * SG<16>::N3 + t * 16);
#pragma unroll
        for (int q = 0; q < 16 / 8; ++q) {
            Pack8<TA> w;
#pragma unroll
            for (int j = 0; j < 8; ++j) w.v[j] = TA(o[8 * q + j]);
            dst[q] = w;
        }
    }
}

// The same transform on the 32^3 grid.  One workgroup of 512 threads per
// slice.  The z stage runs in two passes of 512 grid rows: each pass loads
// its rows with coalesced 16-byte loads into LDS (vin), then thread t turns
// row t into its 8 modes (A, 76 KB).  The y and x stages then work from LDS
// as on the 16^3 grid (Bs reuses vin).  143 KB of LDS: one workgroup per CU.
constexpr int kDft32Threads = 512;
// LDS index of the 8 modes of grid row r = I*32 + J in the 32^3 transforms:
// rows padded to 9 and the 32 rows of one I to 296 float2 values, so that
// both the per-row accesses (z stage) and the per-(I, pz) accesses (y stage)
// spread over the LDS banks.
constexpr int kRowsI = 296;
__device__ __forceinline__ int aidx32(int I, int J, int pz) { return I * kRowsI + J * 9 + pz; }
template <typename TA>
__global__ __launch_bounds__(kDft32Threads) void k_sur_dft32(const TA *__restrict__ v, float scale,
                                                           float2 *__restrict__ X) {
    constexpr int N = 32, N3 = SG<32>::N3, T = kDft32Threads, RS = N + 1;
    __shared__ float2 A[N * kRowsI];          // [I][J][pz], aidx32
    __shared__ float vin[T * RS];             // one pass of rows; then Bs [I][py][pz]
    float2 *Bs = reinterpret_cast<float2 *>(vin);
    const int t = threadIdx.x;
    const int64_t slice = blockIdx.x;
    const TA *src = v + slice * N3;
#pragma unroll 1
    for (int q = 0; q < N * N / T; ++q) {
        const TA *rows = src + int64_t(q) * T * N;
        if constexpr (std::is_same<TA, float>::value) {
#pragma unroll
            for (int w = 0; w < T * N / 4 / T; ++w) {
                const int e4 = t + T * w;                 // float4 index in the pass
                const float4 x = reinterpret_cast<const float4 *>(rows)[e4];
                float *d = vin + (e4 >> 3) * RS + (e4 & 7) * 4;
                d[0] = x.x; d[1] = x.y; d[2] = x.z; d[3] = x.w;
            }
        } else {
#pragma unroll
            for (int w = 0; w < T * N / 8 / T; ++w) {
                const int e8 = t + T * w;                 // 8-value pack index in the pass
                const Pack8<TA> x = reinterpret_cast<const Pack8<TA> *>(rows)[e8];
                float *d = vin + (e8 >> 2) * RS + (e8 & 3) * 8;
#pragma unroll
                for (int j = 0; j < 8; ++j) d[j] = float(x.v[j]);
            }
        }
        __syncthreads();
        float a[N];
#pragma unroll
        for (int K = 0; K < N; ++K) a[K] = vin[t * RS + K];
        const int r = q * T + t;
#pragma unroll
        for (int pz = 0; pz < SM; ++pz) {
            const int k = kidx<N>(pz);
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int K = 0; K < N; ++K) {
                re += a[K] * cosN<N>(k * K);
                im -= a[K] * sinN<N>(k * K);
            }
            A[aidx32(r >> 5, r & 31, pz)] = make_float2(re, im);
        }
        __syncthreads();
    }
    // y: thread t < 256 owns (I, pz) = (t / 8, t % 8): 32 values, the 8 y modes
    if (t < N * SM) {
        const int I = t >> 3, pz = t & 7;
        float2 a[N];
#pragma unroll
        for (int J = 0; J < N; ++J) a[J] = A[aidx32(I, J, pz)];
#pragma unroll
        for (int py = 0; py < SM; ++py) {
            const int k = kidx<N>(py);
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int J = 0; J < N; ++J) {
                const float c = cosN<N>(k * J), sn = sinN<N>(k * J);
                re += a[J].x * c + a[J].y * sn;
                im += a[J].y * c - a[J].x * sn;
            }
            Bs[(I * SM + py) * SM + pz] = make_float2(re, im);
        }
    }
    __syncthreads();
    // x: thread t < 64 owns (py, pz): 32 values, the 8 x modes
    if (t < SM * SM) {
        float2 a[N];
#pragma unroll
        for (int I = 0; I < N; ++I) a[I] = Bs[I * SM * SM + t];
#pragma unroll
        for (int px = 0; px < SM; ++px) {
            const int k = kidx<N>(px);
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int I = 0; I < N; ++I) {
                const float c = cosN<N>(k * I), sn = sinN<N>(k * I);
                re += a[I].x * c + a[I].y * sn;
                im += a[I].y * c - a[I].x * sn;
            }
            X[slice * SM3 + px * SM * SM + t] = make_float2(scale * re, scale * im);
        }
    }
}

// The inverse transform on the 32^3 grid, 512 threads per slice: x and y
// stages in LDS (the mode block Ys shares its space with the y-stage output
// T2, and T1 with the output staging), then two passes of 512 grid rows:
// thread t computes row t's 32 outputs into LDS (ob) and the workgroup writes
// the pass with coalesced 16-byte stores.
template <typename TA>
__global__ __launch_bounds__(kDft32Threads) void k_sur_idft32(const float2 *__restrict__ Y, float scale,
                                                            int accumulate, TA *__restrict__ out) {
    constexpr int N = 32, N3 = SG<32>::N3, T = kDft32Threads, RS = N + 1;
    __shared__ float2 T2[N * kRowsI];         // [I][J][pz] (aidx32); first the modes [px][py][pz]
    __shared__ float ob[T * RS];              // T1 [I][py][pz], then one pass of output rows
    float2 *Ys = T2;
    float2 *T1 = reinterpret_cast<float2 *>(ob);
    const int t = threadIdx.x;
    const int64_t slice = blockIdx.x;
    Ys[t] = Y[slice * SM3 + t];
    __syncthreads();
    // x: thread t < 64 owns (py, pz): 8 modes, the 32 grid values I
    if (t < SM * SM) {
        float2 a[SM];
#pragma unroll
        for (int px = 0; px < SM; ++px) a[px] = Ys[px * SM * SM + t];
#pragma unroll
        for (int I = 0; I < N; ++I) {
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int px = 0; px < SM; ++px) {
                const float c = cosN<N>(kidx<N>(px) * I), sn = sinN<N>(kidx<N>(px) * I);
                re += a[px].x * c - a[px].y * sn;
                im += a[px].x * sn + a[px].y * c;
            }
            T1[I * SM * SM + t] = make_float2(re, im);
        }
    }
    __syncthreads();
    // y: thread t < 256 owns (I, pz) = (t / 8, t % 8): 8 values, the 32 rows J
    if (t < N * SM) {
        const int I = t >> 3, pz = t & 7;
        float2 a[SM];
#pragma unroll
        for (int py = 0; py < SM; ++py) a[py] = T1[(I * SM + py) * SM + pz];
#pragma unroll
        for (int J = 0; J < N; ++J) {
            float re = 0.f, im = 0.f;
#pragma unroll
            for (int py = 0; py < SM; ++py) {
                const float c = cosN<N>(kidx<N>(py) * J), sn = sinN<N>(kidx<N>(py) * J);
                re += a[py].x * c - a[py].y * sn;
                im += a[py].x * sn + a[py].y * c;
            }
            T2[aidx32(I, J, pz)] = make_float2(re, im);
        }
    }
    __syncthreads();
    // z: two passes of 512 rows
#pragma unroll 1
    for (int q = 0; q < N * N / T; ++q) {
        const int r = q * T + t;
        float2 a[SM];
#pragma unroll
        for (int pz = 0; pz < SM; ++pz) a[pz] = T2[aidx32(r >> 5, r & 31, pz)];
#pragma unroll
        for (int K = 0; K < N; ++K) {
            float sum = 0.f;
#pragma unroll
            for (int pz = 0; pz < SM; ++pz)
                sum += a[pz].x * cosN<N>(kidx<N>(pz) * K) - a[pz].y * sinN<N>(kidx<N>(pz) * K);
            ob[t * RS + K] = scale * sum;
        }
        __syncthreads();
        TA *rows = out + slice * N3 + int64_t(q) * T * N;
        if constexpr (std::is_same<TA, float>::value) {
#pragma unroll
            for (int w = 0; w < T * N / 4 / T; ++w) {
                const int e4 = t + T * w;
                const float *sp = ob + (e4 >> 3) * RS + (e4 & 7) * 4;
                float4 x = make_float4(sp[0], sp[1], sp[2], sp[3]);
                float4 *dp = reinterpret_cast<float4 *>(rows) + e4;
                if (accumulate) {
                    const float4 p0 = *dp;
                    x.x += p0.x;
                    x.y += p0.y;
                    x.z += p0.z;
                    x.w += p0.w;
                }
                *dp = x;
            }
        } else {                              // 16-bit inference output (never accumulated)
#pragma unroll
            for (int w = 0; w < T * N / 8 / T; ++w) {
                const int e8 = t + T * w;
                const float *sp = ob + (e8 >> 2) * RS + (e8 & 3) * 8;
                Pack8<TA> x;
#pragma unroll
                for (int j = 0; j < 8; ++j) x.v[j] = TA(sp[j]);
                reinterpret_cast<Pack8<TA> *>(rows)[e8] = x;
            }
        }
        __syncthreads();
    }
}

// Launchers of the per-slice transforms on the N^3 grid.
template <int N, typename TA>
void launch_dft(const TA *v, float scale, float2 *X, unsigned slices, hipStream_t st) {
    if constexpr (N == 16) hipLaunchKernelGGL(k_sur_dft16<TA>, dim3(slices), dim3(256), 0, st, v, scale, X);
    else hipLaunchKernelGGL(k_sur_dft32<TA>, dim3(slices), dim3(kDft32Threads), 0, st, v, scale, X);
}
template <int N, typename TA>
void launch_idft(const float2 *Y, float scale, int accumulate, TA *out, unsigned slices, hipStream_t st) {
    if constexpr (N == 16) hipLaunchKernelGGL(k_sur_idft16<TA>, dim3(slices), dim3(256), 0, st, Y, scale, accumulate, out);
    else hipLaunchKernelGGL(k_sur_idft32<TA>, dim3(slices), dim3(kDft32Threads), 0, st, Y, scale, accumulate, out);
}

// Weight gradient of the mixing: gR(i,o,k) += sum_b conj X(b,i,k) gY(b,o,k).
// One thread per (i, o, k), k fastest, so every access is coalesced.
__global__ __launch_bounds__(256) void k_sur_mix_grad(const float2 *__restrict__ X, const float2 *__restrict__ gY,
                                                      float2 *__restrict__ gR, int64_t batch) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= int64_t(SW) * SW * SM3) return;
    const int k = int(idx & (SM3 - 1));
    const int o = int((idx >> 9) & (SW - 1));
    const int i = int(idx >> 14);
    float2 acc = make_float2(0.f, 0.f);
    for (int64_t b = 0; b < batch; ++b) {
        const float2 x = X[(b * SW + i) * SM3 + k];
        const float2 g = gY[(b * SW + o) * SM3 + k];
        acc.x += x.x * g.x + x.y * g.y;
        acc.y += x.x * g.y - x.y * g.x;
    }
    gR[idx].x += acc.x;
    gR[idx].y += acc.y;
}

// The forward and input-gradient mixing:
// MODE 0: Y(b,o,k) = sum_i X(b,i,k) R(i,o,k);
// MODE 1: gX(b,i,k) = sum_o gY(b,o,k) conj R(i,o,k).
// A workgroup takes 8 consecutive modes and 32 DOFs and stages the 8 modes'
// 32 x 32 weights in LDS (64 KB) once; thread (mode m = t % 8, DOF d = t / 8)
// keeps its 32 outputs in registers.  Within a wave the eight DOFs of a mode
// read the same LDS word (a broadcast), so the weight reads are conflict-free.
template <int MODE>
__global__ __launch_bounds__(256) void k_sur_mix_lds(const float2 *__restrict__ a, const float2 *__restrict__ R,
                                                     float2 *__restrict__ c, int64_t batch) {
    __shared__ float2 Rl[SW * SW * 8];               // [(i*32 + o)*8 + m]
    const int t = threadIdx.x;
    const int k0 = blockIdx.x * 8;
    for (int e = t; e < SW * SW * 8; e += 256) Rl[e] = R[int64_t(e >> 3) * SM3 + k0 + (e & 7)];
    __syncthreads();
    const int m = t & 7;
    const int64_t b = int64_t(blockIdx.y) * 32 + (t >> 3);
    if (b >= batch) return;
    const int k = k0 + m;
    float2 acc[SW];
#pragma unroll
    for (int j = 0; j < SW; ++j) acc[j] = make_float2(0.f, 0.f);
    // all 32 inputs in flight at once: occupancy is set by the LDS tile (two
    // waves per SIMD), so the registers are free
    float2 xs[SW];
#pragma unroll
    for (int j = 0; j < SW; ++j) xs[j] = a[(b * SW + j) * SM3 + k];
#pragma unroll
    for (int j = 0; j < SW; ++j) {
        const float2 x = xs[j];
#pragma unroll
        for (int q = 0; q < SW; ++q) {
            if constexpr (MODE == 0) {      // j = i, q = o
                const float2 r = Rl[(j * SW + q) * 8 + m];
                acc[q].x += x.x * r.x - x.y * r.y;
                acc[q].y += x.x * r.y + x.y * r.x;
            } else {                        // j = o, q = i
                const float2 r = Rl[(q * SW + j) * 8 + m];
                acc[q].x += x.x * r.x + x.y * r.y;
                acc[q].y += x.y * r.x - x.x * r.y;
            }
        }
    }
#pragma unroll
    for (int j = 0; j < SW; ++j) c[(b * SW + j) * SM3 + k] = acc[j];
}

// v_0(b, c, x) = P_c g(b, x) + p0_c
template <int N, typename TA>
__global__ __launch_bounds__(256) void k_sur_lift(const float *__restrict__ g, const float *__restrict__ th,
                                                  TA *__restrict__ v, int64_t total) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= total) return;
    const int64_t b = idx >> SG<N>::LOG3;
    const int x = int(idx & (SG<N>::N3 - 1));
    const float gv = g[idx];
#pragma unroll 8
    for (int c = 0; c < SW; ++c) stA(v, (b * SW + c) * SG<N>::N3 + x, th[c] * gv + th[SW + c]);
}

// z = Wm v + b + s; vout = act ? gelu(z) : z (z stored when zout is set).
// One thread per two neighbouring grid points (total = points, even).
template <int N, typename TA>
__global__ __launch_bounds__(256) void k_sur_pw(const TA *__restrict__ v, const TA *__restrict__ s,
                                                const float *__restrict__ lw, int act, TA *__restrict__ zout,
                                                TA *__restrict__ vout, int64_t total) {
    // the weights are the same for every lane: read at uniform addresses, they
    // arrive through scalar loads instead of LDS
    const int64_t idx = 2 * (int64_t(blockIdx.x) * 256 + threadIdx.x);
    if (idx >= total) return;
    const int64_t b = idx >> SG<N>::LOG3;
    const int x = int(idx & (SG<N>::N3 - 1));
    float2 vi[SW];
#pragma unroll
    for (int i = 0; i < SW; ++i) vi[i] = ld2(v, (b * SW + i) * SG<N>::N3 + x);
#pragma unroll 4
    for (int o = 0; o < SW; ++o) {
        const int64_t at = (b * SW + o) * SG<N>::N3 + x;
        const float2 sv = ld2(s, at);
        float z0 = lw[SW * SW + o] + sv.x, z1 = lw[SW * SW + o] + sv.y;
#pragma unroll
        for (int i = 0; i < SW; ++i) {
            z0 += lw[o * SW + i] * vi[i].x;
            z1 += lw[o * SW + i] * vi[i].y;
        }
        if (zout) st2(zout, at, make_float2(z0, z1));
        st2(vout, at, act ? make_float2(gelu(z0), gelu(z1)) : make_float2(z0, z1));
    }
}

// out(b, x) = Qw . v_L(b, :, x) + q0
template <int N, typename TA>
__global__ __launch_bounds__(256) void k_sur_proj(const TA *__restrict__ v, const float *__restrict__ qw,
                                                  float *__restrict__ out, int64_t total) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= total) return;
    const int64_t b = idx >> SG<N>::LOG3;
    const int x = int(idx & (SG<N>::N3 - 1));
    float s = qw[SW];
#pragma unroll 8
    for (int c = 0; c < SW; ++c) s += qw[c] * ldA(v, (b * SW + c) * SG<N>::N3 + x);
    out[idx] = s;
}

// gv(b, c, x) = Qw_c gout(b, x)
template <int N>
__global__ __launch_bounds__(256) void k_sur_proj_bwd(const float *__restrict__ gout, const float *__restrict__ qw,
                                                      float *__restrict__ gv, int64_t total) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= total) return;
    const int64_t b = idx >> SG<N>::LOG3;
    const int x = int(idx & (SG<N>::N3 - 1));
    const float g = gout[idx];
#pragma unroll 8
    for (int c = 0; c < SW; ++c) gv[(b * SW + c) * SG<N>::N3 + x] = qw[c] * g;
}

// gz = gv * gelu'(z) (or gv for the last layer), written over gv;
// gvin = Wm^T gz.
template <int N, typename TA>
__global__ __launch_bounds__(256) void k_sur_pw_bwd(float *__restrict__ gv, const TA *__restrict__ z,
                                                    const float *__restrict__ lw, int act, float *__restrict__ gvin,
                                                    int64_t total) {
    // weight rows are read at uniform addresses (scalar loads), row o scaled
    // by gz_o and added into the 32 input-channel sums
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= total) return;
    const int64_t b = idx >> SG<N>::LOG3;
    const int x = int(idx & (SG<N>::N3 - 1));
    float a[SW];
#pragma unroll
    for (int i = 0; i < SW; ++i) a[i] = 0.f;
#pragma unroll 4
    for (int o = 0; o < SW; ++o) {
        const int64_t at = (b * SW + o) * SG<N>::N3 + x;
        const float g = gv[at];
        const float gz = act ? g * gelu_grad(ldA(z, at)) : g;
        gv[at] = gz;
#pragma unroll
        for (int i = 0; i < SW; ++i) a[i] += lw[o * SW + i] * gz;
    }
#pragma unroll
    for (int i = 0; i < SW; ++i) gvin[(b * SW + i) * SG<N>::N3 + x] = a[i];
}

// G(o, i) += sum_{b, x} A(b, o, x) V(b, i, x);  gb(o) += sum_{b, x} A(b, o, x).
// Grid T * batch with T = N^3 / 512 tiles per DOF: workgroup w sums the 512
// points from (w % T) * 512 of DOF w / T in LDS tiles of 64.
template <int N, int COUT, int CIN, typename TV>
__global__ __launch_bounds__(256) void k_sur_wgrad(const float *__restrict__ A, const TV *__restrict__ V,
                                                   float *__restrict__ G, float *__restrict__ gb) {
    constexpr int NO = COUT * CIN;
    constexpr int PER = (NO + 255) / 256;
    __shared__ float As[COUT][65];
    __shared__ float Vs[CIN][65];
    const int t = threadIdx.x;
    constexpr int TILES = SG<N>::N3 / 512;
    const int64_t b = blockIdx.x / TILES;
    const int x0 = int(blockIdx.x % TILES) * 512;
    float acc[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) acc[q] = 0.f;
    float accb = 0.f;
    for (int p0 = x0; p0 < x0 + 512; p0 += 64) {
        for (int e = t; e < COUT * 64; e += 256) As[e >> 6][e & 63] = A[(b * COUT + (e >> 6)) * SG<N>::N3 + p0 + (e & 63)];
        for (int e = t; e < CIN * 64; e += 256) Vs[e >> 6][e & 63] = ldA(V, (b * CIN + (e >> 6)) * SG<N>::N3 + p0 + (e & 63));
        __syncthreads();
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int j = t + 256 * q;
            if (j < NO) {
                const int o = j / CIN, i = j % CIN;
                float s = 0.f;
#pragma unroll 16
                for (int p = 0; p < 64; ++p) s += As[o][p] * Vs[i][p];
                acc[q] += s;
            }
        }
        if (t < COUT) {
#pragma unroll 16
            for (int p = 0; p < 64; ++p) accb += As[t][p];
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int j = t + 256 * q;
        if (j < NO) atomicAdd(&G[j], acc[q]);
    }
    if (t < COUT) atomicAdd(&gb[t], accb);
}

// Per-DOF workgroups are dealt round-robin to the eight XCDs; workgroup w
// takes DOF (w % 8) * per + w / 8 so that each XCD works through a contiguous
// DOF range and neighbouring elements, which share AoSoA cache lines, meet in
// one L2.  Launch 8 * per workgroups; the result is count or more when idle.
__device__ __forceinline__ int64_t xcd_slot(int64_t count) {
    const int64_t per = (count + 7) / 8;
    return int64_t(blockIdx.x & 7) * per + int64_t(blockIdx.x >> 3);
}

inline unsigned xcd_grid(int64_t count) { return unsigned(8 * ((count + 7) / 8)); }

__device__ __forceinline__ double block_sum(double v, double *red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// gout = 2 (out - target) / count; sums[0] += sum (out - target)^2,
// sums[1] += sum target^2.
__global__ __launch_bounds__(256) void k_sur_loss(const float *__restrict__ out, const float *__restrict__ tgt,
                                                  float *__restrict__ gout, double *__restrict__ sums, int64_t count) {
    __shared__ double red[256];
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    double d2 = 0, t2 = 0;
    if (idx < count) {
        const float d = out[idx] - tgt[idx];
        gout[idx] = float(2.0 / double(count)) * d;
        d2 = double(d) * d;
        t2 = double(tgt[idx]) * tgt[idx];
    }
    const double s0 = block_sum(d2, red);
    const double s1 = block_sum(t2, red);
    if (threadIdx.x == 0) {
        atomicAdd(&sums[0], s0);
        atomicAdd(&sums[1], s1);
    }
}

__global__ __launch_bounds__(256) void k_sur_adam(float *__restrict__ th, const float *__restrict__ g,
                                                  float *__restrict__ m, float *__restrict__ v, int64_t n, float lr,
                                                  float b1, float b2, float eps, float c1, float c2) {
    const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    th[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// dst(b, x) = src(dof, x) / (rho_b^power * scale), rho_b = rho_in[b] or
// sum_x src(dof, x) cw (written to rho_out).  One workgroup per DOF.
__global__ __launch_bounds__(256) void k_sur_gather(int n, double cw, double scale, int power,
                                                    const double *__restrict__ src, int nupts, int64_t ldim,
                                                    int64_t first, const double *__restrict__ rho_in,
                                                    double *__restrict__ rho_out, float *__restrict__ dst,
                                                    int64_t ndof) {
    __shared__ double red[256];
    const int64_t b = xcd_slot(ndof);
    if (b >= ndof) return;
    const int64_t dof = first + b;
    const int64_t e = dof / nupts;
    const int u = int(dof % nupts);
    const double *base = src + u * ldim + (e / 32) * int64_t(n) * 32 + e % 32;
    double rho;
    if (rho_in) {
        rho = rho_in[b];
    } else {
        double s = 0;
        for (int x = threadIdx.x; x < n; x += 256) s += base[int64_t(x) * 32];
        rho = block_sum(s, red) * cw;
        if (rho_out && threadIdx.x == 0) rho_out[b] = rho;
    }
    double den = scale;
    for (int k = 0; k < power; ++k) den *= rho;
    const double inv = den != 0.0 ? 1.0 / den : 0.0;
    for (int x = threadIdx.x; x < n; x += 256) dst[b * n + x] = float(base[int64_t(x) * 32] * inv);
}

struct SurGram {
    double ginv[25];
    double gdiag[5];
};

// Conservation check and scatter.  For DOF b: moments m_j = cw sum_x q phi_j
// of the collision invariants phi = (1, cx, cy, cz, |c|^2); defect = max_j
// |m_j| / sqrt(G_jj cw sum q^2); DOFs with defect > tol are flagged and not
// written; the rest get Q = rho_b^2 scale (q - sum_j lambda_j phi_j) with
// lambda = G^{-1} m (the least-squares conservation projection).
__global__ __launch_bounds__(256) void k_sur_finish(int N, double c0, double dc, double cw, double scale, double tol,
                                                    const float *__restrict__ q, const double *__restrict__ rho,
                                                    SurGram gm, double *__restrict__ dst, int nupts, int64_t ldim,
                                                    int64_t first, double *__restrict__ defect,
                                                    int32_t *__restrict__ flag, int64_t ndof) {
    __shared__ double red[256];
    const int n = N * N * N;
    const int64_t b = xcd_slot(ndof);
    if (b >= ndof) return;
    const float *qb = q + b * n;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int x = threadIdx.x; x < n; x += 256) {
        const double cx = c0 + dc * (x / (N * N)), cy = c0 + dc * ((x / N) % N), cz = c0 + dc * (x % N);
        const double v = qb[x];
        acc[0] += v;
        acc[1] += v * cx;
        acc[2] += v * cy;
        acc[3] += v * cz;
        acc[4] += v * (cx * cx + cy * cy + cz * cz);
        acc[5] += v * v;
    }
    double mom[5];
    for (int j = 0; j < 5; ++j) mom[j] = block_sum(acc[j], red) * cw;
    const double nq = sqrt(block_sum(acc[5], red) * cw);
    double d = 0;
    for (int j = 0; j < 5; ++j) {
        const double cj = nq > 0 ? fabs(mom[j]) / (sqrt(gm.gdiag[j]) * nq) : 0.0;
        d = cj > d ? cj : d;
    }
    const bool bad = !(d <= tol);
    if (threadIdx.x == 0) {
        defect[b] = d;
        flag[b] = bad ? 1 : 0;
    }
    if (bad) return;
    double lam[5];
    for (int j = 0; j < 5; ++j) {
        double s = 0;
        for (int k = 0; k < 5; ++k) s += gm.ginv[j * 5 + k] * mom[k];
        lam[j] = s;
    }
    const double r = rho[b];
    const double f = r * r * scale;
    const int64_t dof = first + b;
    const int64_t e = dof / nupts;
    const int u = int(dof % nupts);
    double *out = dst + u * ldim + (e / 32) * int64_t(n) * 32 + e % 32;
    for (int x = threadIdx.x; x < n; x += 256) {
        const double cx = c0 + dc * (x / (N * N)), cy = c0 + dc * ((x / N) % N), cz = c0 + dc * (x % N);
        const double p = lam[0] + lam[1] * cx + lam[2] * cy + lam[3] * cz + lam[4] * (cx * cx + cy * cy + cz * cz);
        out[int64_t(x) * 32] = f * (double(qb[x]) - p);
    }
}

// Conservation penalty of a training batch (one workgroup per DOF b): the
// invariant component u_b = Phi^T G^{-1} Phi out_b cw of the network output,
// the orthogonal projection of out_b onto span{1, cx, cy, cz, |c|^2} (G the
// Gram matrix Phi Phi^T cw); gout_b += (2 lambda / count) u_b and sums[2] +=
// |u_b|^2, the gradient and value of lambda |u|^2 / count added to the mean
// square loss.  The targets conserve (they are projected), so this weights
// the error in the invariant directions lambda + 1 times.
__global__ __launch_bounds__(256) void k_sur_cons(int N, double c0, double dc, double cw, SurGram gm, double coef,
                                                  const float *__restrict__ out, float *__restrict__ gout,
                                                  double *__restrict__ sums) {
    __shared__ double red[256];
    const int n = N * N * N;
    const int64_t b = blockIdx.x;
    const float *ob = out + b * n;
    float *gb = gout + b * n;
    double acc[5] = {0, 0, 0, 0, 0};
    for (int x = threadIdx.x; x < n; x += 256) {
        const double cx = c0 + dc * (x / (N * N)), cy = c0 + dc * ((x / N) % N), cz = c0 + dc * (x % N);
        const double v = ob[x];
        acc[0] += v;
        acc[1] += v * cx;
        acc[2] += v * cy;
        acc[3] += v * cz;
        acc[4] += v * (cx * cx + cy * cy + cz * cz);
    }
    double mom[5], lam[5];
    for (int j = 0; j < 5; ++j) mom[j] = block_sum(acc[j], red) * cw;
    for (int j = 0; j < 5; ++j) {
        double t = 0;
        for (int k = 0; k < 5; ++k) t += gm.ginv[j * 5 + k] * mom[k];
        lam[j] = t;
    }
    double uu = 0;
    for (int x = threadIdx.x; x < n; x += 256) {
        const double cx = c0 + dc * (x / (N * N)), cy = c0 + dc * ((x / N) % N), cz = c0 + dc * (x % N);
        const double u = lam[0] + lam[1] * cx + lam[2] * cy + lam[3] * cz + lam[4] * (cx * cx + cy * cy + cz * cz);
        gb[x] += float(coef * u);
        uu += u * u;
    }
    const double tot = block_sum(uu, red);
    if (threadIdx.x == 0) atomicAdd(&sums[2], tot);
}

// ---------------------------------------------------------------------------
// Inference with the lift folded into the first layer and the projection into
// the last (layers >= 2).  With Ghat = DFT(g)/n (one channel):
//   first layer: X_0(i,k) = P_i Ghat(k) + p0_i [k = 0], so
//     Y_0(o,k) = Ghat(k) Rp(o,k) + [k = 0] c0(o),  Rp = sum_i P_i R_0(i,o,.),
//     c0 = sum_i p0_i R_0(i,o,0);  z_0 = A g + a0 + s_0,  A = Wm_0 P,
//     a0 = Wm_0 p0 + b_0;
//   last layer: Qw.s = Re IDFT(sum_o Qw_o Y(o,.)) and sum_o Qw_o Y(o,k) =
//     sum_i X(i,k) Rq(i,k), Rq = sum_o Qw_o R(i,o,.), so
//     out = QW.v + qb + Re IDFT(Y1),  QW = Qw Wm,  qb = Qw.b + q0.
// The folded weights (k_sur_fuse) are laid out as Rp (SW x SM3 complex), c0
// (SW complex), Rq (SW x SM3 complex), A, a0, QW (SW each), qb.
// ---------------------------------------------------------------------------
constexpr int64_t FRC0 = 2LL * SW * SM3, FRQ = FRC0 + 2 * SW, FA = FRQ + 2LL * SW * SM3;
constexpr int64_t FA0 = FA + SW, FQW = FA0 + SW, FQB = FQW + SW, FTOT = FQB + 1;
static_assert(FTOT == FQ_SUR_FUSED_FLOATS, "folded weight layout");

__global__ __launch_bounds__(256) void k_sur_fuse(const float *__restrict__ th, int L, float *__restrict__ fw) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    const float *P = th, *p0 = th + SW;
    const float *l0 = th + 2 * SW, *ll = th + 2 * SW + int64_t(L - 1) * SLAYER;
    const float2 *R0 = reinterpret_cast<const float2 *>(l0), *RL = reinterpret_cast<const float2 *>(ll);
    const float *W0 = l0 + SR, *b0 = l0 + SR + SW * SW, *WL = ll + SR, *bL = ll + SR + SW * SW;
    const float *Qw = th + 2 * SW + int64_t(L) * SLAYER, q0 = Qw[SW];
    int64_t j = idx;
    if (j < SW * SM3) {                      // Rp(o, k)
        const int k = int(j % SM3), o = int(j / SM3);
        float2 a = make_float2(0.f, 0.f);
        for (int i = 0; i < SW; ++i) {
            const float2 r = R0[(int64_t(i) * SW + o) * SM3 + k];
            a.x += P[i] * r.x;
            a.y += P[i] * r.y;
        }
        reinterpret_cast<float2 *>(fw)[j] = a;
        return;
    }
    j -= SW * SM3;
    if (j < SW * SM3) {                      // Rq(i, k)
        const int k = int(j % SM3), i = int(j / SM3);
        float2 a = make_float2(0.f, 0.f);
        for (int o = 0; o < SW; ++o) {
            const float2 r = RL[(int64_t(i) * SW + o) * SM3 + k];
            a.x += Qw[o] * r.x;
            a.y += Qw[o] * r.y;
        }
        reinterpret_cast<float2 *>(fw + FRQ)[j] = a;
        return;
    }
    j -= SW * SM3;
    if (j < SW) {                            // c0(o), A(o), a0(o), QW(i = j)
        const int o = int(j);
        float2 c = make_float2(0.f, 0.f);
        float a = 0.f, a0 = b0[o], qw = 0.f;
        for (int i = 0; i < SW; ++i) {
            const float2 r = R0[(int64_t(i) * SW + o) * SM3];
            c.x += p0[i] * r.x;
            c.y += p0[i] * r.y;
            a += W0[o * SW + i] * P[i];
            a0 += W0[o * SW + i] * p0[i];
            qw += Qw[i] * WL[i * SW + o];     // QW at input channel o
        }
        reinterpret_cast<float2 *>(fw + FRC0)[o] = c;
        fw[FA + o] = a;
        fw[FA0 + o] = a0;
        fw[FQW + o] = qw;
        return;
    }
    if (j == SW) {                           // qb
        float q = q0;
        for (int o = 0; o < SW; ++o) q += Qw[o] * bL[o];
        fw[FQB] = q;
    }
}

// Y(b,o,k) = Xg(b,k) Rp(o,k) + [k = 0] c0(o)
__global__ __launch_bounds__(256) void k_sur_mix_first(const float2 *__restrict__ Xg, const float *__restrict__ fw,
                                                       float2 *__restrict__ Y, int64_t total) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= total) return;
    const int k = int(idx & (SM3 - 1));
    const int o = int((idx >> 9) & (SW - 1));
    const int64_t b = idx >> 14;
    const float2 x = Xg[b * SM3 + k];
    const float2 r = reinterpret_cast<const float2 *>(fw)[o * SM3 + k];
    float2 y = make_float2(x.x * r.x - x.y * r.y, x.x * r.y + x.y * r.x);
    if (k == 0) {
        const float2 c = reinterpret_cast<const float2 *>(fw + FRC0)[o];
        y.x += c.x;
        y.y += c.y;
    }
    Y[idx] = y;
}

// v_1(o) = gelu(A_o g + a0_o + s_0(o)) per point
template <int N, typename TA>
__global__ __launch_bounds__(256) void k_sur_pw_first(const float *__restrict__ g, const TA *__restrict__ s,
                                                      const float *__restrict__ fw, TA *__restrict__ vout,
                                                      int64_t total) {
    const int64_t idx = 2 * (int64_t(blockIdx.x) * 256 + threadIdx.x);   // two points per thread
    if (idx >= total) return;
    const int64_t b = idx >> SG<N>::LOG3;
    const int x = int(idx & (SG<N>::N3 - 1));
    const float2 gv = ld2(g, idx);
#pragma unroll 8
    for (int o = 0; o < SW; ++o) {
        const int64_t at = (b * SW + o) * SG<N>::N3 + x;
        const float2 sv = ld2(s, at);
        st2(vout, at, make_float2(gelu(fw[FA + o] * gv.x + fw[FA0 + o] + sv.x),
                                  gelu(fw[FA + o] * gv.y + fw[FA0 + o] + sv.y)));
    }
}

// Y1(b,k) = sum_i X(b,i,k) Rq(i,k)
__global__ __launch_bounds__(256) void k_sur_mix_last(const float2 *__restrict__ X, const float *__restrict__ fw,
                                                      float2 *__restrict__ Y1, int64_t total) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= total) return;
    const int k = int(idx & (SM3 - 1));
    const int64_t b = idx >> 9;
    const float2 *Rq = reinterpret_cast<const float2 *>(fw + FRQ);
    float2 a = make_float2(0.f, 0.f);
    for (int i = 0; i < SW; ++i) {
        const float2 x = X[(b * SW + i) * SM3 + k];
        const float2 r = Rq[i * SM3 + k];
        a.x += x.x * r.x - x.y * r.y;
        a.y += x.x * r.y + x.y * r.x;
    }
    Y1[idx] = a;
}

// out(b,x) = QW . v(b,:,x) + qb + s1(b,x)
template <int N, typename TA>
__global__ __launch_bounds__(256) void k_sur_pw_last(const TA *__restrict__ v, const TA *__restrict__ s1,
                                                     const float *__restrict__ fw, float *__restrict__ out,
                                                     int64_t total) {
    const int64_t idx = int64_t(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= total) return;
    const int64_t b = idx >> SG<N>::LOG3;
    const int x = int(idx & (SG<N>::N3 - 1));
    float a = fw[FQB] + ldA(s1, idx);
#pragma unroll 8
    for (int i = 0; i < SW; ++i) a += fw[FQW + i] * ldA(v, (b * SW + i) * SG<N>::N3 + x);
    out[idx] = a;
}

inline unsigned nblk(int64_t n) { return unsigned((n + 255) / 256); }

int check_shape(const fq_sur_shape *s, const char *who) {
    if (!s) return fq_internal_fail(FQ_ERR_INVALID, std::string(who) + ": null shape");
    if (s->abi_version != FQ_SUR_ABI_VERSION)
        return fq_internal_fail(FQ_ERR_INVALID, std::string(who) + ": ABI version mismatch");
    if ((s->Nv != 16 && s->Nv != 32) || s->width != SW || s->modes != SM || s->layers < 1 ||
        s->layers > FQ_SUR_MAX_LAYERS)
        return fq_internal_fail(FQ_ERR_UNSUPPORTED, std::string(who) +
                                ": supported shapes are Nv 16 or 32, width 32, modes 8, 1..8 layers");
    return FQ_SUCCESS;
}

int hip_status(const char *who) {
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? FQ_SUCCESS : fq_internal_fail(FQ_ERR_HIP, std::string(who) + ": " + hipGetErrorString(e));
}

struct ActPtrs {
    float *v[FQ_SUR_MAX_LAYERS + 1];
    float *z[FQ_SUR_MAX_LAYERS];
    float2 *X[FQ_SUR_MAX_LAYERS];
    float *s;
    float2 *Y;
};

// Activation layout (floats): training keeps v_0..v_L, z_0..z_{L-1} and
// X_0..X_{L-1}; inference keeps two v buffers and one X.  Both end with the
// s and Y scratch.
int64_t act_floats(int N3, int L, int64_t B, bool training) {
    const int64_t vol = B * SW * N3, md = 2 * B * SW * SM3;
    return training ? (L + 1) * vol + L * vol + L * md + vol + md : 2 * vol + md + vol + md;
}

ActPtrs act_ptrs(int N3, float *acts, int L, int64_t B, bool training) {
    const int64_t vol = B * SW * N3, md = 2 * B * SW * SM3;
    ActPtrs p{};
    float *q = acts;
    if (training) {
        for (int l = 0; l <= L; ++l) p.v[l] = q + l * vol;
        q += (L + 1) * vol;
        for (int l = 0; l < L; ++l) p.z[l] = q + l * vol;
        q += L * vol;
        for (int l = 0; l < L; ++l) p.X[l] = reinterpret_cast<float2 *>(q + l * md);
        q += L * md;
    } else {
        for (int l = 0; l <= L; ++l) p.v[l] = q + (l & 1) * vol;
        q += 2 * vol;
        for (int l = 0; l < L; ++l) p.X[l] = reinterpret_cast<float2 *>(q);
        q += md;
    }
    p.s = q;
    p.Y = reinterpret_cast<float2 *>(q + vol);
    return p;
}

// The folded inference pass with activations stored as A (FP32 or FP16) in
// the FP32-sized regions of the inference activation layout.
template <int N, typename A>
void forward_fused(int L, const float *theta, const float *fused, const float *g, int64_t batch, float *acts,
                   float *out, hipStream_t st) {
    const ActPtrs p = act_ptrs(SG<N>::N3, acts, L, batch, false);
    A *v[2] = {reinterpret_cast<A *>(p.v[0]), reinterpret_cast<A *>(p.v[1])};
    A *sa = reinterpret_cast<A *>(p.s);
    const int64_t pts = batch * SG<N>::N3;
    const unsigned slices = unsigned(batch * SW);
    // first layer from the one-channel transform of g
    launch_dft<N, float>(g, 1.0f / SG<N>::N3, p.X[0], unsigned(batch), st);
    hipLaunchKernelGGL(k_sur_mix_first, dim3(nblk(batch * SW * SM3)), dim3(256), 0, st, p.X[0], fused, p.Y,
                       batch * SW * SM3);
    launch_idft<N, A>(p.Y, 1.0f, 0, sa, slices, st);
    hipLaunchKernelGGL((k_sur_pw_first<N, A>), dim3(nblk(pts / 2)), dim3(256), 0, st, g, sa, fused, v[1], pts);
    for (int l = 1; l < L - 1; ++l) {
        const float *lay = theta + 2 * SW + l * SLAYER;
        const float2 *R = reinterpret_cast<const float2 *>(lay);
        launch_dft<N, A>(v[l & 1], 1.0f / SG<N>::N3, p.X[l], slices, st);
        hipLaunchKernelGGL(k_sur_mix_lds<0>, dim3(SM3 / 8, unsigned((batch + 31) / 32)), dim3(256), 0, st, p.X[l],
                           R, p.Y, batch);
        launch_idft<N, A>(p.Y, 1.0f, 0, sa, slices, st);
        hipLaunchKernelGGL((k_sur_pw<N, A>), dim3(nblk(pts / 2)), dim3(256), 0, st, v[l & 1], sa, lay + SR, 1, nullptr,
                           v[(l + 1) & 1], pts);
    }
    // last layer and projection as one output channel
    launch_dft<N, A>(v[(L - 1) & 1], 1.0f / SG<N>::N3, p.X[L - 1], slices, st);
    hipLaunchKernelGGL(k_sur_mix_last, dim3(nblk(batch * SM3)), dim3(256), 0, st, p.X[L - 1], fused, p.Y,
                       batch * SM3);
    launch_idft<N, A>(p.Y, 1.0f, 0, sa, unsigned(batch), st);
    hipLaunchKernelGGL((k_sur_pw_last<N, A>), dim3(nblk(pts)), dim3(256), 0, st, v[(L - 1) & 1], sa, fused, out, pts);
}

// The unfolded pass (training keeps v_l, z_l and X_l) with activations
// stored as A in the FP32-sized regions of the activation layout.
template <int N, typename A>
void forward_impl(int L, const float *theta, const float *g, int64_t batch, bool tr, float *acts, float *out,
                  hipStream_t st) {
    const ActPtrs p = act_ptrs(SG<N>::N3, acts, L, batch, tr);
    const int64_t pts = batch * SG<N>::N3;
    const unsigned slices = unsigned(batch * SW);
    A *sa = reinterpret_cast<A *>(p.s);
    hipLaunchKernelGGL((k_sur_lift<N, A>), dim3(nblk(pts)), dim3(256), 0, st, g, theta, reinterpret_cast<A *>(p.v[0]), pts);
    for (int l = 0; l < L; ++l) {
        const float *lay = theta + 2 * SW + l * SLAYER;
        const float2 *R = reinterpret_cast<const float2 *>(lay);
        A *vl = reinterpret_cast<A *>(p.v[l]), *vn = reinterpret_cast<A *>(p.v[l + 1]);
        launch_dft<N, A>(vl, 1.0f / SG<N>::N3, p.X[l], slices, st);
        hipLaunchKernelGGL(k_sur_mix_lds<0>, dim3(SM3 / 8, unsigned((batch + 31) / 32)), dim3(256), 0, st, p.X[l],
                           R, p.Y, batch);
        launch_idft<N, A>(p.Y, 1.0f, 0, sa, slices, st);
        hipLaunchKernelGGL((k_sur_pw<N, A>), dim3(nblk(pts / 2)), dim3(256), 0, st, vl, sa, lay + SR, int(l < L - 1),
                           tr ? reinterpret_cast<A *>(p.z[l]) : nullptr, vn, pts);
    }
    hipLaunchKernelGGL((k_sur_proj<N, A>), dim3(nblk(pts)), dim3(256), 0, st, reinterpret_cast<const A *>(p.v[L]),
                       theta + 2 * SW + L * SLAYER, out, pts);
}

// Gradients from the activations of a training forward_impl<A>; the
// gradients themselves (gv, gz, the mode arrays) stay FP32.
template <int N, typename A>
void backward_impl(int L, const float *theta, const float *g, int64_t batch, float *acts, const float *gout,
                   float *grad, float *work, hipStream_t st) {
    const ActPtrs p = act_ptrs(SG<N>::N3, acts, L, batch, true);
    const int64_t pts = batch * SG<N>::N3, vol = batch * SW * SG<N>::N3;
    const unsigned slices = unsigned(batch * SW);
    float *gv = work, *gvin = work + vol;
    float2 *gY = reinterpret_cast<float2 *>(work + 2 * vol);
    float2 *gX = gY + batch * SW * SM3;
    const dim3 wg(unsigned(batch * (SG<N>::N3 / 512)));
    const int64_t qoff = 2 * SW + L * SLAYER;
    hipLaunchKernelGGL((k_sur_wgrad<N, 1, SW, A>), wg, dim3(256), 0, st, gout, reinterpret_cast<const A *>(p.v[L]),
                       grad + qoff, grad + qoff + SW);
    hipLaunchKernelGGL(k_sur_proj_bwd<N>, dim3(nblk(pts)), dim3(256), 0, st, gout, theta + qoff, gv, pts);
    for (int l = L - 1; l >= 0; --l) {
        const int64_t lo = 2 * SW + l * SLAYER;
        const float2 *R = reinterpret_cast<const float2 *>(theta + lo);
        float2 *gR = reinterpret_cast<float2 *>(grad + lo);
        hipLaunchKernelGGL((k_sur_pw_bwd<N, A>), dim3(nblk(pts)), dim3(256), 0, st, gv,
                           reinterpret_cast<const A *>(p.z[l]), theta + lo + SR, int(l < L - 1), gvin, pts);
        hipLaunchKernelGGL((k_sur_wgrad<N, SW, SW, A>), wg, dim3(256), 0, st, gv, reinterpret_cast<const A *>(p.v[l]),
                           grad + lo + SR, grad + lo + SR + SW * SW);
        launch_dft<N, float>(gv, 1.0f, gY, slices, st);
        hipLaunchKernelGGL(k_sur_mix_grad, dim3(nblk(int64_t(SW) * SW * SM3)), dim3(256), 0, st, p.X[l], gY, gR,
                           batch);
        hipLaunchKernelGGL(k_sur_mix_lds<1>, dim3(SM3 / 8, unsigned((batch + 31) / 32)), dim3(256), 0, st, gY, R,
                           gX, batch);
        launch_idft<N, float>(gX, 1.0f / SG<N>::N3, 1, gvin, slices, st);
        float *tmp = gv;
        gv = gvin;
        gvin = tmp;
    }
    hipLaunchKernelGGL((k_sur_wgrad<N, SW, 1, float>), wg, dim3(256), 0, st, gv, g, grad, grad + SW);
}

// Dispatch on the activation storage type.
template <int N>
void forward_acts(int acts_type, int L, const float *theta, const float *g, int64_t batch, bool tr, float *acts,
                  float *out, hipStream_t st) {
    if (acts_type == FQ_SUR_ACTS_FP16) forward_impl<N, _Float16>(L, theta, g, batch, tr, acts, out, st);
    else if (acts_type == FQ_SUR_ACTS_BF16) forward_impl<N, __bf16>(L, theta, g, batch, tr, acts, out, st);
    else forward_impl<N, float>(L, theta, g, batch, tr, acts, out, st);
}
template <int N>
void fused_acts(int acts_type, int L, const float *theta, const float *fused, const float *g, int64_t batch,
                float *acts, float *out, hipStream_t st) {
    if (acts_type == FQ_SUR_ACTS_FP16) forward_fused<N, _Float16>(L, theta, fused, g, batch, acts, out, st);
    else if (acts_type == FQ_SUR_ACTS_BF16) forward_fused<N, __bf16>(L, theta, fused, g, batch, acts, out, st);
    else forward_fused<N, float>(L, theta, fused, g, batch, acts, out, st);
}
template <int N>
void backward_acts(int acts_type, int L, const float *theta, const float *g, int64_t batch, float *acts,
                   const float *gout, float *grad, float *work, hipStream_t st) {
    if (acts_type == FQ_SUR_ACTS_FP16) backward_impl<N, _Float16>(L, theta, g, batch, acts, gout, grad, work, st);
    else if (acts_type == FQ_SUR_ACTS_BF16) backward_impl<N, __bf16>(L, theta, g, batch, acts, gout, grad, work, st);
    else backward_impl<N, float>(L, theta, g, batch, acts, gout, grad, work, st);
}

}  // namespace

extern "C" {

int fq_sur_sizes(const fq_sur_shape *shape, int64_t batch, int32_t training, int64_t *nparams, int64_t *act_size,
                 int64_t *work_size) {
    if (int rc = check_shape(shape, "fq_sur_sizes")) return rc;
    if (batch < 0) return fq_internal_fail(FQ_ERR_INVALID, "fq_sur_sizes: negative batch");
    const int L = shape->layers;
    if (nparams) *nparams = 2 * SW + L * SLAYER + SW + 1;
    const int N3 = shape->Nv * shape->Nv * shape->Nv;
    if (act_size) *act_size = act_floats(N3, L, batch, training != 0);
    if (work_size) *work_size = 2 * batch * SW * N3 + 4 * batch * SW * SM3;
    return FQ_SUCCESS;
}

int fq_sur_forward(const fq_sur_shape *shape, const float *theta, const float *g, int64_t batch, int32_t training,
                   int32_t acts_type, float *acts, float *out, void *stream) {
    if (int rc = check_shape(shape, "fq_sur_forward")) return rc;
    if (batch < 0 || batch > FQ_SUR_MAX_BATCH || !theta || !g || !acts || !out ||
        acts_type < FQ_SUR_ACTS_FP32 || acts_type > FQ_SUR_ACTS_BF16)
        return fq_internal_fail(FQ_ERR_INVALID, "fq_sur_forward: invalid arguments");
    if (batch == 0) return FQ_SUCCESS;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const bool tr = training != 0;
    if (shape->Nv == 16) forward_acts<16>(acts_type, shape->layers, theta, g, batch, tr, acts, out, st);
    else forward_acts<32>(acts_type, shape->layers, theta, g, batch, tr, acts, out, st);
    return hip_status("fq_sur_forward");
}

int fq_sur_fuse(const fq_sur_shape *shape, const float *theta, float *fused, void *stream) {
    if (int rc = check_shape(shape, "fq_sur_fuse")) return rc;
    if (shape->layers < 2) return fq_internal_fail(FQ_ERR_UNSUPPORTED, "fq_sur_fuse: needs at least two layers");
    if (!theta || !fused) return fq_internal_fail(FQ_ERR_INVALID, "fq_sur_fuse: invalid arguments");
    hipLaunchKernelGGL(k_sur_fuse, dim3(nblk(2LL * SW * SM3 + SW + 1)), dim3(256), 0,
                       reinterpret_cast<hipStream_t>(stream), theta, shape->layers, fused);
    return hip_status("fq_sur_fuse");
}

int fq_sur_forward_fused(const fq_sur_shape *shape, const float *theta, const float *fused, const float *g,
                         int64_t batch, int32_t acts_type, float *acts, float *out, void *stream) {
    if (int rc = check_shape(shape, "fq_sur_forward_fused")) return rc;
    if (shape->layers < 2)
        return fq_internal_fail(FQ_ERR_UNSUPPORTED, "fq_sur_forward_fused: needs at least two layers");
    if (batch < 0 || batch > FQ_SUR_MAX_BATCH || !theta || !fused || !g || !acts || !out ||
        acts_type < FQ_SUR_ACTS_FP32 || acts_type > FQ_SUR_ACTS_BF16)
        return fq_internal_fail(FQ_ERR_INVALID, "fq_sur_forward_fused: invalid arguments");
    if (batch == 0) return FQ_SUCCESS;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (shape->Nv == 16) fused_acts<16>(acts_type, shape->layers, theta, fused, g, batch, acts, out, st);
    else fused_acts<32>(acts_type, shape->layers, theta, fused, g, batch, acts, out, st);
    return hip_status("fq_sur_forward_fused");
}

int fq_sur_backward(const fq_sur_shape *shape, const float *theta, const float *g, int64_t batch, int32_t acts_type,
                    float *acts, const float *gout, float *grad, float *work, void *stream) {
    if (int rc = check_shape(shape, "fq_sur_backward")) return rc;
    if (batch < 0 || batch > FQ_SUR_MAX_BATCH || !theta || !g || !acts || !gout || !grad || !work ||
        acts_type < FQ_SUR_ACTS_FP32 || acts_type > FQ_SUR_ACTS_BF16)
        return fq_internal_fail(FQ_ERR_INVALID, "fq_sur_backward: invalid arguments");
    if (batch == 0) return FQ_SUCCESS;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (shape->Nv == 16) backward_acts<16>(acts_type, shape->layers, theta, g, batch, acts, gout, grad, work, st);
    else backward_acts<32>(acts_type, shape->layers, theta, g, batch, acts, gout, grad, work, st);
    return hip_status("fq_sur_backward");
}

int fq_sur_loss(int64_t count, const float *out, const float *target, float *gout, double *sums, void *stream) {
    if (count < 0 || !out || !target || !gout || !sums)
        return fq_internal_fail(FQ_ERR_INVALID, "fq_sur_loss: invalid arguments");
    if (count == 0) return FQ_SUCCESS;
    hipLaunchKernelGGL(k_sur_loss, dim3(nblk(count)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), out,
                       target, gout, sums, count);
    return hip_status("fq_sur_loss");
}

int fq_sur_cons(int32_t Nv, double L, double cw, double lambda, int64_t batch, const float *out,
                const double *gram_inv, float *gout, double *sums, void *stream) {
    if ((Nv != 16 && Nv != 32) || !(L > 0) || !(cw > 0) || !(lambda >= 0) || batch < 0 || batch > FQ_SUR_MAX_BATCH ||
        !out || !gram_inv || !gout || !sums)
        return fq_internal_fail(FQ_ERR_INVALID, "fq_sur_cons: invalid arguments");
    if (batch == 0 || lambda == 0) return FQ_SUCCESS;
    SurGram gm;
    for (int j = 0; j < 25; ++j) gm.ginv[j] = gram_inv[j];
    for (int j = 0; j < 5; ++j) gm.gdiag[j] = 0;
    const double count = double(batch) * Nv * Nv * Nv;
    hipLaunchKernelGGL(k_sur_cons, dim3(unsigned(batch)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), Nv,
                       -L + L / Nv, 2 * L / Nv, cw, gm, 2.0 * lambda / count, out, gout, sums);
    return hip_status("fq_sur_cons");
}

int fq_sur_adam(int64_t nparams, float *theta, const float *grad, float *m, float *v, float lr, float beta1,
                float beta2, float eps, int32_t step, void *stream) {
    if (nparams < 0 || !theta || !grad || !m || !v || step < 1 || !(lr >= 0) || !(beta1 >= 0 && beta1 < 1) ||
        !(beta2 >= 0 && beta2 < 1) || !(eps > 0))
        return fq_internal_fail(FQ_ERR_INVALID, "fq_sur_adam: invalid arguments");
    if (nparams == 0) return FQ_SUCCESS;
    const float c1 = float(1.0 / (1.0 - pow(double(beta1), step)));
    const float c2 = float(1.0 / (1.0 - pow(double(beta2), step)));
    hipLaunchKernelGGL(k_sur_adam, dim3(nblk(nparams)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), theta,
                       grad, m, v, nparams, lr, beta1, beta2, eps, c1, c2);
    return hip_status("fq_sur_adam");
}

int fq_sur_gather(int32_t Nv, double cw, double scale, int32_t power, const double *src, int32_t nupts, int64_t ldim,
                  int32_t neles, int64_t first_dof, int64_t ndof, const double *rho_in, double *rho_out, float *dst,
                  void *stream) {
    if (Nv <= 0 || !(cw > 0) || !(scale != 0) || power < 0 || power > 4 || !src || !dst || nupts <= 0 ||
        neles <= 0 || first_dof < 0 || ndof < 0 || first_dof + ndof > int64_t(nupts) * neles)
        return fq_internal_fail(FQ_ERR_INVALID, "fq_sur_gather: invalid arguments");
    const int64_t nparr = (int64_t(neles) + 31) / 32 * 32;
    if (ldim < int64_t(Nv) * Nv * Nv * nparr) return fq_internal_fail(FQ_ERR_INVALID, "fq_sur_gather: ldim too small");
    if (ndof == 0) return FQ_SUCCESS;
    hipLaunchKernelGGL(k_sur_gather, dim3(xcd_grid(ndof)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream),
                       Nv * Nv * Nv, cw, scale, power, src, nupts, ldim, first_dof, rho_in, rho_out, dst, ndof);
    return hip_status("fq_sur_gather");
}

int fq_sur_finish(int32_t Nv, double L, double cw, double scale, double tol, const float *q, const double *rho,
                  const double *gram_inv, const double *gram_diag, double *dst, int32_t nupts, int64_t ldim,
                  int32_t neles, int64_t first_dof, int64_t ndof, double *defect, int32_t *flag, void *stream) {
    if (Nv <= 0 || !(L > 0) || !(cw > 0) || !q || !rho || !gram_inv || !gram_diag || !dst || !defect || !flag ||
        nupts <= 0 || neles <= 0 || first_dof < 0 || ndof < 0 || first_dof + ndof > int64_t(nupts) * neles)
        return fq_internal_fail(FQ_ERR_INVALID, "fq_sur_finish: invalid arguments");
    const int64_t nparr = (int64_t(neles) + 31) / 32 * 32;
    if (ldim < int64_t(Nv) * Nv * Nv * nparr) return fq_internal_fail(FQ_ERR_INVALID, "fq_sur_finish: ldim too small");
    if (ndof == 0) return FQ_SUCCESS;
    SurGram gm;
    for (int j = 0; j < 25; ++j) gm.ginv[j] = gram_inv[j];
    for (int j = 0; j < 5; ++j) gm.gdiag[j] = gram_diag[j];
    hipLaunchKernelGGL(k_sur_finish, dim3(xcd_grid(ndof)), dim3(256), 0, reinterpret_cast<hipStream_t>(stream), Nv,
                       -L + L / Nv, 2 * L / Nv, cw, scale, tol, q, rho, gm, dst, nupts, ldim, first_dof, defect, flag,
                       ndof);
    return hip_status("fq_sur_finish");
}

}  // extern "C"
