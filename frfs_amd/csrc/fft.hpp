// In-register and lane-pair FFT building blocks for gfx950 (wave64).
//
// Conventions
//   S = -1 : e^{-2 pi i jk/N}  (the reference's CUFFT_FORWARD,
//            frfs/solvers/dgfs/cufft.py:182)
//   S = +1 : e^{+2 pi i jk/N}  (CUFFT_INVERSE, cufft.py:183)
//   Transforms are unnormalised.
//
// A length-N pencil lives on a PAIR of adjacent lanes (lane ^ 1), each lane
// holding E = N/2 complex values in registers.  Two data layouts exist:
//   block  : lane half a holds x[a*E + e] in slot e
//   cyclic : lane half a holds x[a + 2*e] in slot e
// fft2_dif : block in  -> cyclic out (exchange, twiddle, local FFT)
// fft2_dit : cyclic in -> block out  (local FFT, twiddle, exchange)
// The pair exchange is a DPP quad_perm [1,0,3,2] lane swap (one VALU op per
// 32-bit word, no LDS traffic).
//
// Twiddle factors are compile-time constants evaluated by a constexpr
// octant-reduced Taylor series (no run-time sin/cos, no tables).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include <utility>

namespace fq {

template <typename T> struct cpx { T x, y; };

template <typename T> __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T> __device__ __forceinline__ cpx<T> cmul(cpx<T> a, cpx<T> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> __device__ __forceinline__ cpx<T> cscale(cpx<T> a, T s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
// t += a * b as four fused multiply-adds
template <typename T> __device__ __forceinline__ void cmac(cpx<T> &t, cpx<T> a, cpx<T> b) {
    t.x = fmadd(a.x, b.x, t.x);
    t.x = fmadd(-a.y, b.y, t.x);
    t.y = fmadd(a.x, b.y, t.y);
    t.y = fmadd(a.y, b.x, t.y);
}

// ---------------------------------------------------------------------------
// compile-time loop
// ---------------------------------------------------------------------------
template <int I, int N, typename F> __device__ __forceinline__ void sfor(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        sfor<I + 1, N>(f);
    }
}

// ---------------------------------------------------------------------------
// constexpr twiddles: cos/sin(2 pi k / E) with exact rational octant reduction
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double ct_sin(double x) {
    double term = x, sum = x;
    for (int n = 1; n < 14; ++n) { term *= -x * x / double((2 * n) * (2 * n + 1)); sum += term; }
    return sum;
}
constexpr double ct_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 14; ++n) { term *= -x * x / double((2 * n - 1) * (2 * n)); sum += term; }
    return sum;
}

template <int E, int K> struct Twiddle {
    static constexpr int k = ((K % E) + E) % E;
    static constexpr int m = 8 * k;
    static constexpr int o = m / E;
    static constexpr int rem = m - o * E;
    static constexpr double phi = (kPi / 4.0) * double(rem) / double(E);
    static constexpr double psi = (kPi / 4.0) * double(E - rem) / double(E);
    static constexpr double cphi = ct_cos(phi), sphi = ct_sin(phi);
    static constexpr double cpsi = ct_cos(psi), spsi = ct_sin(psi);
    // cos / sin of 2 pi k / E
    static constexpr double c = o == 0 ? cphi : o == 1 ? spsi : o == 2 ? -sphi : o == 3 ? -cpsi
                              : o == 4 ? -cphi : o == 5 ? -spsi : o == 6 ? sphi : cpsi;
    static constexpr double s = o == 0 ? sphi : o == 1 ? cpsi : o == 2 ? cphi : o == 3 ? spsi
                              : o == 4 ? -sphi : o == 5 ? -cpsi : o == 6 ? -cphi : -spsi;
    static constexpr bool one = (k == 0);
    static constexpr bool quarter = ((4 * k) % E == 0);   // +-1, +-i
    static constexpr int qidx = quarter ? (4 * k) / E : 0;   // multiple of pi/2
};

// v * e^{S * 2 pi i K / E}
template <int E, int S, int K, typename T> __device__ __forceinline__ cpx<T> twmul(cpx<T> v) {
    using W = Twiddle<E, K>;
    if constexpr (W::one) {
        return v;
    } else if constexpr (W::quarter) {
        // e^{S i pi/2 * q}
        constexpr int q = ((W::qidx * S) % 4 + 4) % 4;
        if constexpr (q == 1) return {-v.y, v.x};
        else if constexpr (q == 2) return {-v.x, -v.y};
        else return {v.y, -v.x};
    } else {
        const T c = T(W::c), s = T(S * W::s);
        return {v.x * c - v.y * s, v.x * s + v.y * c};
    }
}

// ---------------------------------------------------------------------------
// In-register FFT of length E, natural order in and out.
// Mixed radix (4, 2, 3) decimation: n = m*a + b, k = c + r*d.
// ---------------------------------------------------------------------------
template <int E, int S> struct LFFT;

template <int S> struct LFFT<1, S> {
    template <typename T> __device__ __forceinline__ static void run(cpx<T> *) {}
};
template <int S> struct LFFT<2, S> {
    template <typename T> __device__ __forceinline__ static void run(cpx<T> *x) {
        cpx<T> a = x[0], b = x[1];
        x[0] = a + b;
        x[1] = a - b;
    }
};
template <int S> struct LFFT<3, S> {
    template <typename T> __device__ __forceinline__ static void run(cpx<T> *x) {
        // X1 = x0 + w x1 + w^2 x2, w = e^{S 2 pi i/3} = -1/2 + S i sqrt(3)/2
        constexpr double h = 0.86602540378443864676372317075293618;
        const T s = T(S * h);
        cpx<T> t = x[1] + x[2];
        cpx<T> d = x[1] - x[2];
        cpx<T> x0 = x[0];
        x[0] = x0 + t;
        cpx<T> m = {x0.x - T(0.5) * t.x, x0.y - T(0.5) * t.y};
        // i*s*d
        cpx<T> r = {-s * d.y, s * d.x};
        x[1] = m + r;
        x[2] = m - r;
    }
};
template <int S> struct LFFT<4, S> {
    template <typename T> __device__ __forceinline__ static void run(cpx<T> *x) {
        cpx<T> a = x[0] + x[2], b = x[0] - x[2];
        cpx<T> c = x[1] + x[3], d = x[1] - x[3];
        // d * (S i)
        cpx<T> di = (S > 0) ? cpx<T>{-d.y, d.x} : cpx<T>{d.y, -d.x};
        x[0] = a + c;
        x[2] = a - c;
        x[1] = b + di;
        x[3] = b - di;
    }
};

template <int E> struct Radix {
    static constexpr int r = (E % 4 == 0 && E > 4) ? 4 : (E % 2 == 0 ? 2 : (E % 3 == 0 ? 3 : E));
};

template <int E, int S> struct LFFT {
    static_assert(E % 2 == 0 || E % 3 == 0, "unsupported FFT length");
    template <typename T> __device__ __forceinline__ static void run(cpx<T> *x) {
        constexpr int r = Radix<E>::r;
        constexpr int m = E / r;
        cpx<T> y[E];
        sfor<0, m>([&](auto bb) {
            constexpr int b = decltype(bb)::value;
            cpx<T> t[r];
            sfor<0, r>([&](auto aa) { constexpr int a = decltype(aa)::value; t[a] = x[m * a + b]; });
            LFFT<r, S>::run(t);
            sfor<0, r>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                y[c * m + b] = twmul<E, S, b * c>(t[c]);
            });
        });
        sfor<0, r>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            LFFT<m, S>::run(y + c * m);
        });
        sfor<0, r>([&](auto cc) {
            constexpr int c = decltype(cc)::value;
            sfor<0, m>([&](auto dd) {
                constexpr int d = decltype(dd)::value;
                x[c + r * d] = y[c * m + d];
            });
        });
    }
};

// ---------------------------------------------------------------------------
// Lane-pair exchange: value of lane ^ 1 (DPP quad_perm [1,0,3,2]).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int swap_adj_i32(int v) {
    // every lane of the quad is a valid source, so no "old" value is needed
    return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);
}
__device__ __forceinline__ float swap_adj(float v) {
    return __builtin_bit_cast(float, swap_adj_i32(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double swap_adj(double v) {
    uint64_t u = __builtin_bit_cast(uint64_t, v);
    int lo = swap_adj_i32(int(uint32_t(u)));
    int hi = swap_adj_i32(int(uint32_t(u >> 32)));
    uint64_t r = (uint64_t(uint32_t(hi)) << 32) | uint64_t(uint32_t(lo));
    return __builtin_bit_cast(double, r);
}
template <typename T> __device__ __forceinline__ cpx<T> swap_adj(cpx<T> v) { return {swap_adj(v.x), swap_adj(v.y)}; }
// Value of lane ^ 2 (DPP quad_perm [2,3,0,1]).
__device__ __forceinline__ float swap_q2(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, true));
}
__device__ __forceinline__ cpx<float> swap_q2(cpx<float> v) { return {swap_q2(v.x), swap_q2(v.y)}; }
// partner's value + (hi ? -v : v).  FP32: a sign select feeding an add whose
// other operand is the DPP swap, which folds into one v_add_f32_dpp.  FP64
// has no DPP form of its add, so the sign enters one FMA.
template <typename T> __device__ __forceinline__ cpx<T> pair_combine(cpx<T> v, bool hi) {
    const cpx<T> p = swap_adj(v);
    if constexpr (std::is_same_v<T, float>) {
        return {p.x + (hi ? -v.x : v.x), p.y + (hi ? -v.y : v.y)};
    } else {
        const T sg = hi ? T(-1) : T(1);
        return {fmadd(sg, v.x, p.x), fmadd(sg, v.y, p.y)};
    }
}

// Runtime-selected twiddle for the high lane only: lane half 0 multiplies by 1.
template <int N, int S, int K, typename T> __device__ __forceinline__ cpx<T> twmul_hi(cpx<T> v, bool hi) {
    cpx<T> t = twmul<N, S, K>(v);
    return hi ? t : v;
}

// block in -> cyclic out.  half = lane & 1.
template <int N, int S, typename T> __device__ __forceinline__ void fft2_dif(cpx<T> *v, int half) {
    constexpr int E = N / 2;
    const bool hi = half != 0;
    sfor<0, E>([&](auto bb) {
        constexpr int b = decltype(bb)::value;
        // half 0: x0 + x1 ; half 1: x0 - x1  (x0 = lo lane's value, x1 = hi lane's)
        v[b] = twmul_hi<N, S, b>(pair_combine(v[b], hi), hi);
    });
    LFFT<E, S>::run(v);
}

// cyclic in -> block out.
template <int N, int S, typename T> __device__ __forceinline__ void fft2_dit(cpx<T> *v, int half) {
    constexpr int E = N / 2;
    const bool hi = half != 0;
    LFFT<E, S>::run(v);
    sfor<0, E>([&](auto jj) {
        constexpr int j = decltype(jj)::value;
        cpx<T> w = twmul_hi<N, S, j>(v[j], hi);      // hi lane: Y1[j] * w^j
        // lane 0: Y0 + wY1 ; lane 1: Y0 - wY1
        v[j] = pair_combine(w, hi);
    });
}

}  // namespace fq
