// Complex arithmetic and the constant bank shared by every kernel.
//
// The bank holds, for every line length the kernels use, the table W_N^m = exp(-2 pi i m/N),
// m in [0,N), plus the Rader-17 convolution spectra and the folded-17 cos/sin rows.  It is
// computed on the host in long double by the plan step (plan.cpp) and uploaded once per
// plan; kernels read it through a __restrict__ const pointer with compile-time offsets, so
// every access is a uniform scalar load that lands in SGPRs.
#pragma once
#include <hip/hip_runtime.h>

#define FX_HD __host__ __device__ __forceinline__

namespace fx {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
    T r, i;
};

template <typename T>
FX_HD cx<T> operator+(cx<T> a, cx<T> b) { return {a.r + b.r, a.i + b.i}; }
template <typename T>
FX_HD cx<T> operator-(cx<T> a, cx<T> b) { return {a.r - b.r, a.i - b.i}; }
template <typename T>
FX_HD cx<T> operator-(cx<T> a) { return {-a.r, -a.i}; }
template <typename T>
FX_HD cx<T> scale(cx<T> a, T s) { return {a.r * s, a.i * s}; }
template <typename T>
FX_HD cx<T> cmul(cx<T> a, cx<T> w)
{
    return {a.r * w.r - a.i * w.i, a.r * w.i + a.i * w.r};
}
// a * (S i): S = -1 gives a*(-i) = (a.i, -a.r); S = +1 gives a*(+i) = (-a.i, a.r).
template <int S, typename T>
FX_HD cx<T> mul_si(cx<T> a)
{
    if constexpr (S < 0) return {a.i, -a.r};
    else return {-a.i, a.r};
}

// Line lengths that have a W table in the bank, in bank order.
constexpr int kLens[] = {2, 3, 4, 6, 8, 9, 12, 16, 17, 24, 36, 48};
constexpr int kNumLens = sizeof(kLens) / sizeof(kLens[0]);

constexpr int woff(int n)
{
    int off = 0;
    for (int k = 0; k < kNumLens; ++k) {
        if (kLens[k] == n) return off;
        off += kLens[k];
    }
    return -1;
}
constexpr int kBankW = woff(48) + 48;

template <typename T>
struct Bank {
    cx<T> w[kBankW];          // W_N^m = exp(-2 pi i m / N), forward sign
    cx<T> rader17[2][16];     // [0] forward, [1] backward: DFT16(b')/16, b'[q] = W17^(3^-q)
    T cos17[17];              // cos(2 pi m / 17)
    T sin17[17];              // sin(2 pi m / 17)
};

}  // namespace fx
