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

// Streaming HBM access.  A transform reads each input element once and writes each output
// element once, so those accesses carry the nontemporal hint (the `nt` bit of the gfx950
// global load/store): the caches give the lines the streaming policy instead of keeping
// them.  A plain copy with one 78.6 KB volume per workgroup (17^3 complex128, batch 1024)
// takes 36.3 us, 33.4 us with hinted stores, 32.9 us with both hinted.  It pays where every
// store instruction of a wave writes whole lines (the fused x pass, pencil pass B: 8^3 at
// batch 4096 12.9 -> 11.3 us, 36^3 and 24x24x48 2-5%); where a store writes short runs
// that later instructions complete (the pipelined y pass when a z run is shorter than a
// 128 B line: 6^3, 12^3, 17^3) it gains nothing or loses (17^3 with two volumes per
// workgroup: 15% slower), so those stores stay plain.  Pencil pass B reads its intermediate with the
// hint (the lines are read once, then overwritten: 36^3 -2.4%, 24x24x48 -7%), but pass A
// stores it plainly: with hinted stores it leaves the Infinity Cache before pass B reads it
// back, and 36^3 went from 69 to 77.5 us.
template <typename T>
FX_HD cx<T> ld_stream(const cx<T>* p)
{
    using V2 = T __attribute__((ext_vector_type(2)));
    const V2 v = __builtin_nontemporal_load(reinterpret_cast<const V2*>(p));
    return {v.x, v.y};
}
template <typename T>
FX_HD void st_stream(cx<T>* p, cx<T> w)
{
    using V2 = T __attribute__((ext_vector_type(2)));
    __builtin_nontemporal_store(V2{w.r, w.i}, reinterpret_cast<V2*>(p));
}

// H = 0: a plain access; H != 0: the streaming form above.  Stores move the element as one
// two-component vector either way (a struct copy can be split into two half-width stores).
template <int H, typename T>
FX_HD cx<T> ld_hint(const cx<T>* p)
{
    if constexpr (H != 0) return ld_stream(p);
    else return *p;
}
template <int H, typename T>
FX_HD void st_hint(cx<T>* p, cx<T> w)
{
    if constexpr (H != 0) {
        st_stream(p, w);
    } else {
        using V2 = T __attribute__((ext_vector_type(2)));
        *reinterpret_cast<V2*>(p) = V2{w.r, w.i};
    }
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
