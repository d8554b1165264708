// Register-resident DFT codelets for the line lengths the 3D kernels use.
//
//   dft<N,S,T,V17>(x, bank)   in-place DFT of x[0..N-1], exponent sign S (-1 forward,
//                             +1 backward), unnormalised.
//
// Every index in here is a compile-time constant once the fixed-trip loops are unrolled,
// so the arrays stay in VGPRs and the twiddle reads are scalar loads from the bank.
//   2, 3, 4    butterflies (3 is the Winograd form: one real scale by -1/2, one by sqrt(3)/2)
//   8, 9, 16   Cooley-Tukey N1 x N2 with bank twiddles W_N^(n2 k1)
//   6, 12, 24, 36, 48   Good-Thomas prime-factor (2x3, 4x3, 8x3, 4x9, 16x3): no twiddles,
//              the Ruritanian input map n = (N2 n1 + N1 n2) mod N and CRT output map
//              k = (E1 k1 + E2 k2) mod N are pure register renaming
//              (index maps as in bench/geom/strategies/L36_pfa.md:10-24)
// The fourth parameter (V17) selects the line algorithm: for 17 as below; for the composite
// lengths 1 takes the factors in the other order (8 = 4x2, 16 = 2x8, 6 = 3x2, 12 = 3x4,
// 24 = 3x8, 36 = 9x4, 48 = 3x16), a radix order the create-time race can choose.
//   17         V17 = 0: Rader with generator 3 -- a 16-point cyclic convolution done as two
//              forward DFT16s and a pointwise product with the plan-time spectrum
//              (construction of bench/geom/strategies/L17_rader.md:21-55);
//              V17 = 1: conjugate-pair folded dense form (u = x_j + x_{17-j},
//              v = x_j - x_{17-j}, real cos/sin rows), the alternative of
//              bench/gpu/impl_4/L17_dmma.cu:165-215.
#pragma once
#include "cx.hpp"

namespace fx {

constexpr int inv_mod(int a, int m)
{
    for (int x = 1; x < m; ++x)
        if ((a * x) % m == 1) return x;
    return m == 1 ? 0 : -1;
}
constexpr int pow_mod(int g, int e, int m)
{
    int r = 1 % m;
    for (int k = 0; k < e; ++k) r = (r * g) % m;
    return r;
}

// perm17.p[u] = 3^u mod 17: the generator powers Rader-17 gathers and scatters by.
struct Perm17 {
    int p[16];
};
constexpr Perm17 make_perm17()
{
    Perm17 r{};
    int v = 1;
    for (int u = 0; u < 16; ++u) {
        r.p[u] = v;
        v = (v * 3) % 17;
    }
    return r;
}
constexpr Perm17 kPerm17 = make_perm17();

template <int N, int S, typename T, int V17 = 0>
FX_HD void dft(cx<T>* x, const Bank<T>* __restrict__ bk);

// a * W_N^m with W_N = exp(S 2 pi i / N).  m is a compile-time constant after unrolling, so
// the trivial rotations fold away and the rest read one bank entry.
template <int N, int S, typename T>
FX_HD cx<T> tw(cx<T> a, int m, const Bank<T>* __restrict__ bk)
{
    m %= N;
    if (m == 0) return a;
    if (2 * m == N) return -a;
    if (4 * m == N) return mul_si<S>(a);
    if (4 * m == 3 * N) return mul_si<-S>(a);
    cx<T> w = bk->w[woff(N) + m];
    if constexpr (S > 0) w.i = -w.i;
    return cmul(a, w);
}

template <typename T>
FX_HD void dft2(cx<T>* x)
{
    cx<T> a = x[0], b = x[1];
    x[0] = a + b;
    x[1] = a - b;
}

template <int S, typename T>
FX_HD void dft3(cx<T>* x, const Bank<T>* __restrict__ bk)
{
    const cx<T> w = bk->w[woff(3) + 1];  // (-1/2, -sqrt(3)/2)
    cx<T> t1 = x[1] + x[2];
    cx<T> d = x[1] - x[2];
    cx<T> t2 = {x[0].r + w.r * t1.r, x[0].i + w.r * t1.i};
    // (S i) * (sqrt(3)/2) * d
    cx<T> u = mul_si<S>(scale(d, -w.i));
    x[0] = x[0] + t1;
    x[1] = t2 + u;
    x[2] = t2 - u;
}

template <int S, typename T>
FX_HD void dft4(cx<T>* x)
{
    cx<T> a = x[0] + x[2], b = x[0] - x[2];
    cx<T> c = x[1] + x[3], d = mul_si<S>(x[1] - x[3]);
    x[0] = a + c;
    x[2] = a - c;
    x[1] = b + d;
    x[3] = b - d;
}

// Cooley-Tukey: n = N2 n1 + n2, k = k1 + N1 k2.
template <int N1, int N2, int S, typename T>
FX_HD void ct(cx<T>* x, const Bank<T>* __restrict__ bk)
{
    constexpr int N = N1 * N2;
    cx<T> y[N];
#pragma unroll
    for (int n2 = 0; n2 < N2; ++n2) {
        cx<T> t[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) t[n1] = x[N2 * n1 + n2];
        dft<N1, S, T>(t, bk);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) y[k1 * N2 + n2] = tw<N, S, T>(t[k1], n2 * k1, bk);
    }
#pragma unroll
    for (int k1 = 0; k1 < N1; ++k1) {
        cx<T> t[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) t[n2] = y[k1 * N2 + n2];
        dft<N2, S, T>(t, bk);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) x[k1 + N1 * k2] = t[k2];
    }
}

// Good-Thomas prime-factor algorithm, gcd(N1, N2) = 1.
template <int N1, int N2, int S, typename T>
FX_HD void pfa(cx<T>* x, const Bank<T>* __restrict__ bk)
{
    constexpr int N = N1 * N2;
    constexpr int E1 = (N2 * inv_mod(N2 % N1, N1)) % N;
    constexpr int E2 = (N1 * inv_mod(N1 % N2, N2)) % N;
    static_assert(inv_mod(N2 % N1, N1) >= 0 && inv_mod(N1 % N2, N2) >= 0, "coprime factors");
    cx<T> y[N];
#pragma unroll
    for (int n1 = 0; n1 < N1; ++n1) {
        cx<T> t[N2];
#pragma unroll
        for (int n2 = 0; n2 < N2; ++n2) t[n2] = x[(N2 * n1 + N1 * n2) % N];
        dft<N2, S, T>(t, bk);
#pragma unroll
        for (int k2 = 0; k2 < N2; ++k2) y[n1 * N2 + k2] = t[k2];
    }
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) {
        cx<T> t[N1];
#pragma unroll
        for (int n1 = 0; n1 < N1; ++n1) t[n1] = y[n1 * N2 + k2];
        dft<N1, S, T>(t, bk);
#pragma unroll
        for (int k1 = 0; k1 < N1; ++k1) x[(E1 * k1 + E2 * k2) % N] = t[k1];
    }
}

// Rader-17, generator 3:  a[u] = x[3^u], A = DFT16(a), X[0] = x[0] + A[0],
// P = A * Bhat (+ x[0] folded into P[0]), X[3^w] = DFT16(P)[w].  Both DFT16s are forward;
// the direction lives entirely in the spectrum Bhat.  Bhat[0] = -1/16 and Bhat[8] =
// sqrt(17)/16 are real in both directions, so those two products are real scalings.
template <int S, typename T>
FX_HD void rader17(cx<T>* x, const Bank<T>* __restrict__ bk)
{
    cx<T> a[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) a[u] = x[kPerm17.p[u]];
    const cx<T> x0 = x[0];
    dft<16, -1, T>(a, bk);
    const cx<T> X0 = x0 + a[0];
    const cx<T>* __restrict__ R = bk->rader17[S < 0 ? 0 : 1];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        if (s == 0 || s == 8) a[s] = scale(a[s], R[s].r);
        else a[s] = cmul(a[s], R[s]);
    }
    a[0] = a[0] + x0;
    dft<16, -1, T>(a, bk);
    x[0] = X0;
#pragma unroll
    for (int w = 0; w < 16; ++w) x[kPerm17.p[w]] = a[w];
}

// Folded dense 17: X_k = A_k + S i B_k, X_{17-k} = A_k - S i B_k with
// A_k = x0 + sum_j cos(2 pi jk/17) u_j,  B_k = sum_j sin(2 pi jk/17) v_j,  j,k in [1,8].
template <int S, typename T>
FX_HD void fold17(cx<T>* x, const Bank<T>* __restrict__ bk)
{
    cx<T> u[9], v[9];
    const cx<T> x0 = x[0];
    cx<T> X0 = x0;
#pragma unroll
    for (int j = 1; j <= 8; ++j) {
        u[j] = x[j] + x[17 - j];
        v[j] = x[j] - x[17 - j];
        X0 = X0 + u[j];
    }
#pragma unroll
    for (int k = 1; k <= 8; ++k) {
        cx<T> A = x0, B = {T(0), T(0)};
#pragma unroll
        for (int j = 1; j <= 8; ++j) {
            const int m = (j * k) % 17;
            const T c = bk->cos17[m], s = bk->sin17[m];
            A.r += c * u[j].r;
            A.i += c * u[j].i;
            B.r += s * v[j].r;
            B.i += s * v[j].i;
        }
        const cx<T> sb = mul_si<S>(B);
        x[k] = A + sb;
        x[17 - k] = A - sb;
    }
    x[0] = X0;
}

// Folded dense 17 that hands each output to put(k, X_k) as soon as it is final, so a
// caller can store it immediately and the 17 results never coexist with u and v.
template <int S, typename T, typename Put>
FX_HD void fold17_put(const cx<T>* x, const Bank<T>* __restrict__ bk, Put put)
{
    cx<T> u[9], v[9];
    const cx<T> x0 = x[0];
    cx<T> X0 = x0;
#pragma unroll
    for (int j = 1; j <= 8; ++j) {
        u[j] = x[j] + x[17 - j];
        v[j] = x[j] - x[17 - j];
        X0 = X0 + u[j];
    }
    put(0, X0);
#pragma unroll
    for (int k = 1; k <= 8; ++k) {
        cx<T> A = x0, B = {T(0), T(0)};
#pragma unroll
        for (int j = 1; j <= 8; ++j) {
            const int m = (j * k) % 17;
            const T c = bk->cos17[m], s = bk->sin17[m];
            A.r += c * u[j].r;
            A.i += c * u[j].i;
            B.r += s * v[j].r;
            B.i += s * v[j].i;
        }
        const cx<T> sb = mul_si<S>(B);
        put(k, A + sb);
        put(17 - k, A - sb);
    }
}

// dft followed by put(k, X_k) for every k; the folded 17 interleaves the two.
template <int N, int S, typename T, int V17, typename Put>
FX_HD void dft_put(cx<T>* x, const Bank<T>* __restrict__ bk, Put put)
{
    if constexpr (N == 17 && V17 == 1) {
        fold17_put<S>(x, bk, put);
    } else {
        dft<N, S, T, V17>(x, bk);
#pragma unroll
        for (int k = 0; k < N; ++k) put(k, x[k]);
    }
}

template <int N, int S, typename T, int V17>
FX_HD void dft(cx<T>* x, const Bank<T>* __restrict__ bk)
{
    if constexpr (N == 1) {
    } else if constexpr (N == 2) {
        dft2(x);
    } else if constexpr (N == 3) {
        dft3<S>(x, bk);
    } else if constexpr (N == 4) {
        dft4<S>(x);
    } else if constexpr (N == 6) {
        if constexpr (V17 == 0) pfa<2, 3, S>(x, bk);
        else pfa<3, 2, S>(x, bk);
    } else if constexpr (N == 8) {
        if constexpr (V17 == 0) ct<2, 4, S>(x, bk);
        else ct<4, 2, S>(x, bk);
    } else if constexpr (N == 9) {
        ct<3, 3, S>(x, bk);
    } else if constexpr (N == 12) {
        if constexpr (V17 == 0) pfa<4, 3, S>(x, bk);
        else pfa<3, 4, S>(x, bk);
    } else if constexpr (N == 16) {
        if constexpr (V17 == 0) ct<4, 4, S>(x, bk);
        else ct<2, 8, S>(x, bk);
    } else if constexpr (N == 17) {
        if constexpr (V17 == 0) rader17<S>(x, bk);
        else fold17<S>(x, bk);
    } else if constexpr (N == 24) {
        if constexpr (V17 == 0) pfa<8, 3, S>(x, bk);
        else pfa<3, 8, S>(x, bk);
    } else if constexpr (N == 36) {
        if constexpr (V17 == 0) pfa<4, 9, S>(x, bk);
        else pfa<9, 4, S>(x, bk);
    } else if constexpr (N == 48) {
        if constexpr (V17 == 0) pfa<16, 3, S>(x, bk);
        else pfa<3, 16, S>(x, bk);
    } else {
        static_assert(N == 1, "unsupported line length");
    }
}

}  // namespace fx
