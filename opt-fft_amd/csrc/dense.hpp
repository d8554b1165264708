// Dense row-column floor: every axis pass computes each output of a line as a direct
// O(N) sum over the line, X[k] = sum_j x[j] W_N^(j k mod N), with no factorization.  One
// thread owns one line: it reads the N inputs, writes the N outputs in place (the first
// pass reads `in` and writes `out`, the other two work in place on `out`).  Three launches
// through HBM, deliberately not clever: the floor that validates the harness and that every
// other kernel is measured against (the role of bench/gpu/impl_4/baseline_gpu.cu:1-8,
// written here line-per-thread so it needs no scratch buffer).
#pragma once
#include "kernels.hpp"

namespace fx {

template <int N0_, int N1_, int N2_, typename T_, int S_>
struct DenseCfg {
    using T = T_;
    static constexpr int N0 = N0_, N1 = N1_, N2 = N2_, S = S_;
    static constexpr int V = N0 * N1 * N2;
    static constexpr int NT = 256;
};

template <class C>
struct Dense {
    using T = typename C::T;

    // Line `l` of axis AX (0 = x, 1 = y, 2 = z) over nvol volumes, from src to dst.
    template <int AX>
    static FX_HD void line(long l, long nvol, const cx<T>* src, cx<T>* dst,
                           const Bank<T>* __restrict__ bk, T scl)
    {
        constexpr int N = AX == 0 ? C::N0 : AX == 1 ? C::N1 : C::N2;
        constexpr long STRIDE = AX == 0 ? (long)C::N1 * C::N2 : AX == 1 ? C::N2 : 1;
        constexpr long PER_VOL = C::V / N;  // lines of this axis per volume
        if (l >= nvol * PER_VOL) return;
        const long v = l / PER_VOL;
        const long r = l - v * PER_VOL;
        long base;
        if (AX == 0) base = r;                                   // (y, z)
        else if (AX == 1) base = (r / C::N2) * C::N1 * C::N2 + r % C::N2;  // (x, z)
        else base = r * C::N2;                                   // (x, y)
        base += v * C::V;
        cx<T> x[N];
        for (int j = 0; j < N; ++j) x[j] = src[base + j * STRIDE];
        for (int k = 0; k < N; ++k) {
            cx<T> acc{T(0), T(0)};
            for (int j = 0; j < N; ++j) {
                cx<T> w = bk->w[woff(N) + (j * k) % N];
                if (C::S > 0) w.i = -w.i;
                acc = acc + cmul(x[j], w);
            }
            dst[base + k * STRIDE] = scale(acc, scl);
        }
    }
};

template <class C, int AX>
__global__ void __launch_bounds__(C::NT)
    k_dense(const cx<typename C::T>* src, cx<typename C::T>* dst, long nvol,
            const Bank<typename C::T>* __restrict__ bk, typename C::T scl)
{
    Dense<C>::template line<AX>((long)blockIdx.x * C::NT + threadIdx.x, nvol, src, dst, bk, scl);
}

}  // namespace fx

#include <fx_dense_launch.hpp>
