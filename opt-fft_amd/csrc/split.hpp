// Single-pass kernel for volumes that fit in a CU's LDS and registers together but not in
// LDS alone: 24^3 complex128 is 221 KB and 36^3 complex64 373 KB, against 160 KB of LDS.
// The two-pass pencil moves every element through HBM twice; here each element is read
// once and written once.
//
// Per volume (one workgroup, LPT x-lines per thread):
//   * x-lines (y, z) HBM -> VGPRs (a wave's load instruction reads 64 consecutive elements),
//     x DFT in registers;
//   * R rounds, each over XR = N0 / R x-planes: the round's outputs of the x DFT go to LDS
//     (XR planes, rows padded by PZ), z lines and then y lines are transformed there, and the
//     y pass stores straight to HBM (a store instruction writes runs of N2 elements).  The
//     z and y lines run in two in-place stages (ct2_stage1/2), so a thread holds 8 (or 9)
//     points of a line rather than all 24 (or 36);
//   * the next volume's x-lines are loaded into the VGPRs freed by the last round's LDS
//     write, so they are in flight during that round's z and y passes.
// During every round but the last, the x-planes of the later rounds stay in VGPRs beside the
// line being transformed (at 24^3 complex128, R = 2: 12 of 24 points of a line, 48 VGPRs;
// at 36^3 complex64, R = 3, two lines: 24 of 36 points each, 96 VGPRs, after round 0).
// Workgroups of the persistent form (P > 0) claim volumes from the launch's claim queues.
#pragma once
#include <utility>

#include "kernels.hpp"

namespace fx {

// SP1 / SP2: the next volume's x-line points [0, SP1) are loaded after the last round's LDS
// write, [SP1, SP2) after its z pass and [SP2, N0) after its y pass, so that fewer
// prefetched registers are live while the line passes run.
template <int N0_, int N1_, int N2_, typename T_, int S_, int NT_, int R_, int PZ_, int P_,
          int SP1_, int SP2_, int ALG_ = 0>
struct SplitCfg {
    using T = T_;
    static constexpr int N0 = N0_, N1 = N1_, N2 = N2_, S = S_, NT = NT_, R = R_, PZ = PZ_;
    static constexpr int P = P_, V17 = ALG_, VPW = 1, SP1 = SP1_, SP2 = SP2_;
    static_assert(0 <= SP1_ && SP1_ <= SP2_ && SP2_ <= N0_, "every x-line point loaded once");
    static constexpr int V = N0 * N1 * N2;
    static constexpr int PLANE = N1 * N2;
    static constexpr int XR = N0 / R;      // x-planes per round
    static constexpr int YS = N2 + PZ;     // LDS pitch between z rows
    static constexpr int XS = N1 * YS;     // LDS pitch between x-planes
    static constexpr int LDS = XR * XS;    // elements
    static constexpr int LPT = (PLANE + NT - 1) / NT;  // x-lines per thread
    static constexpr int WPE = waves_per_eu(LDS * (int)sizeof(cx<T>), NT, P);
    static constexpr int HN = 3;           // nontemporal loads and stores
    static_assert(N0 % R == 0, "the rounds split the x-planes evenly");
};

// No instruction is scheduled across this point (the groups of a two-stage line would
// otherwise have their LDS reads hoisted together, holding all N points in registers).
FX_HD void sched_fence()
{
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_amdgcn_sched_barrier(0);
#endif
}

// Line DFT of N = A * B points in LDS (point n at b0 + n * st), in two stages that hold A,
// then B, points in registers instead of N (Cooley-Tukey over n = a * B + b):
//   stage 1, for each b: DFT_A over a of x[a * B + b], times W_N^(b * k1), written back to
//     positions k1 * B + b (the same positions: in place);
//   stage 2, for each k1: DFT_B over b of positions k1 * B + b gives X[k1 + A * k2], handed
//     to put(k, X_k) as soon as computed.
// Written back in place by stage 2, X_k sits at position (k % A) * B + k / A.
template <int A, int B, int S, typename T>
FX_HD void ct2_stage1(cx<T>* sm, int b0, int st, const Bank<T>* __restrict__ bk)
{
#pragma unroll
    for (int b = 0; b < B; ++b) {
        cx<T> t[A];
#pragma unroll
        for (int a = 0; a < A; ++a) t[a] = sm[b0 + (a * B + b) * st];
        dft<A, S, T>(t, bk);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) sm[b0 + (k1 * B + b) * st] = tw<A * B, S, T>(t[k1], b * k1, bk);
        sched_fence();
    }
}
template <int A, int B, int S, typename T, typename Put>
FX_HD void ct2_stage2(const cx<T>* sm, int b0, int st, const Bank<T>* __restrict__ bk, Put put)
{
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
        cx<T> t[B];
#pragma unroll
        for (int b = 0; b < B; ++b) t[b] = sm[b0 + (k1 * B + b) * st];
        dft<B, S, T>(t, bk);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) put(k1 + A * k2, t[k2]);
        sched_fence();
    }
}
// Factors of the two-stage line passes.
template <int N>
struct Ct2;
template <>
struct Ct2<24> {
    static constexpr int A = 3, B = 8;
};
template <>
struct Ct2<36> {
    static constexpr int A = 4, B = 9;
};

template <class C>
struct Split {
    using T = typename C::T;
    using Regs = cx<T>[C::LPT][C::N0];

    // Loads of points [J0, J1) of volume b's x-lines.  Lanes past the last line re-read a
    // valid element.
    template <int J0 = 0, int J1 = C::N0>
    static FX_HD void load(int tid, long b, const cx<T>* __restrict__ in, Regs& R)
    {
        const cx<T>* __restrict__ src = in + b * (long)C::V;
#pragma unroll
        for (int i = 0; i < C::LPT; ++i) {
            int l = tid + i * C::NT;
            l = l < C::PLANE ? l : C::PLANE - 1;
#pragma unroll
            for (int j = J0; j < J1; ++j) R[i][j] = ld_hint<C::HN & 1>(src + j * C::PLANE + l);
        }
    }
    // x DFT of every line in registers; each line's round-0 x-planes go to LDS as soon as
    // it is transformed, so with two lines per thread the first line holds only its later
    // rounds' points while the second is transformed.
    static FX_HD void xdft_put0(int tid, Regs& R, cx<T>* sm, const Bank<T>* __restrict__ bk)
    {
#pragma unroll
        for (int i = 0; i < C::LPT; ++i) {
            const int l = tid + i * C::NT;
            if (l < C::PLANE) {
                dft<C::N0, C::S, T, C::V17>(R[i], bk);
                const int y = l / C::N2;
                const int z = l - y * C::N2;
#pragma unroll
                for (int xl = 0; xl < C::XR; ++xl) sm[xl * C::XS + y * C::YS + z] = R[i][xl];
            }
        }
    }
    // Round RD: x-planes [RD * XR, (RD + 1) * XR) of the x DFT's output to LDS.
    template <int RD>
    static FX_HD void put(int tid, const Regs& R, cx<T>* sm)
    {
#pragma unroll
        for (int i = 0; i < C::LPT; ++i) {
            const int l = tid + i * C::NT;
            if (l < C::PLANE) {
                const int y = l / C::N2;
                const int z = l - y * C::N2;
#pragma unroll
                for (int xl = 0; xl < C::XR; ++xl) sm[xl * C::XS + y * C::YS + z] = R[i][RD * C::XR + xl];
            }
        }
    }
    // z lines (xl, y), contiguous in LDS, in two stages (Ct2): the outputs stay in the
    // stage order, z frequency k at row position (k % A) * B + k / A.
    static FX_HD void zpass(int tid, cx<T>* sm, const Bank<T>* __restrict__ bk)
    {
        using F = Ct2<C::N2>;
        for (int l = tid; l < C::XR * C::N1; l += C::NT) {
            const int xl = l / C::N1;
            const int y = l - xl * C::N1;
            const int b0 = xl * C::XS + y * C::YS;
            ct2_stage1<F::A, F::B, C::S>(sm, b0, 1, bk);
            ct2_stage2<F::A, F::B, C::S>(sm, b0, 1, bk, [&](int k, cx<T> w) {
                sm[b0 + (k % F::A) * F::B + k / F::A] = w;
            });
        }
    }
    // y lines (xl, row position p), stride YS, in two stages: LDS -> HBM at x = rd * XR + xl,
    // z frequency kz = p / B + A * (p % B), every output multiplied by scl.  Lanes take
    // consecutive positions p, so a store instruction covers whole z rows (in permuted lane
    // order).
    static FX_HD void ypass(int tid, int rd, long b, cx<T>* sm, cx<T>* __restrict__ out,
                            const Bank<T>* __restrict__ bk, T scl)
    {
        using Fz = Ct2<C::N2>;
        using Fy = Ct2<C::N1>;
        cx<T>* __restrict__ dst = out + b * (long)C::V + (long)rd * C::XR * C::PLANE;
        for (int l = tid; l < C::XR * C::N2; l += C::NT) {
            const int xl = l / C::N2;
            const int p = l - xl * C::N2;
            const int kz = p / Fz::B + Fz::A * (p % Fz::B);
            const int b0 = xl * C::XS + p;
            ct2_stage1<Fy::A, Fy::B, C::S>(sm, b0, C::YS, bk);
            cx<T>* __restrict__ o = dst + xl * C::PLANE + kz;
            // Always scaled: a product by 1 is exact, and one store path keeps the
            // registers of the second one free.
            ct2_stage2<Fy::A, Fy::B, C::S>(sm, b0, C::YS, bk, [&](int k, cx<T> w) {
                st_hint<0>(o + k * C::N2, scale(w, scl));
            });
        }
    }
};

// One round that is not the last: x-planes to LDS (round 0's are already there, written by
// xdft_put0), z pass, y pass to HBM.  Barriers between the phases and after the y pass (the
// next round overwrites the LDS).
template <class C, int RD>
__device__ __forceinline__ void split_round(int tid, long b, typename Split<C>::Regs& R,
                                            cx<typename C::T>* sm,
                                            cx<typename C::T>* __restrict__ out,
                                            const Bank<typename C::T>* __restrict__ bk,
                                            typename C::T scl)
{
    using K = Split<C>;
    if constexpr (RD > 0) K::template put<RD>(tid, R, sm);
    __syncthreads();
    K::zpass(tid, sm, bk);
    __syncthreads();
    K::ypass(tid, RD, b, sm, out, bk, scl);
    __syncthreads();
}

template <class C, int... RD>
__device__ __forceinline__ void split_rounds(std::integer_sequence<int, RD...>, int tid, long b,
                                             typename Split<C>::Regs& R, cx<typename C::T>* sm,
                                             cx<typename C::T>* __restrict__ out,
                                             const Bank<typename C::T>* __restrict__ bk,
                                             typename C::T scl)
{
    (split_round<C, RD>(tid, b, R, sm, out, bk, scl), ...);
}

// Persistent (P > 0, claims through `work`) or one volume per workgroup (P = 0, work NULL,
// no prefetch).
template <class C>
__global__ void __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::WPE, 8)))
    k_split(const cx<typename C::T>* __restrict__ in, cx<typename C::T>* __restrict__ out,
            long nvol, const Bank<typename C::T>* __restrict__ bk, typename C::T scl,
            unsigned* __restrict__ work)
{
    using K = Split<C>;
    __shared__ cx<typename C::T> sm[C::LDS];
    __shared__ long s_next;
    typename K::Regs R;
    long b = blockIdx.x;
    if (b >= nvol) return;
    K::load((int)threadIdx.x, b, in, R);
    for (;;) {
        // The thread index passes through an empty asm each iteration, so the compiler does
        // not keep its address arithmetic (24 load offsets, LDS slots) live across the loop.
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        long claim = 0;
        if (tid == 0) claim = claim_block(work, nvol, b);
        K::xdft_put0(tid, R, sm, bk);
        split_rounds<C>(std::make_integer_sequence<int, C::R - 1>{}, tid, b, R, sm, out, bk, scl);
        if constexpr (C::R > 1) K::template put<C::R - 1>(tid, R, sm);
        if (tid == 0) s_next = claim;
        __syncthreads();
        const long nb = C::P > 0 ? s_next : nvol;
        if (C::P > 0 && nb < nvol) K::template load<0, C::SP1>(tid, nb, in, R);
        K::zpass(tid, sm, bk);
        if (C::P > 0 && nb < nvol) K::template load<C::SP1, C::SP2>(tid, nb, in, R);
        __syncthreads();
        K::ypass(tid, C::R - 1, b, sm, out, bk, scl);
        if (nb >= nvol) break;
        K::template load<C::SP2, C::N0>(tid, nb, in, R);
        __syncthreads();
        b = nb;
    }
}

}  // namespace fx

#include <fx_split_launch.hpp>
