// 8^3 with the volume in one wavefront's registers: the y and z axes through lane
// exchanges instead of LDS.
//
// Lane l (0..63) of a wave owns the eight x-points of line (y, z) = (l / 8, l % 8): the x
// pass loads them straight from HBM (for fixed x the wave reads 64 consecutive elements, one
// contiguous 1 KiB run at complex128) and runs the 8-point DFT in registers.  The z axis
// lives in lane bits 0-2 and the y axis in bits 3-5, so each is an 8-point DFT across lanes:
// three radix-2 decimation-in-frequency stages, the stage on lane bit m pairing lane l with
// l ^ m through a lane exchange (X in Shfl8Cfg),
//       lane bit clear:  a + b          lane bit set:  (b - a) * W_{2h}^(j mod h)
// (a this lane's value, b the partner's, h = the stage's half size, j the lane's index
// along the axis), which leaves frequency bitrev3(j) on lane j.  PERM = 1 then permutes the
// lanes back to natural order with one more shuffle and stores contiguous runs; PERM = 0
// stores from the bit-reversed lanes (each store instruction still covers one contiguous
// 1 KiB run, its lanes permuted).  No LDS and no barrier: waves are independent, W of
// them per workgroup, one volume each.
//
// The phases are __host__ __device__ functions of (lane, registers, partner registers), so
// the host emulator runs the same arithmetic with the exchanges done over an array of 64
// lanes.  Launchers come from <fx_shfl_launch.hpp>, found on the include path.
#pragma once
#include "codelets.hpp"
#include "registry.hpp"

namespace fx {

// X: how a stage reaches its partner lane, 0 through ds_bpermute (the LDS crossbar), 1 with
// VALU moves: DPP within a row of 16 lanes (l ^ 1, l ^ 2: quad permutations; l ^ 4: a
// half-row mirror then a quad reversal; l ^ 8: a row rotation by 8) and gfx950's
// v_permlane16_swap / v_permlane32_swap for l ^ 16 and l ^ 32.
template <typename T_, int S_, int W_, int PERM_, int X_ = 1>
struct Shfl8Cfg {
    using T = T_;
    static constexpr int S = S_, W = W_, PERM = PERM_, X = X_;
    static constexpr int NT = 64 * W;
};

FX_HD int bitrev3(int j) { return ((j & 1) << 2) | (j & 2) | ((j >> 2) & 1); }

template <class C>
struct Shfl8 {
    using T = typename C::T;
    using Regs = cx<T>[8];

    // x pass: load the lane's line (8 strided points) and transform it in registers.
    static FX_HD void load_x(int lane, const cx<T>* __restrict__ vol, Regs r,
                             const Bank<T>* __restrict__ bk)
    {
#pragma unroll
        for (int x = 0; x < 8; ++x) r[x] = ld_stream(vol + x * 64 + lane);
        dft<8, C::S, T>(r, bk);
    }

    // One radix-2 DIF stage on lane bit m of the axis whose index is lane bits [base, base+3).
    static FX_HD void stage(int lane, int m, int base, Regs r, const Regs partner,
                            const Bank<T>* __restrict__ bk)
    {
        const int j = (lane >> base) & 7;
        const int h = m >> base;  // 4, 2 or 1
        const bool hi = (lane & m) != 0;
        // twiddle W_8^((j mod h) * 4 / h) for the upper half, 1 for the lower
        const int e = hi ? (j & (h - 1)) * (4 / h) : 0;
        cx<T> w = bk->w[woff(8) + e];
        if (C::S > 0) w.i = -w.i;
#pragma unroll
        for (int x = 0; x < 8; ++x) {
            const cx<T> t = hi ? partner[x] - r[x] : r[x] + partner[x];
            r[x] = hi ? cmul(t, w) : t;
        }
    }

    // Source lane of the natural-order permutation: lane l takes frequency (ky, kz) =
    // (l / 8, l % 8) from the lane holding it.
    static FX_HD int natural_src(int lane)
    {
        return (bitrev3(lane >> 3) << 3) | bitrev3(lane & 7);
    }

    // Store: frequency (kx, ky, kz) at kx * 64 + ky * 8 + kz.
    static FX_HD void store(int lane, cx<T>* __restrict__ vol, const Regs r, T scl)
    {
        const int o = C::PERM ? lane : natural_src(lane);
#pragma unroll
        for (int x = 0; x < 8; ++x) st_stream(vol + x * 64 + o, scale(r[x], scl));
    }
};

#ifdef __HIP_DEVICE_COMPILE__
template <typename T>
__device__ __forceinline__ cx<T> shfl_xor_cx(cx<T> v, int m)
{
    return cx<T>{__shfl_xor(v.r, m), __shfl_xor(v.i, m)};
}
template <int CTRL>
__device__ __forceinline__ unsigned dpp_move(unsigned v)
{
    return (unsigned)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, 0xF, 0xF, false);
}
// The value of lane l ^ M, with VALU moves (the permlane swaps return this lane's value and
// its partner's in some order, so their XOR with this lane's value is the partner's; the
// two results are combined in integer operations only: a float bit-cast of the second one
// compiled to a copy of the first in this compiler's gfx950 code).
template <int M>
__device__ __forceinline__ unsigned lane_xor(unsigned v)
{
    if constexpr (M == 32) {
        const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        return r[0] ^ r[1] ^ v;
    } else if constexpr (M == 16) {
        const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        return r[0] ^ r[1] ^ v;
    } else if constexpr (M == 8) {
        return dpp_move<0x128>(v);              // row_ror:8
    } else if constexpr (M == 4) {
        return dpp_move<0x1B>(dpp_move<0x141>(v));  // row_half_mirror, quad_perm [3,2,1,0]
    } else if constexpr (M == 2) {
        return dpp_move<0x4E>(v);               // quad_perm [2,3,0,1]
    } else {
        static_assert(M == 1, "lane bit");
        return dpp_move<0xB1>(v);               // quad_perm [1,0,3,2]
    }
}
template <int M, typename T>
__device__ __forceinline__ cx<T> lane_xor_cx(cx<T> v)
{
    constexpr int NW = (int)(sizeof(cx<T>) / 4);
    struct Words {
        unsigned w[NW];
    };
    Words a = __builtin_bit_cast(Words, v);
#pragma unroll
    for (int i = 0; i < NW; ++i) a.w[i] = lane_xor<M>(a.w[i]);
    return __builtin_bit_cast(cx<T>, a);
}
// One radix-2 stage on lane bit M (axis bits from BASE): partner values, then the butterfly.
template <class C, int M, int BASE>
__device__ __forceinline__ void shfl_stage(int lane, typename Shfl8<C>::Regs r,
                                           typename Shfl8<C>::Regs p,
                                           const Bank<typename C::T>* __restrict__ bk)
{
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        if constexpr (C::X == 1) p[x] = lane_xor_cx<M>(r[x]);
        else p[x] = shfl_xor_cx(r[x], M);
    }
    Shfl8<C>::stage(lane, M, BASE, r, p, bk);
}
template <typename T>
__device__ __forceinline__ cx<T> shfl_cx(cx<T> v, int src)
{
    return cx<T>{__shfl(v.r, src), __shfl(v.i, src)};
}
#endif

template <class C>
__global__ void __launch_bounds__(C::NT)
    k_shfl8(const cx<typename C::T>* __restrict__ in, cx<typename C::T>* __restrict__ out,
            long nvol, const Bank<typename C::T>* __restrict__ bk, typename C::T scl)
{
#ifdef __HIP_DEVICE_COMPILE__
    using K = Shfl8<C>;
    using T = typename C::T;
    const int lane = threadIdx.x & 63;
    const long v = (long)blockIdx.x * C::W + (threadIdx.x >> 6);
    if (v >= nvol) return;  // whole waves exit: no barrier below
    typename K::Regs r, p;
    K::load_x(lane, in + v * 512, r, bk);
    shfl_stage<C, 4, 0>(lane, r, p, bk);
    shfl_stage<C, 2, 0>(lane, r, p, bk);
    shfl_stage<C, 1, 0>(lane, r, p, bk);
    shfl_stage<C, 32, 3>(lane, r, p, bk);
    shfl_stage<C, 16, 3>(lane, r, p, bk);
    shfl_stage<C, 8, 3>(lane, r, p, bk);
    if (C::PERM) {
        const int src = K::natural_src(lane);
#pragma unroll
        for (int x = 0; x < 8; ++x) r[x] = shfl_cx(r[x], src);
    }
    K::store(lane, out + v * 512, r, (T)scl);
#endif
}

}  // namespace fx

#include <fx_shfl_launch.hpp>
