// DFT-as-dense-matrix variant on the matrix cores (complex64, L <= 17).
//
// A block of VPW volumes is staged in LDS (complex float).  Each axis pass is, per tile of
// 16 lines, the real product
//       [Yr]   [ Wr  -Wi ] [Xr]          W[j][i] = exp(S 2 pi i j i / N)
//       [Yi] = [ Wi   Wr ] [Xi]
// of a (2N x 2N) matrix (rows zero-padded to a multiple of 16, columns to a multiple of
// the instruction's K) with a (2N x 16) operand.  K is 16 (v_mfma_f32_16x16x16*) or 32
// (gfx950 v_mfma_f32_16x16x32_*, KW = 32, for 2N > 16: a 17-point tile is 3 x 2
// instructions instead of 3 x 3, but its K pads 34 to 64 instead of 48, and those kernels
// hold more operand registers; they measured 10-25% slower at N = 12 and 17 and stay in the
// race as the "k32" variants).  Both operands are split into a high and a
// low half-precision part (x = xh + xl, xl = x - xh, each rounded to nearest even by the
// conversion instructions) and three products xh*Wh + xl*Wh + xh*Wl are accumulated in
// FP32 -- the compensated scheme; the dropped xl*Wl term and the rounding of the low parts
// are ~2^-18 of the result (bf16) or ~2^-24 (fp16).  SPLIT = 0 uses bf16 parts (~16
// significant bits, the FP32 exponent range), SPLIT = 1 fp16 parts (~22 bits).  fp16 parts
// hold only |x| < 65504 and the low part of a small x falls below the normal range, so with
// SPLIT = 1 every tile of 16 lines is multiplied, before the split, by the power of two that
// brings its largest magnitude into [0.5, 1) (a wave-wide maximum through DPP lane moves),
// and the result by the inverse power; both scalings are exact.  The matrix fragments come
// from the plan-time bank and live in VGPRs for the whole pass.
//
// Operand layouts, 64 lanes, Q = K/4 values per lane, q = 0..Q-1:
//   A: lane l holds A[l % 16][Q (l / 16) + q];  B: B[Q (l / 16) + q][l % 16];
//   D: D[4 (l / 16) + i][l % 16],  i = 0..3  (both K widths).
#pragma once
#include <stdint.h>

#include "kernels.hpp"

namespace fx {

typedef short mf_v4s __attribute__((ext_vector_type(4)));
typedef _Float16 mf_v4h __attribute__((ext_vector_type(4)));
typedef __bf16 mf_v8b __attribute__((ext_vector_type(8)));
typedef _Float16 mf_v8h __attribute__((ext_vector_type(8)));
typedef float mf_v4f __attribute__((ext_vector_type(4)));

// Split and product for SPLIT (0 bf16, 1 fp16) and the instruction's K width.
template <int SPLIT, int KW>
struct MfmaOps;
template <>
struct MfmaOps<0, 16> {
    using V = mf_v4s;
    static constexpr int Q = 4;
    __device__ __forceinline__ static void split(float x, V& hi, V& lo, int q)
    {
        const __bf16 h = (__bf16)x;
        hi[q] = __builtin_bit_cast(short, h);
        lo[q] = __builtin_bit_cast(short, (__bf16)(x - (float)h));
    }
    __device__ __forceinline__ static mf_v4f mma(V a, V b, mf_v4f c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
};
template <>
struct MfmaOps<1, 16> {
    using V = mf_v4h;
    static constexpr int Q = 4;
    __device__ __forceinline__ static void split(float x, V& hi, V& lo, int q)
    {
        const _Float16 h = (_Float16)x;
        hi[q] = h;
        lo[q] = (_Float16)(x - (float)h);
    }
    __device__ __forceinline__ static mf_v4f mma(V a, V b, mf_v4f c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
    }
};
template <>
struct MfmaOps<0, 32> {
    using V = mf_v8b;
    static constexpr int Q = 8;
    __device__ __forceinline__ static void split(float x, V& hi, V& lo, int q)
    {
        const __bf16 h = (__bf16)x;
        hi[q] = h;
        lo[q] = (__bf16)(x - (float)h);
    }
    __device__ __forceinline__ static mf_v4f mma(V a, V b, mf_v4f c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct MfmaOps<1, 32> {
    using V = mf_v8h;
    static constexpr int Q = 8;
    __device__ __forceinline__ static void split(float x, V& hi, V& lo, int q)
    {
        const _Float16 h = (_Float16)x;
        hi[q] = h;
        lo[q] = (_Float16)(x - (float)h);
    }
    __device__ __forceinline__ static mf_v4f mma(V a, V b, mf_v4f c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

// max(x, x moved by the DPP control CTRL) on the rows in ROWS (others keep x).
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_max(float x)
{
    const int y = __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, x), __builtin_bit_cast(int, x),
                                              CTRL, ROWS, 0xF, false);
    return fmaxf(x, __builtin_bit_cast(float, y));
}

// Largest value of v over the wave, uniform: quad, half-row and row butterflies, then the
// row broadcasts fold rows 0-2 into row 3, whose lane 63 holds the maximum.
__device__ __forceinline__ float wave_max(float v)
{
    v = dpp_max<0xB1, 0xF>(v);   // quad_perm [1,0,3,2]
    v = dpp_max<0x4E, 0xF>(v);   // quad_perm [2,3,0,1]
    v = dpp_max<0x141, 0xF>(v);  // row_half_mirror
    v = dpp_max<0x140, 0xF>(v);  // row_mirror
    v = dpp_max<0x142, 0xA>(v);  // row_bcast:15 into rows 1 and 3
    v = dpp_max<0x143, 0xC>(v);  // row_bcast:31 into rows 2 and 3
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

// Matrix fragments of one line length N for this lane: hi and lo parts, TM x TK tiles.
// KW_: instruction K, 16 (also for 0) or 32 (only where 2N > 16; 16 otherwise).
template <int N, int S, int SPLIT, int KW_ = 0>
struct MfmaMatrix {
    static constexpr int K2 = 2 * N;
    static constexpr int KW = (KW_ == 32 && K2 > 16) ? 32 : 16;   // instruction K
    static constexpr int TM = (K2 + 15) / 16;      // tiles along m (16 rows each)
    static constexpr int TK = (K2 + KW - 1) / KW;  // tiles along k (KW columns each)
    using Ops = MfmaOps<SPLIT, KW>;
    static constexpr int Q = Ops::Q;
    typename Ops::V hi[TM][TK], lo[TM][TK];

    // A[m][k] of the real block matrix.
    __device__ __forceinline__ static float entry(int m, int k, const Bank<float>* __restrict__ bk)
    {
        if (m >= K2 || k >= K2) return 0.f;
        const int j = m < N ? m : m - N, i = k < N ? k : k - N;
        const cx<float> w = bk->w[woff(N) + (j * i) % N];
        const float wr = w.r, wi = S < 0 ? w.i : -w.i;
        if (m < N) return k < N ? wr : -wi;
        return k < N ? wi : wr;
    }
    __device__ __forceinline__ void build(int lane, const Bank<float>* __restrict__ bk)
    {
#pragma unroll
        for (int mt = 0; mt < TM; ++mt)
#pragma unroll
            for (int kt = 0; kt < TK; ++kt)
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const float a = entry(16 * mt + lane % 16, KW * kt + Q * (lane / 16) + q, bk);
                    Ops::split(a, hi[mt][kt], lo[mt][kt], q);
                }
    }
};

// One axis pass over nlines lines (line l starts at base(l), N points STRIDE apart), in
// place in LDS, tiles of 16 lines per wave.
template <int N, int STRIDE, int S, int SPLIT, int NT, int KW_, typename BaseFn>
__device__ __forceinline__ void mfma_pass(cx<float>* sm, int nlines, BaseFn base,
                                          const MfmaMatrix<N, S, SPLIT, KW_>& A, float scl)
{
    using M = MfmaMatrix<N, S, SPLIT, KW_>;
    using Ops = typename M::Ops;
    constexpr int TM = M::TM, TK = M::TK, KW = M::KW, Q = M::Q;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ntiles = (nlines + 15) / 16;
    float* smf = (float*)sm;
    for (int tile = wave; tile < ntiles; tile += NT / 64) {
        const int line = tile * 16 + lane % 16;
        const bool valid = line < nlines;
        const int b = base(valid ? line : nlines - 1);
        float v[TK][Q];
        float mx = 0.f;
#pragma unroll
        for (int kt = 0; kt < TK; ++kt)
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int k = KW * kt + Q * (lane / 16) + q;
                float x = 0.f;
                if (k < N) x = smf[2 * (b + k * STRIDE)];
                else if (k < 2 * N) x = smf[2 * (b + (k - N) * STRIDE) + 1];
                v[kt][q] = x;
                mx = fmaxf(mx, fabsf(x));
            }
        float up = 1.f, down = scl;
        if constexpr (SPLIT == 1) {
            mx = wave_max(mx);
            int e = 0;
            if (mx > 0.f) (void)frexpf(mx, &e);
            up = ldexpf(1.f, -e);
            down = ldexpf(scl, e);
        }
        typename Ops::V bh[TK], bl[TK];
#pragma unroll
        for (int kt = 0; kt < TK; ++kt)
#pragma unroll
            for (int q = 0; q < Q; ++q) Ops::split(v[kt][q] * up, bh[kt], bl[kt], q);
#pragma unroll
        for (int mt = 0; mt < TM; ++mt) {
            mf_v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < TK; ++kt) {
                acc = Ops::mma(A.hi[mt][kt], bh[kt], acc);
                acc = Ops::mma(A.hi[mt][kt], bl[kt], acc);
                acc = Ops::mma(A.lo[mt][kt], bh[kt], acc);
            }
            if (valid) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int m = 16 * mt + 4 * (lane / 16) + q;
                    if (m < N) smf[2 * (b + m * STRIDE)] = acc[q] * down;
                    else if (m < 2 * N) smf[2 * (b + (m - N) * STRIDE) + 1] = acc[q] * down;
                }
            }
        }
    }
}

template <class C, int SPLIT>
struct MfmaCfg : C {
    static constexpr int SPLITV = SPLIT;
};

// One block of VPW volumes: stage, the three axis passes, store.  CONJ: conjugate on load
// and on store (the backward transform through forward matrices, for the mixed launch).
// Ends with a barrier, so a caller may reuse `sm` at once.  KW: instruction K (0 or 16, 32).
template <class C, int SPLIT, int KW = 0>
__device__ __forceinline__ void mfma_block(long bid, long nvol, const cx<float>* __restrict__ in,
                                           cx<float>* __restrict__ out,
                                           const Bank<float>* __restrict__ bk, float scl,
                                           bool conj, cx<float>* sm)
{
    using K = Fused<C>;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int nv = K::nvols(bid, nvol);
    K::stage(tid, bid, nvol, in, sm);
    __syncthreads();
    if (conj) {
        for (int e = tid; e < nv * C::V; e += C::NT) sm[K::slot(e)].i = -sm[K::slot(e)].i;
        __syncthreads();
    }
    {
        MfmaMatrix<C::N2, C::S, SPLIT, KW> A;
        A.build(lane, bk);
        mfma_pass<C::N2, 1, C::S, SPLIT, C::NT>(
            sm, nv * C::N0 * C::N1,
            [](int l) {
                const int v = l / (C::N0 * C::N1);
                const int r = l - v * (C::N0 * C::N1);
                return v * C::VS + (r / C::N1) * C::XS + (r % C::N1) * C::YS;
            },
            A, 1.f);
    }
    __syncthreads();
    {
        MfmaMatrix<C::N1, C::S, SPLIT, KW> A;
        A.build(lane, bk);
        mfma_pass<C::N1, C::YS, C::S, SPLIT, C::NT>(
            sm, nv * C::N0 * C::N2,
            [](int l) {
                const int v = l / (C::N0 * C::N2);
                const int r = l - v * (C::N0 * C::N2);
                return v * C::VS + (r / C::N2) * C::XS + (r % C::N2);
            },
            A, 1.f);
    }
    __syncthreads();
    {
        MfmaMatrix<C::N0, C::S, SPLIT, KW> A;
        A.build(lane, bk);
        mfma_pass<C::N0, C::XS, C::S, SPLIT, C::NT>(
            sm, nv * C::PLANE,
            [](int l) {
                const int v = l / C::PLANE;
                const int q = l - v * C::PLANE;
                return v * C::VS + (q / C::N2) * C::YS + (q % C::N2);
            },
            A, scl);
    }
    __syncthreads();
    cx<float>* __restrict__ dst = out + bid * C::VPW * (long)C::V;
    const float ci = conj ? -1.f : 1.f;
    for (int e = tid; e < nv * C::V; e += C::NT) {
        const cx<float> y = sm[K::slot(e)];
        st_stream(dst + e, cx<float>{y.r, ci * y.i});
    }
    __syncthreads();
}

template <class C, int SPLIT, int KW = 0>
__global__ void __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::WPE, 8)))
    k_mfma(const cx<float>* __restrict__ in, cx<float>* __restrict__ out, long nvol,
           const Bank<float>* __restrict__ bk, float scl)
{
    static_assert(sizeof(typename C::T) == 4, "the matrix-core variant is complex64");
    __shared__ cx<float> sm[C::LDS];
    mfma_block<C, SPLIT, KW>(blockIdx.x, nvol, in, out, bk, scl, false, sm);
}

}  // namespace fx

#include <fx_mfma_launch.hpp>
