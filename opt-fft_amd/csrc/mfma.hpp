// DFT-as-dense-matrix variant on the matrix cores (complex64, L <= 17).
//
// A block of VPW volumes is staged in LDS (complex float).  Each axis pass is, per tile of
// 16 lines, the real product
//       [Yr]   [ Wr  -Wi ] [Xr]          W[j][i] = exp(S 2 pi i j i / N)
//       [Yi] = [ Wi   Wr ] [Xi]
// of a (2N x 2N) matrix (zero-padded to a multiple of 16) with a (2N x 16) operand, on
// v_mfma_f32_16x16x16 instructions.  Both operands are split into a high and a low
// half-precision part (x = xh + xl, xl = x - xh) and three products xh*Wh + xl*Wh + xh*Wl
// are accumulated in FP32 -- the compensated scheme; the dropped xl*Wl term is ~2^-32
// (bf16) or ~2^-44 (fp16) relative.  SPLIT = 0 uses bf16 parts (fp32 range, ~16 significant
// bits), SPLIT = 1 fp16 parts (~22 bits, range limited to |x| < 65504).  The matrix
// fragments come from the plan-time bank and live in VGPRs for the whole kernel.
//
// Operand layouts of v_mfma_f32_16x16x16{bf16_1k,f16} (64 lanes, 4 values per lane):
//   A: lane l holds A[l % 16][4 (l / 16) + i];  B: B[4 (l / 16) + i][l % 16];
//   D: D[4 (l / 16) + i][l % 16],  i = 0..3.
#pragma once
#include <stdint.h>

#include "kernels.hpp"

namespace fx {

typedef short mf_v4s __attribute__((ext_vector_type(4)));
typedef _Float16 mf_v4h __attribute__((ext_vector_type(4)));
typedef float mf_v4f __attribute__((ext_vector_type(4)));

// Round-to-nearest-even float -> bfloat16 bits (finite inputs).
__device__ __forceinline__ short bf16_bits(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (short)(u >> 16);
}
__device__ __forceinline__ float bf16_value(short h)
{
    return __uint_as_float((uint32_t)(unsigned short)h << 16);
}

template <int SPLIT>
struct MfmaOps;
template <>
struct MfmaOps<0> {
    using V = mf_v4s;
    // Lane q of hi/lo receives the high and low parts of x.
    __device__ __forceinline__ static void split(float x, V& hi, V& lo, int q)
    {
        const short h = bf16_bits(x);
        hi[q] = h;
        lo[q] = bf16_bits(x - bf16_value(h));
    }
    __device__ __forceinline__ static mf_v4f mma(V a, V b, mf_v4f c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
    }
};
template <>
struct MfmaOps<1> {
    using V = mf_v4h;
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

// Matrix fragments of one line length N for this lane: hi and lo parts, MT x KT tiles.
template <int N, int S, int SPLIT>
struct MfmaMatrix {
    static constexpr int K2 = 2 * N;
    static constexpr int T = (K2 + 15) / 16;  // tiles along m and k
    using Ops = MfmaOps<SPLIT>;
    typename Ops::V hi[T][T], lo[T][T];

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
        for (int mt = 0; mt < T; ++mt)
#pragma unroll
            for (int kt = 0; kt < T; ++kt)
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float a = entry(16 * mt + lane % 16, 16 * kt + 4 * (lane / 16) + q, bk);
                    Ops::split(a, hi[mt][kt], lo[mt][kt], q);
                }
    }
};

// One axis pass over nlines lines (line l starts at base(l), N points STRIDE apart), in
// place in LDS, tiles of 16 lines per wave.
template <int N, int STRIDE, int S, int SPLIT, int NT, typename BaseFn>
__device__ __forceinline__ void mfma_pass(cx<float>* sm, int nlines, BaseFn base,
                                          const MfmaMatrix<N, S, SPLIT>& A, float scl)
{
    using M = MfmaMatrix<N, S, SPLIT>;
    using Ops = MfmaOps<SPLIT>;
    constexpr int T = M::T;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ntiles = (nlines + 15) / 16;
    float* smf = (float*)sm;
    for (int tile = wave; tile < ntiles; tile += NT / 64) {
        const int line = tile * 16 + lane % 16;
        const bool valid = line < nlines;
        const int b = base(valid ? line : nlines - 1);
        typename Ops::V bh[T], bl[T];
#pragma unroll
        for (int kt = 0; kt < T; ++kt)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = 16 * kt + 4 * (lane / 16) + q;
                float v = 0.f;
                if (k < N) v = smf[2 * (b + k * STRIDE)];
                else if (k < 2 * N) v = smf[2 * (b + (k - N) * STRIDE) + 1];
                Ops::split(v, bh[kt], bl[kt], q);
            }
#pragma unroll
        for (int mt = 0; mt < T; ++mt) {
            mf_v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int kt = 0; kt < T; ++kt) {
                acc = Ops::mma(A.hi[mt][kt], bh[kt], acc);
                acc = Ops::mma(A.hi[mt][kt], bl[kt], acc);
                acc = Ops::mma(A.lo[mt][kt], bh[kt], acc);
            }
            if (valid) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int m = 16 * mt + 4 * (lane / 16) + q;
                    if (m < N) smf[2 * (b + m * STRIDE)] = acc[q] * scl;
                    else if (m < 2 * N) smf[2 * (b + (m - N) * STRIDE) + 1] = acc[q] * scl;
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
// Ends with a barrier, so a caller may reuse `sm` at once.
template <class C, int SPLIT>
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
        MfmaMatrix<C::N2, C::S, SPLIT> A;
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
        MfmaMatrix<C::N1, C::S, SPLIT> A;
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
        MfmaMatrix<C::N0, C::S, SPLIT> A;
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

template <class C, int SPLIT>
__global__ void __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::WPE, 8)))
    k_mfma(const cx<float>* __restrict__ in, cx<float>* __restrict__ out, long nvol,
           const Bank<float>* __restrict__ bk, float scl)
{
    static_assert(sizeof(typename C::T) == 4, "the matrix-core variant is complex64");
    __shared__ cx<float> sm[C::LDS];
    mfma_block<C, SPLIT>(blockIdx.x, nvol, in, out, bk, scl, false, sm);
}

}  // namespace fx

#include <fx_mfma_launch.hpp>
