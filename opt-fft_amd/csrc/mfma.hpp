// DFT-as-dense-matrix variant on the matrix cores (complex64, L <= 17).
//
// A block of VPW volumes is staged in LDS (complex float).  Each axis pass is, per tile of
// 16 lines, the real product Y = A X of a (2N x 2N) matrix with a (2N x 16) operand whose
// rows interleave real and imaginary parts (row 2j holds Re x_j, row 2j+1 Im x_j; the same
// for the outputs), so that
//       A[2o][2j] = Wr   A[2o][2j+1] = -Wi   A[2o+1][2j] = Wi   A[2o+1][2j+1] = Wr,
// W[o][j] = exp(S 2 pi i o j / N), rows zero-padded to a multiple of 16 and columns to a
// multiple of the instruction's K.  K is 16 (v_mfma_f32_16x16x16*) or 32 (gfx950
// v_mfma_f32_16x16x32_*, KW = 32, for 2N > 16: a 17-point tile is 3 x 2 instructions instead
// of 3 x 3, but its K pads 34 to 64 instead of 48; the race picks per geometry).  With the interleaved rows a
// lane's operand values are whole complex points, one 8-byte LDS read each, at per-lane
// offsets computed once per pass; the padded positions read point N-1 again (their matrix
// columns are zero, so any finite value gives the same sums).  Both operands are split into
// a high and a low half-precision part (x = xh + xl, xl = x - xh, each rounded to nearest
// even by the conversion instructions) and three products xh*Wh + xl*Wh + xh*Wl are
// accumulated in FP32 -- the compensated scheme; the dropped xl*Wl term and the rounding of
// the low parts are ~2^-18 of the result (bf16) or ~2^-24 (fp16).  SPLIT = 0 uses bf16 parts
// (~16 significant bits, the FP32 exponent range), SPLIT = 1 fp16 parts (~22 bits).  fp16
// parts hold only |x| < 65504 and the low part of a small x falls below the normal range, so
// with SPLIT = 1 every line is multiplied, before the split, by the power of two that brings
// its largest magnitude into [0.5, 1) (a maximum over the four lanes holding the line,
// through two lane-swap moves), and its outputs by the inverse power; both scalings are
// exact.  The matrix fragments come from the plan-time bank, are built once per workgroup
// and stay in VGPRs.  17-point lines may instead take the conjugate-pair fold (Fold17Matrix
// below).
//
// Operand layouts, 64 lanes, Q = K/4 values per lane, q = 0..Q-1:
//   A: lane l holds A[l % 16][Q (l / 16) + q];  B: B[Q (l / 16) + q][l % 16];
//   D: D[4 (l / 16) + i][l % 16],  i = 0..3  (both K widths).
// So in a k-tile lane l holds Q/2 consecutive points of line l % 16, and in an m-tile two
// consecutive output points.
#pragma once
#include <stdint.h>

#include <type_traits>

#include "kernels.hpp"

namespace fx {

typedef short mf_v4s __attribute__((ext_vector_type(4)));
typedef _Float16 mf_v4h __attribute__((ext_vector_type(4)));
typedef __bf16 mf_v8b __attribute__((ext_vector_type(8)));
typedef _Float16 mf_v8h __attribute__((ext_vector_type(8)));
typedef float mf_v4f __attribute__((ext_vector_type(4)));
typedef float mf_v2f __attribute__((ext_vector_type(2)));
typedef __bf16 mf_v2b __attribute__((ext_vector_type(2)));
typedef _Float16 mf_v2h __attribute__((ext_vector_type(2)));
typedef unsigned mf_v2u __attribute__((ext_vector_type(2)));
typedef unsigned mf_v4u __attribute__((ext_vector_type(4)));

// Two values split at once into packed high and low parts (one pair conversion each, the
// same rounding as the element-wise split): SPLIT 0 bf16, 1 fp16.
template <int SPLIT>
__device__ __forceinline__ void split_pair(float x0, float x1, unsigned& hi, unsigned& lo)
{
    using H2 = typename std::conditional<SPLIT == 0, mf_v2b, mf_v2h>::type;
    const mf_v2f x = {x0, x1};
    const H2 h = __builtin_convertvector(x, H2);
    const mf_v2f r = x - __builtin_convertvector(h, mf_v2f);
    hi = __builtin_bit_cast(unsigned, h);
    lo = __builtin_bit_cast(unsigned, __builtin_convertvector(r, H2));
}

// Q/2 packed words as the instruction's operand vector.
template <typename V, int NW>
__device__ __forceinline__ V pack_words(const unsigned (&w)[NW])
{
    if constexpr (NW == 2) return __builtin_bit_cast(V, mf_v2u{w[0], w[1]});
    else return __builtin_bit_cast(V, mf_v4u{w[0], w[1], w[2], w[3]});
}

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

// The lane exchanges below use gfx950's v_permlane16_swap (rows 0 <-> 1 and 2 <-> 3 of 16
// lanes) and v_permlane32_swap (rows 0, 1 <-> 2, 3) with both operands holding v: every lane
// gets back its own value and that of lane l ^ 16 (or l ^ 32), in an order that differs
// between the rows.  VALU moves, no LDS round trip.  The two results are combined in integer
// operations only: a float bit-cast of the second result compiled to a copy of the first in
// this compiler's gfx950 code (r0 + r1 became 2 r0), while integer uses of both are kept.
__device__ __forceinline__ unsigned max_xor16(unsigned v)
{
    const auto r = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return r[0] > r[1] ? r[0] : r[1];
}
__device__ __forceinline__ unsigned max_xor32(unsigned v)
{
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return r[0] > r[1] ? r[0] : r[1];
}
// v + v(l ^ 16), the partner's bits recovered as r0 ^ r1 ^ v; both lanes of a pair add the
// same two values, in either order alike.
__device__ __forceinline__ float sum_xor16(float v)
{
    const unsigned u = __builtin_bit_cast(unsigned, v);
    const auto r = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    return v + __builtin_bit_cast(float, r[0] ^ r[1] ^ u);
}

// Largest of v over the four lanes that hold one line's operand values (l, l ^ 16, l ^ 32,
// l ^ 48).  v >= 0, so its bit pattern orders like its value.
__device__ __forceinline__ float line_max(float v)
{
    return __builtin_bit_cast(float, max_xor32(max_xor16(__builtin_bit_cast(unsigned, v))));
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

    // A[m][k] of the real block matrix, rows and columns interleaved (re, im) per point.
    __device__ __forceinline__ static float entry(int m, int k, const Bank<float>* __restrict__ bk)
    {
        const int o = m >> 1, j = k >> 1;
        if (o >= N || j >= N) return 0.f;
        const cx<float> w = bk->w[woff(N) + (o * j) % N];
        const float wr = w.r, wi = S < 0 ? w.i : -w.i;
        if ((m & 1) == 0) return (k & 1) == 0 ? wr : -wi;
        return (k & 1) == 0 ? wi : wr;
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

// 17-point lines by the conjugate-pair fold (KW = kFold17 in the kernel templates; 17^3
// only).  With a_j = x_j + x_{17-j}, b_j = x_j - x_{17-j} (j = 1..8), c = cos(2 pi jk/17),
// s = S sin(2 pi jk/17):
//     X_k = x_0 + sum_j (a_j c + i b_j s),  X_{17-k} = x_0 + sum_j (a_j c - i b_j s),  k = 1..8,
//     X_0 = x_0 + sum_j a_j.
// Both sums are one K = 32 product each, [a; b] (32 rows, re/im interleaved) by a 16 x 32
// matrix whose 16 rows are X_1..X_8 (P) or X_16..X_9 (M) less x_0, so a tile of 16 lines is
// 2 x 3 v_mfma_f32_16x16x32 instructions where the dense 34 x 34 form needs 3 x 3 x 3 of
// K = 16 (or 3 x 2 x 3 of K = 32 with K padded to 64).  Lane group g = l / 16 holds rows
// 8g..8g+7 of [a; b]: a_{1..4}, a_{5..8}, b_{1..4}, b_{5..8}.
constexpr int kFold17 = 1;

template <int S, int SPLIT>
struct Fold17Matrix {
    using Ops = MfmaOps<SPLIT, 32>;
    typename Ops::V ph, pl, mh, ml;   // P and M, high and low parts

    // Row m (output point m/2 + 1, re/im), column c (a or b of point (c % 16)/2 + 1, re/im).
    __device__ __forceinline__ static float entry(int m, int c, bool minus,
                                                  const Bank<float>* __restrict__ bk)
    {
        const int k = m / 2 + 1, j = (c % 16) / 2 + 1;
        const cx<float> w = bk->w[woff(17) + (j * k) % 17];   // exp(-2 pi i jk/17)
        const float cs = w.r, sn = (S < 0 ? w.i : -w.i) * (minus ? -1.f : 1.f);   // S sin
        const bool re_out = (m & 1) == 0, re_in = (c & 1) == 0;
        if (c < 16) return re_out == re_in ? cs : 0.f;          // a_j c
        if (re_out) return re_in ? 0.f : -sn;                    // Re(i b s) = -Im(b) s
        return re_in ? sn : 0.f;                                 // Im(i b s) =  Re(b) s
    }
    __device__ __forceinline__ void build(int lane, const Bank<float>* __restrict__ bk)
    {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int m = lane % 16, c = 8 * (lane / 16) + q;
            Ops::split(entry(m, c, false, bk), ph, pl, q);
            Ops::split(entry(m, c, true, bk), mh, ml, q);
        }
    }
};

// The distinct line matrices of one geometry, built once and used by every pass: z (N2),
// and y (N1) and x (N0) only where their lengths differ from those already held.
template <class C, int SPLIT, int KW>
struct MfmaMats {
    using MZ = MfmaMatrix<C::N2, C::S, SPLIT, KW>;
    using MY = MfmaMatrix<C::N1, C::S, SPLIT, KW>;
    using MX = MfmaMatrix<C::N0, C::S, SPLIT, KW>;
    static constexpr bool kOwnY = C::N1 != C::N2;
    static constexpr bool kOwnX = C::N0 != C::N2 && C::N0 != C::N1;
    MZ az;
    MY ay;
    MX ax;
    __device__ __forceinline__ void build(int lane, const Bank<float>* __restrict__ bk)
    {
        az.build(lane, bk);
        if constexpr (kOwnY) ay.build(lane, bk);
        if constexpr (kOwnX) ax.build(lane, bk);
    }
    __device__ __forceinline__ const MZ& z() const { return az; }
    __device__ __forceinline__ const MY& y() const
    {
        if constexpr (kOwnY) return ay;
        else return az;
    }
    __device__ __forceinline__ const MX& x() const
    {
        if constexpr (C::N0 == C::N2) return az;
        else if constexpr (C::N0 == C::N1) return ay;
        else return ax;
    }
};

template <class C, int SPLIT>
struct MfmaMats<C, SPLIT, kFold17> {
    static_assert(C::N0 == 17 && C::N1 == 17 && C::N2 == 17, "the fold form is for 17^3");
    using MF = Fold17Matrix<C::S, SPLIT>;
    MF a;
    __device__ __forceinline__ void build(int lane, const Bank<float>* __restrict__ bk)
    {
        a.build(lane, bk);
    }
    __device__ __forceinline__ const MF& z() const { return a; }
    __device__ __forceinline__ const MF& y() const { return a; }
    __device__ __forceinline__ const MF& x() const { return a; }
};

// One axis pass over nlines lines (line l starts at base(l), N points STRIDE apart, in LDS),
// tiles of 16 lines per wave; output point o of line l goes to store(base(l), l, o, value)
// (in place in LDS, or to HBM).
template <int N, int STRIDE, int S, int SPLIT, int NT, int KW_, typename BaseFn, typename StoreFn>
__device__ __forceinline__ void mfma_pass(const cx<float>* sm, int nlines, BaseFn base,
                                          const MfmaMatrix<N, S, SPLIT, KW_>& A, float scl,
                                          StoreFn store)
{
    static_assert(NT % 64 == 0, "whole waves: every lane of a tile takes part in the lane exchanges");
    using M = MfmaMatrix<N, S, SPLIT, KW_>;
    using Ops = typename M::Ops;
    constexpr int TM = M::TM, TK = M::TK, KW = M::KW, Q = M::Q, P = Q / 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4;
    const int ntiles = (nlines + 15) / 16;
    // Point offsets of this lane's operand values (padded positions repeat point N - 1).
    int goff[TK][P];
#pragma unroll
    for (int kt = 0; kt < TK; ++kt)
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int j = (KW * kt + Q * g) / 2 + p;
            goff[kt][p] = (j < N ? j : N - 1) * STRIDE;
        }
    for (int tile = wave; tile < ntiles; tile += NT / 64) {
        const int line = tile * 16 + (lane & 15);
        const bool valid = line < nlines;
        const int b = base(valid ? line : nlines - 1);
        float v[TK][Q];
        float mx = 0.f;
#pragma unroll
        for (int kt = 0; kt < TK; ++kt)
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const cx<float> x = sm[b + goff[kt][p]];
                v[kt][2 * p] = x.r;
                v[kt][2 * p + 1] = x.i;
                mx = fmaxf(mx, fmaxf(fabsf(x.r), fabsf(x.i)));
            }
        float up = 1.f, down = scl;
        if constexpr (SPLIT == 1) {
            mx = line_max(mx);
            int e = 0;
            if (mx > 0.f) (void)frexpf(mx, &e);
            up = ldexpf(1.f, -e);
            down = ldexpf(scl, e);
        }
        typename Ops::V bh[TK], bl[TK];
#pragma unroll
        for (int kt = 0; kt < TK; ++kt) {
            unsigned wh[P], wl[P];
#pragma unroll
            for (int p = 0; p < P; ++p)
                split_pair<SPLIT>(v[kt][2 * p] * up, v[kt][2 * p + 1] * up, wh[p], wl[p]);
            bh[kt] = pack_words<typename Ops::V>(wh);
            bl[kt] = pack_words<typename Ops::V>(wl);
        }
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
                for (int p = 0; p < 2; ++p) {
                    const int o = 8 * mt + 2 * g + p;
                    if (o < N) store(b, line, o, cx<float>{acc[2 * p] * down, acc[2 * p + 1] * down});
                }
            }
        }
    }
}

// The fold form of the pass above for 17-point lines (same interface).
template <int N, int STRIDE, int S, int SPLIT, int NT, typename BaseFn, typename StoreFn>
__device__ __forceinline__ void mfma_pass(const cx<float>* sm, int nlines, BaseFn base,
                                          const Fold17Matrix<S, SPLIT>& A, float scl,
                                          StoreFn store)
{
    static_assert(N == 17, "fold form");
    static_assert(NT % 64 == 0, "whole waves: every lane of a tile takes part in the lane exchanges");
    using Ops = MfmaOps<SPLIT, 32>;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4;
    const int ntiles = (nlines + 15) / 16;
    // This lane's points j = j0..j0+3 and their partners 17 - j; a (g < 2) or b (g >= 2).
    const int j0 = 4 * (g & 1) + 1;
    const float sg = g < 2 ? 1.f : -1.f;
    for (int tile = wave; tile < ntiles; tile += NT / 64) {
        const int line = tile * 16 + (lane & 15);
        const bool valid = line < nlines;
        const int b = base(valid ? line : nlines - 1);
        const cx<float> x0 = sm[b];
        float v[8];
        float mx = 0.f, sr = 0.f, si = 0.f;
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            const cx<float> u = sm[b + (j0 + p) * STRIDE];
            const cx<float> w = sm[b + (17 - j0 - p) * STRIDE];
            v[2 * p] = fmaf(sg, w.r, u.r);
            v[2 * p + 1] = fmaf(sg, w.i, u.i);
            sr += v[2 * p];
            si += v[2 * p + 1];
            mx = fmaxf(mx, fmaxf(fabsf(v[2 * p]), fabsf(v[2 * p + 1])));
        }
        // X_0: the a sums of groups 0 and 1 (groups 2 and 3 exchange b sums, unused).
        sr = sum_xor16(sr);
        si = sum_xor16(si);
        float up = 1.f, down = scl;
        if constexpr (SPLIT == 1) {
            mx = line_max(mx);
            int e = 0;
            if (mx > 0.f) (void)frexpf(mx, &e);
            up = ldexpf(1.f, -e);
            down = ldexpf(scl, e);
        }
        unsigned wh[4], wl[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) split_pair<SPLIT>(v[2 * p] * up, v[2 * p + 1] * up, wh[p], wl[p]);
        const typename Ops::V bh = pack_words<typename Ops::V>(wh);
        const typename Ops::V bl = pack_words<typename Ops::V>(wl);
        mf_v4f accp = {0.f, 0.f, 0.f, 0.f}, accm = {0.f, 0.f, 0.f, 0.f};
        accp = Ops::mma(A.ph, bh, accp);
        accm = Ops::mma(A.mh, bh, accm);
        accp = Ops::mma(A.ph, bl, accp);
        accm = Ops::mma(A.mh, bl, accm);
        accp = Ops::mma(A.pl, bh, accp);
        accm = Ops::mma(A.ml, bh, accm);
        if (valid) {
            const float xr = x0.r * scl, xi = x0.i * scl;
#pragma unroll
            for (int p = 0; p < 2; ++p) {
                const int k = 2 * g + 1 + p;
                store(b, line, k, cx<float>{fmaf(accp[2 * p], down, xr), fmaf(accp[2 * p + 1], down, xi)});
                store(b, line, 17 - k, cx<float>{fmaf(accm[2 * p], down, xr), fmaf(accm[2 * p + 1], down, xi)});
            }
            if (g == 0) store(b, line, 0, cx<float>{(x0.r + sr) * scl, (x0.i + si) * scl});
        }
    }
}

template <class C, int SPLIT>
struct MfmaCfg : C {
    static constexpr int SPLITV = SPLIT;
};

// The three axis passes of block `bid`, staged in LDS: z and y in place; x with XD = 1 from
// LDS straight to `out` (for a fixed output point the 16 lines of a tile are consecutive in
// (y, z), so each store instruction writes four runs of 16 elements), with XD = 0 in place,
// then one coalesced copy of the block to `out`.  CONJ: conjugate on store (the backward
// transform through forward matrices, for the mixed launch).  Ends with a barrier, so a
// caller may reuse `sm` at once.
template <class C, int SPLIT, int KW, int XD>
__device__ __forceinline__ void mfma_passes(long bid, long nvol, cx<float>* __restrict__ out,
                                            float scl, bool conj,
                                            const MfmaMats<C, SPLIT, KW>& A, cx<float>* sm)
{
    const int nv = Fused<C>::nvols(bid, nvol);
    mfma_pass<C::N2, 1, C::S, SPLIT, C::NT>(
        sm, nv * C::N0 * C::N1,
        [](int l) {
            const unsigned u = (unsigned)l, v = u / (C::N0 * C::N1), r = u - v * (C::N0 * C::N1);
            return (int)(v * C::VS + (r / C::N1) * C::XS + (r % C::N1) * C::YS);
        },
        A.z(), 1.f, [sm](int b, int, int o, cx<float> y) { sm[b + o] = y; });
    __syncthreads();
    mfma_pass<C::N1, C::YS, C::S, SPLIT, C::NT>(
        sm, nv * C::N0 * C::N2,
        [](int l) {
            const unsigned u = (unsigned)l, v = u / (C::N0 * C::N2), r = u - v * (C::N0 * C::N2);
            return (int)(v * C::VS + (r / C::N2) * C::XS + (r % C::N2));
        },
        A.y(), 1.f, [sm](int b, int, int o, cx<float> y) { sm[b + o * C::YS] = y; });
    __syncthreads();
    cx<float>* __restrict__ dst = out + bid * C::VPW * (long)C::V;
    const float ci = conj ? -1.f : 1.f;
    auto xbase = [](int l) {
        const unsigned u = (unsigned)l, v = u / C::PLANE, q = u - v * C::PLANE;
        return (int)(v * C::VS + (q / C::N2) * C::YS + (q % C::N2));
    };
    if constexpr (XD == 1) {
        mfma_pass<C::N0, C::XS, C::S, SPLIT, C::NT>(
            sm, nv * C::PLANE, xbase, A.x(), scl, [dst, ci](int, int l, int o, cx<float> y) {
                const unsigned u = (unsigned)l, v = u / C::PLANE;
                st_stream(dst + (long)(u + v * (C::V - C::PLANE)) + (long)o * C::PLANE,
                          cx<float>{y.r, ci * y.i});
            });
    } else {
        mfma_pass<C::N0, C::XS, C::S, SPLIT, C::NT>(
            sm, nv * C::PLANE, xbase, A.x(), scl,
            [sm](int b, int, int o, cx<float> y) { sm[b + o * C::XS] = y; });
        __syncthreads();
        for (int e = threadIdx.x; e < nv * C::V; e += C::NT) {
            const cx<float> y = sm[Fused<C>::slot(e)];
            st_stream(dst + e, cx<float>{y.r, ci * y.i});
        }
    }
    __syncthreads();
}

// One block of VPW volumes: stage (conjugated on load when CONJ), then the passes.
template <class C, int SPLIT, int KW, int XD>
__device__ __forceinline__ void mfma_block(long bid, long nvol, const cx<float>* __restrict__ in,
                                           cx<float>* __restrict__ out, float scl, bool conj,
                                           const MfmaMats<C, SPLIT, KW>& A, cx<float>* sm)
{
    using K = Fused<C>;
    const int tid = threadIdx.x;
    const int nv = K::nvols(bid, nvol);
    K::stage(tid, bid, nvol, in, sm);
    __syncthreads();
    if (conj) {
        for (int e = tid; e < nv * C::V; e += C::NT) sm[K::slot(e)].i = -sm[K::slot(e)].i;
        __syncthreads();
    }
    mfma_passes<C, SPLIT, KW, XD>(bid, nvol, out, scl, conj, A, sm);
}

// One block per workgroup (grid = nblocks).
template <class C, int SPLIT, int KW, int XD>
__global__ void __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::WPE, 8)))
    k_mfma(const cx<float>* __restrict__ in, cx<float>* __restrict__ out, long nvol,
           const Bank<float>* __restrict__ bk, float scl)
{
    static_assert(sizeof(typename C::T) == 4, "the matrix-core variant is complex64");
    __shared__ cx<float> sm[C::LDS];
    MfmaMats<C, SPLIT, KW> A;
    A.build(threadIdx.x & 63, bk);
    mfma_block<C, SPLIT, KW, XD>(blockIdx.x, nvol, in, out, scl, false, A, sm);
}

// Persistent and pipelined: at most P workgroups per CU, each walking blocks
// blockIdx.x + k gridDim.x (every block costs the same, so no claims) with its matrices
// built once.  The next block's elements are loaded into registers (E per thread, the
// coalesced order of the staging copy) while this block's passes run, and written to LDS
// when they end.
template <class C, int P>
struct MfmaPipe {
    static constexpr int E = (C::VPW * C::V + C::NT - 1) / C::NT;
    static constexpr int WPE = waves_per_eu(C::LDS * (int)sizeof(cx<float>), C::NT, P);
};

template <class C, int SPLIT, int KW, int P, int XD>
__global__ void __launch_bounds__(C::NT)
    __attribute__((amdgpu_waves_per_eu(MfmaPipe<C, P>::WPE, 8)))
    k_mfma_pipe(const cx<float>* __restrict__ in, cx<float>* __restrict__ out, long nvol,
                long nblocks, const Bank<float>* __restrict__ bk, float scl)
{
    static_assert(sizeof(typename C::T) == 4, "the matrix-core variant is complex64");
    using K = Fused<C>;
    constexpr int E = MfmaPipe<C, P>::E;
    __shared__ cx<float> sm[C::LDS];
    const int tid = threadIdx.x;
    MfmaMats<C, SPLIT, KW> A;
    A.build(tid & 63, bk);
    cx<float> pf[E];
    // Loads past a partial last block repeat its last element (no branch per element).
    auto load = [&](long b) {
        const int n = K::nvols(b, nvol) * C::V;
        const cx<float>* __restrict__ src = in + b * C::VPW * (long)C::V;
#pragma unroll
        for (int k = 0; k < E; ++k) {
            const int e = tid + k * C::NT;
            pf[k] = ld_hint<1>(src + (e < n ? e : n - 1));
        }
    };
    long bid = blockIdx.x;
    load(bid);
    for (;;) {
        {
            const int n = K::nvols(bid, nvol) * C::V;
#pragma unroll
            for (int k = 0; k < E; ++k)
                if (tid + k * C::NT < n) sm[K::slot(tid + k * C::NT)] = pf[k];
        }
        __syncthreads();
        const long nxt = bid + gridDim.x;
        if (nxt < nblocks) load(nxt);
        mfma_passes<C, SPLIT, KW, XD>(bid, nvol, out, scl, false, A, sm);
        if (nxt >= nblocks) break;
        bid = nxt;
    }
}

}  // namespace fx

#include <fx_mfma_launch.hpp>
