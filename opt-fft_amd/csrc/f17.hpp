// Balanced persistent kernel for 17^3 volumes (complex128 and complex64).
//
// A 17^3 axis pass is 289 lines of 17 points.  With one line per lane that is 4.5 waves, and
// the hardware puts a workgroup's fifth wave on the same SIMD as its first, so one SIMD of
// four carries twice the others' work in every pass.  Here a workgroup is exactly four waves
// (one per SIMD) and each pass is split so the four waves carry equal work:
//   * lane t (0..255) transforms line t whole (folded dense 17, outputs stored as computed);
//   * lines 256..287: wave w takes lines 256 + 8w .. 256 + 8w + 7, eight lanes per line, each
//     lane computing one output pair (X_k, X_17-k), k = 1 + lane % 8, from the folded sums
//     u_j = x_j + x_17-j, v_j = x_j - x_17-j (the lane with k = 1 also stores X_0);
//   * line 288 is done the same way by eight lanes of wave (pass % 4), so over the three
//     passes of a volume the extra round falls on a different wave each time.
// The eight lanes of a split line belong to one wave: all read the line before any writes
// it (one wave's LDS operations execute in program order).
//
// Data path per volume: one coalesced linear read HBM -> VGPRs (the next volume's read is
// issued right after this volume is staged, so it is in flight through all three passes),
// stage VGPRs -> LDS, z pass and y pass LDS -> LDS, x pass LDS -> HBM with the output scale
// fused (lanes of a wave hold consecutive (y, z): every store instruction writes contiguous
// runs).  Workgroups claim volumes from the launch's counter (claim_block).
#pragma once
#include "kernels.hpp"

namespace fx {

template <typename T_, int S_, int P_>
struct F17Cfg {
    using T = T_;
    static constexpr int N0 = 17, N1 = 17, N2 = 17, S = S_, VPW = 1, NT = 256, PERSIST = P_;
    static constexpr int N = 17, PL = 289, V = 4913;
    static constexpr int U = (V + NT - 1) / NT;  // staged elements per thread (20)
    static constexpr int LDS = V;                // cx<T> elements of the volume
    static constexpr int PCT = 8 * 16;           // pair-constant table: [k-1][c_1..c_8, s_1..s_8]
    static constexpr int LDS_BYTES = V * (int)sizeof(cx<T>) + PCT * (int)sizeof(T) + 16;
    static constexpr int WPE = waves_per_eu(LDS_BYTES, NT, P_);
};

template <class C>
struct F17 {
    using T = typename C::T;
    static constexpr int N = C::N, PL = C::PL, V = C::V, NT = C::NT, U = C::U;

    struct Regs {
        cx<T> r[U];
    };
    // Result of one lane's share of a split line, held between the read and the write.
    struct Pair {
        cx<T> y0, yk, ykm;
        int line, k;
        bool on;
    };

    // pct[(k-1)*16 + (j-1)] = cos(2 pi jk/17), pct[(k-1)*16 + 8 + (j-1)] = sin(2 pi jk/17).
    static FX_HD void fill_pct(int t, const Bank<T>* __restrict__ bk, T* pct)
    {
        if (t >= C::PCT) return;
        const int k = t / 16 + 1, i = t % 16, j = i % 8 + 1;
        const int m = (j * k) % 17;
        pct[t] = i < 8 ? bk->cos17[m] : bk->sin17[m];
    }

    // Unconditional clamped loads: every lane issues U loads back to back.
    static FX_HD void load(int t, long v, const cx<T>* __restrict__ in, Regs& R)
    {
        const cx<T>* __restrict__ src = in + v * (long)V;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            int e = t + u * NT;
            e = e < V ? e : V - 1;
            R.r[u] = ld_stream(src + e);
        }
    }

    // Component-wise stores: a predicated whole-struct copy becomes a memcpy that keeps the
    // compiler from holding R in registers.
    static FX_HD void stage(int t, const Regs& R, cx<T>* sm)
    {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = t + u * NT;
            if (e < V) {
                sm[e].r = R.r[u].r;
                sm[e].i = R.r[u].i;
            }
        }
    }

    // First element and stride of line `l` of pass D (0: z-lines, 1: y-lines, 2: x-lines).
    template <int D>
    static FX_HD int base(int l)
    {
        if constexpr (D == 0) return 17 * l;
        else if constexpr (D == 1) return PL * (l / 17) + l % 17;
        else return l;
    }
    template <int D>
    static constexpr int stride()
    {
        return D == 0 ? 1 : (D == 1 ? 17 : PL);
    }

    // Whole line t.  D < 2: in place in LDS; D == 2: to dst (this volume in HBM), scaled.
    template <int D>
    static FX_HD void main_task(int t, cx<T>* sm, const Bank<T>* __restrict__ bk,
                                cx<T>* __restrict__ dst, T scl)
    {
        const int b = base<D>(t);
        constexpr int st = stride<D>();
        cx<T> x[17];
#pragma unroll
        for (int k = 0; k < 17; ++k) x[k] = sm[b + k * st];
        if constexpr (D < 2) {
            fold17_put<C::S>(x, bk, [&](int k, const cx<T>& y) { sm[b + k * st] = y; });
        } else {
            fold17_put<C::S>(x, bk, [&](int k, const cx<T>& y) {
                st_stream(dst + b + k * st, cx<T>{y.r * scl, y.i * scl});
            });
        }
    }

    // Split line of round `round` (0: lines 256..287 in every wave; 1: line 288 in wave
    // D % 4): read the line and compute this lane's outputs.
    template <int D>
    static FX_HD void pair_read(int t, int round, const cx<T>* sm, const T* pct, Pair& p)
    {
        const int w = t / 64, l = t % 64;
        if (round == 0) {
            p.line = 256 + 8 * w + l / 8;
            p.on = true;
        } else {
            p.line = 288;
            p.on = (w == D % 4) && l < 8;
        }
        p.k = l % 8 + 1;
        if (!p.on) return;
        const int b = base<D>(p.line);
        constexpr int st = stride<D>();
        const T* __restrict__ c = pct + (p.k - 1) * 16;
        const cx<T> x0 = sm[b];
        cx<T> A = x0, B = {T(0), T(0)}, X0 = x0;
#pragma unroll
        for (int j = 1; j <= 8; ++j) {
            const cx<T> a = sm[b + j * st], z = sm[b + (17 - j) * st];
            const cx<T> u = a + z, v = a - z;
            X0 = X0 + u;
            A.r += c[j - 1] * u.r;
            A.i += c[j - 1] * u.i;
            B.r += c[8 + j - 1] * v.r;
            B.i += c[8 + j - 1] * v.i;
        }
        const cx<T> sb = mul_si<C::S>(B);
        p.y0 = X0;
        p.yk = A + sb;
        p.ykm = A - sb;
    }

    template <int D>
    static FX_HD void pair_write(const Pair& p, cx<T>* sm, cx<T>* __restrict__ dst, T scl)
    {
        if (!p.on) return;
        const int b = base<D>(p.line);
        constexpr int st = stride<D>();
        if constexpr (D < 2) {
            sm[b + p.k * st] = p.yk;
            sm[b + (17 - p.k) * st] = p.ykm;
            if (p.k == 1) sm[b] = p.y0;
        } else {
            st_stream(dst + b + p.k * st, cx<T>{p.yk.r * scl, p.yk.i * scl});
            st_stream(dst + b + (17 - p.k) * st, cx<T>{p.ykm.r * scl, p.ykm.i * scl});
            if (p.k == 1) st_stream(dst + b, cx<T>{p.y0.r * scl, p.y0.i * scl});
        }
    }
};

template <class C, int D>
__device__ __forceinline__ void f17_pass(int t, cx<typename C::T>* sm, const typename C::T* pct,
                                         const Bank<typename C::T>* __restrict__ bk,
                                         cx<typename C::T>* __restrict__ dst, typename C::T scl)
{
    using K = F17<C>;
    K::template main_task<D>(t, sm, bk, dst, scl);
    typename K::Pair p;
    K::template pair_read<D>(t, 0, sm, pct, p);
    K::template pair_write<D>(p, sm, dst, scl);
    if ((t >> 6) == D % 4) {
        K::template pair_read<D>(t, 1, sm, pct, p);
        K::template pair_write<D>(p, sm, dst, scl);
    }
}

template <class C>
__global__ void __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::WPE, 8)))
    k_f17(const cx<typename C::T>* __restrict__ in, cx<typename C::T>* __restrict__ out,
          long nvol, const Bank<typename C::T>* __restrict__ bk, typename C::T scl,
          unsigned* __restrict__ work)
{
    using T = typename C::T;
    using K = F17<C>;
    __shared__ cx<T> sm[C::LDS];
    __shared__ T pct[C::PCT];
    __shared__ long s_next;
    typename K::Regs R;
    const int t = threadIdx.x;
    long b = blockIdx.x;
    if (b >= nvol) return;
    K::fill_pct(t, bk, pct);
    K::load(t, b, in, R);
    for (;;) {
        long claim = 0;
        if (t == 0) claim = claim_block(work, nvol, b);  // in flight while staging
        K::stage(t, R, sm);
        if (t == 0) s_next = claim;
        __syncthreads();
        const long nb = s_next;
        if (nb < nvol) K::load(t, nb, in, R);
        f17_pass<C, 0>(t, sm, pct, bk, nullptr, scl);
        __syncthreads();
        f17_pass<C, 1>(t, sm, pct, bk, nullptr, scl);
        __syncthreads();
        f17_pass<C, 2>(t, sm, pct, bk, out + b * (long)C::V, scl);
        if (nb >= nvol) break;
        __syncthreads();
        b = nb;
    }
}

}  // namespace fx

#include <fx_f17_launch.hpp>
