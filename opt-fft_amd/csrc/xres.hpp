// Single-pass kernel for volumes larger than a CU holds (36^3 complex128, 746 KB; 24x24x48
// complex128, 442 KB; also 24^3 complex128 and the complex64 36^3 and 24x24x48): the x DFT
// split by output residue.  With N0 = R * M, output x-plane R q + h of
// X[k] = sum_j x[j] W^(j k) is
//
//   X[R q + h] = sum_{n < M} W_M^(n q) * ( W_N0^(n h) * sum_{m < R} x[n + M m] W_R^(m h) ),
//
// so a workgroup (v, h) reads all N0 x-planes of volume v, forms the M sums of residue h in
// registers (a thread per (y, z) point, every load instruction of a wave 64 consecutive
// elements), and an M-point DFT over n gives the x transform of its M output planes.  Those
// M planes (N1 x N2, rows padded by PZ) are held in LDS, z lines and then y lines are
// transformed there, and the y pass writes planes R q + h to HBM.  Every element is written
// once; it is read R times, by the R workgroups of its volume, which are consecutive
// workgroups of one XCD (an 8-XCD part dispatches workgroup g to XCD g mod 8), so they run
// together and most reads of a line after the first are served by that XCD's L2.  The
// inputs are therefore read with plain loads (HL = 0), the outputs stored nontemporal
// (HS = 1).  What bounds the kernel is the rate at which one CU reads from L2.
#pragma once

#include "kernels.hpp"

namespace fx {

// OR: order of the x pass's points.  0: every workgroup of a volume walks the plane from
// point 0, so the R of them request the same lines together; 1: residue h starts at its
// thread block h mod PPT (the blocks of NT points taken in rotated order); 2: residue h
// starts at point h * PLANE / R (the plane walked with a wrap).  1 and 2 spread the R
// workgroups' requests over different lines (and L2 channels) at the price of keeping each
// line in L2 longer.
template <int N0_, int N1_, int N2_, typename T_, int S_, int R_, int NT_, int OR_ = 0,
          int PZ_ = 1, int HL_ = 0, int HS_ = 1, int ALG_ = 0>
struct XResCfg {
    using T = T_;
    static constexpr int N0 = N0_, N1 = N1_, N2 = N2_, S = S_, R = R_, NT = NT_, PZ = PZ_;
    static constexpr int OR = OR_;
    static constexpr int PPT = (N1 * N2 + NT - 1) / NT;   // (y, z) points per thread
    static constexpr int HL = HL_, HS = HS_, ALG = ALG_, V17 = ALG_, VPW = 1, P = 0;
    static_assert(N0 % R == 0, "the residues split the x-planes evenly");
    static constexpr int M = N0 / R;       // x-planes per workgroup
    static constexpr int V = N0 * N1 * N2;
    static constexpr int PLANE = N1 * N2;
    static constexpr int YS = N2 + PZ;     // LDS pitch between z rows
    static constexpr int SS = N1 * YS;     // LDS pitch between x-planes
    static constexpr int LDS = M * SS;     // elements
    static constexpr int WPE = waves_per_eu(LDS * (int)sizeof(cx<T>), NT, 0);
    static constexpr int kXcds = 8;
};

template <class C>
struct XRes {
    using T = typename C::T;

    // Workgroup g -> (volume v, residue h): g = kXcds * i + x, volume kXcds * (i / R) + x,
    // residue i mod R.
    static FX_HD void task(long g, long& v, int& h)
    {
        const long x = g % C::kXcds, i = g / C::kXcds;
        const long vq = i / C::R;
        h = (int)(i - vq * C::R);
        v = vq * C::kXcds + x;
    }
    static FX_HD long grid(long nvol) { return (nvol + C::kXcds - 1) / C::kXcds * C::kXcds * C::R; }

    // W_N0^e with the transform's sign.
    static FX_HD cx<T> w0(int e, const Bank<T>* __restrict__ bk)
    {
        cx<T> w = bk->w[woff(C::N0) + e];
        if constexpr (C::S > 0) w.i = -w.i;
        return w;
    }

    // Phase 1: residue h of the x DFT of volume v, HBM -> VGPRs -> LDS planes 0 .. M-1.
    static FX_HD void xpass(int tid, long v, int h, const cx<T>* __restrict__ in, cx<T>* sm,
                            const Bank<T>* __restrict__ bk)
    {
        const cx<T>* __restrict__ src = in + v * (long)C::V;
        cx<T> wr[C::R], tw[C::M];
#pragma unroll
        for (int m = 1; m < C::R; ++m) wr[m] = w0(C::M * ((m * h) % C::R), bk);
#pragma unroll
        for (int n = 1; n < C::M; ++n) tw[n] = w0(n * h, bk);
        for (int kk = 0; kk < C::PPT; ++kk) {
            int p;
            if constexpr (C::OR == 1) {
                p = tid + ((kk + h) % C::PPT) * C::NT;
                if (p >= C::PLANE) continue;
            } else {
                p = tid + kk * C::NT;
                if (p >= C::PLANE) continue;
                if constexpr (C::OR == 2) {
                    p += h * (C::PLANE / C::R);
                    p = p < C::PLANE ? p : p - C::PLANE;
                }
            }
            cx<T> a[C::M];
#pragma unroll
            for (int n = 0; n < C::M; ++n) a[n] = ld_hint<C::HL>(src + n * C::PLANE + p);
#pragma unroll
            for (int m = 1; m < C::R; ++m) {
#pragma unroll
                for (int n = 0; n < C::M; ++n) {
                    const cx<T> x = ld_hint<C::HL>(src + (n + C::M * m) * C::PLANE + p);
                    a[n] = a[n] + cmul(x, wr[m]);
                }
            }
#pragma unroll
            for (int n = 1; n < C::M; ++n) a[n] = cmul(a[n], tw[n]);
            dft<C::M, C::S, T, C::ALG>(a, bk);
            const int y = p / C::N2;
            const int z = p - y * C::N2;
#pragma unroll
            for (int q = 0; q < C::M; ++q) sm[q * C::SS + y * C::YS + z] = a[q];
        }
    }
    // Phase 2: z lines (plane s, y).
    static FX_HD void zpass(int tid, cx<T>* sm, const Bank<T>* __restrict__ bk)
    {
        lds_pass<C::N2, 1, C::S, C::ALG, C::NT>(
            tid, sm, C::M * C::N1,
            [](int l) {
                const int s = l / C::N1;
                const int y = l - s * C::N1;
                return s * C::SS + y * C::YS;
            },
            bk);
    }
    // Phase 3: y lines (plane s, z), LDS -> output x-plane R s + h (coalesced across z).
    static FX_HD void ypass(int tid, long v, int h, const cx<T>* sm, cx<T>* __restrict__ out,
                            const Bank<T>* __restrict__ bk, T scl)
    {
        cx<T>* __restrict__ dst = out + v * (long)C::V;
        for (int l = tid; l < C::M * C::N2; l += C::NT) {
            const int s = l / C::N2;
            const int z = l - s * C::N2;
            cx<T> t[C::N1];
#pragma unroll
            for (int j = 0; j < C::N1; ++j) t[j] = sm[s * C::SS + j * C::YS + z];
            dft<C::N1, C::S, T, C::ALG>(t, bk);
            cx<T>* __restrict__ o = dst + (long)(C::R * s + h) * C::PLANE + z;
            if (scl != T(1)) {
#pragma unroll
                for (int k = 0; k < C::N1; ++k) st_hint<C::HS>(o + k * C::N2, scale(t[k], scl));
            } else {
#pragma unroll
                for (int k = 0; k < C::N1; ++k) st_hint<C::HS>(o + k * C::N2, t[k]);
            }
        }
    }
};

template <class C>
__global__ void __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::WPE, 8)))
    k_xres(const cx<typename C::T>* __restrict__ in, cx<typename C::T>* __restrict__ out,
           long nvol, const Bank<typename C::T>* __restrict__ bk, typename C::T scl)
{
    using K = XRes<C>;
    __shared__ cx<typename C::T> sm[C::LDS];
    long v;
    int h;
    K::task(blockIdx.x, v, h);
    if (v >= nvol) return;   // uniform over the workgroup, before any barrier
    const int tid = threadIdx.x;
    K::xpass(tid, v, h, in, sm, bk);
    __syncthreads();
    K::zpass(tid, sm, bk);
    __syncthreads();
    K::ypass(tid, v, h, sm, out, bk, scl);
}

}  // namespace fx

#include <fx_xres_launch.hpp>
