// The two 3D kernel shapes, as templates over geometry, precision, direction and tiling.
//
// Fused (volume fits in LDS; L = 6, 8, 12, 17, 8x8x16):
//   one workgroup owns VPW consecutive volumes.  One coalesced HBM read stages them into
//   LDS (z rows padded by PZ elements, x planes by PY, to keep the strided line reads off a
//   single bank), the z and y line passes run LDS -> VGPR -> LDS, and the x pass reads LDS
//   and writes HBM directly: lanes of a wave take consecutive (y,z), so every store of the
//   x pass is one contiguous, coalesced run.  One HBM read and one HBM write per element.
//
// Pencil (volume exceeds LDS; L = 24, 36, 24x24x48 at complex128):
//   pass A: one workgroup owns SPW consecutive x-slabs (contiguous n1*n2 planes), stages
//   them in LDS, runs z then y; the y pass writes the slab straight to `out`, coalesced.
//   pass B: one thread per x-line, in place on `out`: every load and store instruction of a
//   wave covers consecutive (y,z), i.e. contiguous addresses, so no LDS is needed at all.
//   Pass B re-reads what pass A wrote; chunking keeps that intermediate cache resident.
//
// Persistent pipelined forms (k_fpipe for fused geometries, k_spipe for pencil pass A): one
// or a few workgroups per CU claim blocks from a per-launch counter (claim_block); a
// block's first line pass reads its lines straight from HBM into VGPRs, and the next
// block's lines are issued as soon as those registers are free, so they load while the
// current block is in its remaining passes.  PipeCfg / PencilCfg ST = 1 additionally
// spreads each workgroup's HBM requests over the iteration (next lines in two halves,
// stagger_start); PipeCfg ST = 2 issues them in three parts, the last after the y pass, so
// fewer prefetched registers are live during it.
//
// HBM accesses that touch an element once use the nontemporal forms of cx.hpp (ld_hint /
// st_hint, HN in the configs): inputs, outputs written in whole lines, and the pencil
// intermediate as pass B reads it.
//
// Each kernel body is a sequence of phase functions separated by workgroup barriers.  A
// phase only reads LDS locations written in an earlier phase or by the same thread, so the
// phases are plain __host__ __device__ functions of (thread, block).  Launchers come from
// <fx_launch.hpp>, found on the include path.
#pragma once
#include "claim_layout.hpp"
#include "codelets.hpp"
#include "registry.hpp"

namespace fx {

template <typename T>
constexpr int prec_of() { return sizeof(T) == 8 ? kC128 : kC64; }

// Copy TOT contiguous elements HBM -> LDS, slot(e) giving the padded LDS index.  Loads
// are issued in groups of up to 16 per thread before any LDS write, so each wave keeps
// 16 x 1 KiB (complex128) of reads in flight.
template <int TOT, int NT, int HL = 1, typename T, typename SlotFn>
FX_HD void stage_full(int tid, const cx<T>* __restrict__ src, cx<T>* sm, SlotFn slot)
{
    constexpr int ITER = (TOT + NT - 1) / NT;
    constexpr int GROUP = ITER < 16 ? ITER : 16;
#pragma unroll
    for (int g = 0; g < ITER; g += GROUP) {
        cx<T> buf[GROUP];
        // Loads are unconditional (a tail lane re-reads the last element) so the compiler
        // keeps buf in VGPRs and issues the whole group before the first wait.
#pragma unroll
        for (int k = 0; k < GROUP; ++k) {
            const int e = tid + (g + k) * NT;
            if (g + k < ITER) buf[k] = ld_hint<HL>(src + ((TOT % NT == 0 || e < TOT) ? e : TOT - 1));
        }
#pragma unroll
        for (int k = 0; k < GROUP; ++k) {
            const int e = tid + (g + k) * NT;
            if (g + k < ITER && (TOT % NT == 0 || e < TOT)) sm[slot(e)] = buf[k];
        }
    }
}

// In-LDS line pass: line l starts at base(l), its N points are STRIDE apart.
template <int N, int STRIDE, int S, int V17, int NT, typename T, typename BaseFn>
FX_HD void lds_pass(int tid, cx<T>* sm, int nlines, BaseFn base, const Bank<T>* __restrict__ bk)
{
    for (int l = tid; l < nlines; l += NT) {
        const int b = base(l);
        cx<T> t[N];
#pragma unroll
        for (int j = 0; j < N; ++j) t[j] = sm[b + j * STRIDE];
        dft_put<N, S, T, V17>(t, bk, [&](int k, cx<T> v) { sm[b + k * STRIDE] = v; });
    }
}

// Waves per SIMD a kernel must sustain so that its LDS-limited number of workgroups fits
// on a CU (160 KB of LDS, 4 SIMDs, at most 8 waves each): the VGPR budget follows from it.
constexpr int waves_per_eu(int lds_bytes, int nt, int cap_wgs)
{
    const int waves = (nt + 63) / 64;
    int wgs = lds_bytes > 0 ? 163840 / lds_bytes : 32;
    if (wgs > 32 / waves) wgs = 32 / waves;
    if (cap_wgs > 0 && wgs > cap_wgs) wgs = cap_wgs;
    if (wgs < 1) wgs = 1;
    const int w = (wgs * waves + 3) / 4;
    return w < 1 ? 1 : (w > 8 ? 8 : w);
}

// ------------------------------------------------------------------------------- fused
template <int N0_, int N1_, int N2_, typename T_, int S_, int VPW_, int NT_, int PZ_, int PY_,
          int V17_>
struct FusedCfg {
    using T = T_;
    static constexpr int N0 = N0_, N1 = N1_, N2 = N2_, S = S_, VPW = VPW_, NT = NT_;
    static constexpr int PZ = PZ_, PY = PY_, V17 = V17_;
    static constexpr int V = N0 * N1 * N2;
    static constexpr int PLANE = N1 * N2;
    static constexpr int YS = N2 + PZ;           // LDS pitch between z rows
    static constexpr int XS = N1 * YS + PY;      // LDS pitch between x planes
    static constexpr int VS = N0 * XS;           // LDS pitch between volumes
    static constexpr int LDS = VPW * VS;         // elements
    static constexpr int WPE = waves_per_eu(LDS * (int)sizeof(cx<T>), NT, 0);
    static constexpr int HN = 3;                 // streaming hints: 1 loads, 2 stores
};

template <class C>
struct Fused {
    using T = typename C::T;

    // Volumes held by block `bid` (VPW except for the last block).
    static FX_HD int nvols(long bid, long nvol)
    {
        const long left = nvol - bid * C::VPW;
        return (int)(left < C::VPW ? left : C::VPW);
    }
    static FX_HD int slot(int e)
    {
        const int v = e / C::V;
        const int r = e - v * C::V;
        const int x = r / C::PLANE;
        const int q = r - x * C::PLANE;
        const int y = q / C::N2;
        const int z = q - y * C::N2;
        return v * C::VS + x * C::XS + y * C::YS + z;
    }
    // Phase 1: one coalesced read of the block's contiguous volumes into LDS.
    static FX_HD void stage(int tid, long bid, long nvol, const cx<T>* __restrict__ in, cx<T>* sm)
    {
        const int nv = nvols(bid, nvol);
        const cx<T>* __restrict__ src = in + bid * C::VPW * (long)C::V;
        if (nv == C::VPW)
            stage_full<C::VPW * C::V, C::NT, C::HN & 1>(tid, src, sm, [](int e) { return slot(e); });
        else
            for (int e = tid; e < nv * C::V; e += C::NT) sm[slot(e)] = ld_hint<C::HN & 1>(src + e);
    }
    // Phase 2: z lines (v, x, y), contiguous in LDS.
    static FX_HD void zpass(int tid, long bid, long nvol, cx<T>* sm, const Bank<T>* __restrict__ bk)
    {
        lds_pass<C::N2, 1, C::S, C::V17, C::NT>(
            tid, sm, nvols(bid, nvol) * C::N0 * C::N1,
            [](int l) {
                const int v = l / (C::N0 * C::N1);
                const int r = l - v * (C::N0 * C::N1);
                const int x = r / C::N1;
                const int y = r - x * C::N1;
                return v * C::VS + x * C::XS + y * C::YS;
            },
            bk);
    }
    // Phase 3: y lines (v, x, z), stride YS.
    static FX_HD void ypass(int tid, long bid, long nvol, cx<T>* sm, const Bank<T>* __restrict__ bk)
    {
        lds_pass<C::N1, C::YS, C::S, C::V17, C::NT>(
            tid, sm, nvols(bid, nvol) * C::N0 * C::N2,
            [](int l) {
                const int v = l / (C::N0 * C::N2);
                const int r = l - v * (C::N0 * C::N2);
                const int x = r / C::N2;
                const int z = r - x * C::N2;
                return v * C::VS + x * C::XS + z;
            },
            bk);
    }
    // Phase 4: x lines (v, y, z), LDS -> HBM; a wave's stores are contiguous in (y, z).
    static FX_HD void xpass(int tid, long bid, long nvol, const cx<T>* sm, cx<T>* __restrict__ out,
                            const Bank<T>* __restrict__ bk, T scl)
    {
        const int nv = nvols(bid, nvol);
        cx<T>* __restrict__ dst = out + bid * C::VPW * (long)C::V;
        const bool scaled = scl != T(1);
        for (int l = tid; l < nv * C::PLANE; l += C::NT) {
            const int v = l / C::PLANE;
            const int q = l - v * C::PLANE;
            const int y = q / C::N2;
            const int z = q - y * C::N2;
            const int b = v * C::VS + y * C::YS + z;
            cx<T> t[C::N0];
#pragma unroll
            for (int j = 0; j < C::N0; ++j) t[j] = sm[b + j * C::XS];
            dft<C::N0, C::S, T, C::V17>(t, bk);
            cx<T>* __restrict__ o = dst + (long)v * C::V + q;
            if (scaled) {
#pragma unroll
                for (int k = 0; k < C::N0; ++k)
                    st_hint<C::HN & 2>(o + k * C::PLANE, scale(t[k], scl));
            } else {
#pragma unroll
                for (int k = 0; k < C::N0; ++k) st_hint<C::HN & 2>(o + k * C::PLANE, t[k]);
            }
        }
    }
};

template <class C>
__global__ void __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::WPE, 8)))
    k_fused(const cx<typename C::T>* __restrict__ in, cx<typename C::T>* __restrict__ out,
            long nvol, const Bank<typename C::T>* __restrict__ bk, typename C::T scl)
{
    using K = Fused<C>;
    __shared__ cx<typename C::T> sm[C::LDS];
    const int tid = threadIdx.x;
    const long bid = blockIdx.x;
    K::stage(tid, bid, nvol, in, sm);
    __syncthreads();
    K::zpass(tid, bid, nvol, sm, bk);
    __syncthreads();
    K::ypass(tid, bid, nvol, sm, bk);
    __syncthreads();
    K::xpass(tid, bid, nvol, sm, out, bk, scl);
}

// ---------------------------------------------------------------------- fused, pipelined
// Persistent, software-pipelined form of the fused kernel.  A workgroup walks blocks
// b, b + grid, ...; for each block:
//   x pass: the x-lines (v, y, z) come straight from HBM into VGPRs -- lanes hold
//           consecutive (y, z), so each load instruction reads one contiguous run -- are
//           transformed and written to LDS;
//   prefetch: the next block's x-lines are issued into the same VGPRs at once, and stay in
//           flight through the barriers (which wait on LDS only) and the two passes below;
//   z pass: LDS -> LDS;
//   y pass: LDS -> HBM; lanes hold consecutive z, so the stores are contiguous runs.
// Four LDS operations per element instead of six, and HBM reads overlap the arithmetic.
template <class C>
struct FusedPipe {
    using T = typename C::T;
    static constexpr int NLX = C::VPW * C::PLANE;          // x-lines per block
    static constexpr int LPT = (NLX + C::NT - 1) / C::NT;  // x-lines per thread
    using Regs = cx<T>[LPT][C::N0];

    static FX_HD int nvols(long bid, long nvol) { return Fused<C>::nvols(bid, nvol); }

    // Issue the loads of block b's x-lines.  Lanes past the last line re-read a valid
    // element (their values are never used).
    // Points J0 .. J1-1 of each line (all of them by default).
    template <int J0 = 0, int J1 = C::N0>
    static FX_HD void load_x(int tid, long b, long nvol, const cx<T>* __restrict__ in, Regs& R)
    {
        const int nl = nvols(b, nvol) * C::PLANE;
        const cx<T>* __restrict__ src = in + b * C::VPW * (long)C::V;
#pragma unroll
        for (int i = 0; i < LPT; ++i) {
            int l = tid + i * C::NT;
            l = l < nl ? l : nl - 1;
            const int v = l / C::PLANE;
            const int q = l - v * C::PLANE;
            const cx<T>* __restrict__ p = src + (long)v * C::V + q;
#pragma unroll
            for (int j = J0; j < J1; ++j) R[i][j] = ld_hint<C::HN & 1>(p + j * C::PLANE);
        }
    }
    // x DFT in registers, results to their LDS slots.
    static FX_HD void xpass(int tid, long b, long nvol, Regs& R, cx<T>* sm,
                            const Bank<T>* __restrict__ bk)
    {
        const int nl = nvols(b, nvol) * C::PLANE;
#pragma unroll
        for (int i = 0; i < LPT; ++i) {
            const int l = tid + i * C::NT;
            if (l < nl) {
                const int v = l / C::PLANE;
                const int q = l - v * C::PLANE;
                const int y = q / C::N2;
                const int z = q - y * C::N2;
                const int s0 = v * C::VS + y * C::YS + z;
                dft_put<C::N0, C::S, T, C::V17>(R[i], bk,
                                                [&](int k, cx<T> w) { sm[s0 + k * C::XS] = w; });
            }
        }
    }
    static FX_HD void zpass(int tid, long b, long nvol, cx<T>* sm, const Bank<T>* __restrict__ bk)
    {
        Fused<C>::zpass(tid, b, nvol, sm, bk);
    }
    // y lines (v, x, z): LDS -> HBM.
    // cj: conjugate every output (the backward transform through the forward path).
    static FX_HD void ypass(int tid, long b, long nvol, const cx<T>* sm, cx<T>* __restrict__ out,
                            const Bank<T>* __restrict__ bk, T scl, bool cj = false)
    {
        const int nv = nvols(b, nvol);
        cx<T>* __restrict__ dst = out + b * C::VPW * (long)C::V;
        const bool scaled = scl != T(1);
        for (int l = tid; l < nv * C::N0 * C::N2; l += C::NT) {
            const int v = l / (C::N0 * C::N2);
            const int r = l - v * (C::N0 * C::N2);
            const int x = r / C::N2;
            const int z = r - x * C::N2;
            const int s0 = v * C::VS + x * C::XS + z;
            cx<T> t[C::N1];
#pragma unroll
            for (int j = 0; j < C::N1; ++j) t[j] = sm[s0 + j * C::YS];
            cx<T>* __restrict__ o = dst + (long)v * C::V + x * C::PLANE + z;
            // Hinted only where a z run fills whole 128 B lines, so that every store
            // instruction of a wave writes complete lines; shorter runs are completed by
            // later instructions and lose with the streaming policy.
            constexpr int H = (C::HN & 2) && (C::N2 * (int)sizeof(cx<T>)) % 128 == 0;
            if (cj)
                dft_put<C::N1, C::S, T, C::V17>(t, bk, [&](int k, cx<T> w) {
                    st_hint<H>(o + k * C::N2, cx<T>{w.r * scl, -(w.i * scl)});
                });
            else if (scaled)
                dft_put<C::N1, C::S, T, C::V17>(
                    t, bk, [&](int k, cx<T> w) { st_hint<H>(o + k * C::N2, scale(w, scl)); });
            else
                dft_put<C::N1, C::S, T, C::V17>(
                    t, bk, [&](int k, cx<T> w) { st_hint<H>(o + k * C::N2, w); });
        }
    }
};

// Start of a staggered persistent workgroup: those with bit 3 of the index set wait ~1.6 us.
// Workgroups are placed on the 8 XCDs round-robin by index, so this delays half of the CUs
// of every XCD (delaying whole XCDs, odd indices, measured 1 us slower at 17^3).  Other
// patterns on the two-volume 17^3 complex128 kernel at batch 1024, one device: half the
// workgroups 0.8 us late 33.9 us, 3.2 us late 33.0 us, four groups 0 / 0.8 / 1.6 / 2.4 us
// 32.8 us, against 32.6 us for this one.
__device__ __forceinline__ void stagger_start()
{
    if ((blockIdx.x >> 3) & 1) __builtin_amdgcn_s_sleep(60);  // 60 x 64 cycles
}

// A pipelined kernel launched with at most P workgroups per CU (P > 0) may spend the
// registers of P resident workgroups rather than of as many as the LDS footprint allows.
// ST = 1 staggers the HBM traffic: the next block's x-lines are issued in two halves (the
// first ceil(N0/2) points after the x pass, the rest after the z pass) and half of the
// workgroups start late (stagger_start), so the CUs' read and write phases spread out
// instead of every CU reading at once.  17^3 complex128 at batch 1024: 36.3 us against
// 38.5 us; kernels of ~15 us (8^3, 6^3, 8x8x16) lose more to the late start than they gain.
// ST = 2 issues the next block's x-lines in three parts (after the x, the z and the y
// pass: points [0, ceil(N0/2)), up to N0 - ceil(N0/6), the rest), so that the last part
// does not occupy VGPRs during the y pass: the two-volume 640-thread 17^3 kernel (168 VGPRs
// at 10 waves per CU) otherwise spills two prefetched points and waits for their loads to
// land before spilling them.  Split 9/5/3 at N0 = 17 (161 VGPRs): 36.5-36.9 us against
// 36.7-37.5 us for 6/6/5 on the same device.
template <class C, int P, int ST_ = 0, int HN_ = 3>
struct PipeCfg : C {
    static constexpr int WPE =
        P > 0 ? waves_per_eu(C::LDS * (int)sizeof(cx<typename C::T>), C::NT, P) : C::WPE;
    static constexpr int ST = ST_;
    static constexpr int HN = HN_;
    // x-line points [0, SPLIT) are issued after the x pass, [SPLIT, SPLIT2) after the z
    // pass, [SPLIT2, N0) after the y pass.
    static constexpr int SPLIT = ST_ ? (C::N0 + 1) / 2 : C::N0;
    static constexpr int SPLIT2 = ST_ == 2 ? C::N0 - (C::N0 + 5) / 6 : C::N0;
    static_assert(ST_ >= 0 && ST_ <= 2, "ST: 0, 1 (halves) or 2 (three parts)");
    static_assert(0 < SPLIT && SPLIT <= SPLIT2 && SPLIT2 <= C::N0,
                  "every x-line point is loaded exactly once");
};

// C with the streaming hints off.  The complex64 mixed-geometry launch (6^3 .. 17^3 blocks
// through one persistent kernel, multi.hpp) measured 221 us with hinted x-line loads against
// 207-209 us with plain ones on the same device.
template <class C, int HN_>
struct HintCfg : C {
    static constexpr int HN = HN_;
};
template <class C>
using PlainCfg = HintCfg<C, 0>;

// Claim queues of a launch of `grid` workgroups over nblocks >= grid blocks: Q =
// min(kClaimQueues, grid) counters (claim_layout.hpp), block b in queue b % Q, workgroup w
// drawing from queue w % Q after starting with block w.  Queue q holds nq blocks and gq
// workgroups; each of them draws tickets until one is past the queue's end, so the queue
// hands out exactly nq tickets, 0 .. nq-1 (nq - gq blocks and one refusal per workgroup),
// and the holder of the last one puts its counter back to zero for the next launch.
struct ClaimQueue {
    long q, nqueues, gq, nq;
};
FX_HD ClaimQueue claim_queue(long w, long grid, long nblocks)
{
    const long Q = grid < kClaimQueues ? grid : kClaimQueues;
    const long q = w % Q;
    return ClaimQueue{q, Q, (grid - q + Q - 1) / Q, (nblocks - q + Q - 1) / Q};
}
// Block of ticket t of queue c; >= nblocks once the queue is exhausted.
FX_HD long claim_queue_block(const ClaimQueue& c, long t)
{
    return c.q + c.nqueues * (c.gq + t);
}

// Next block of a persistent workgroup (first block: blockIdx.x).  Claiming hands blocks
// to whichever workgroup of the queue is free first; a fixed stride (work == NULL) lets the
// slowest CU set the time.  `work`: kClaimWords zeroed counters, left at zero.
__device__ __forceinline__ long claim_block(unsigned* work, long nblocks, long b)
{
    if (!work) return b + gridDim.x;
    const ClaimQueue c = claim_queue(blockIdx.x, gridDim.x, nblocks);
    unsigned* ctr = work + c.q * kClaimLine;
    const unsigned t = atomicAdd(ctr, 1u);
    if ((long)t == c.nq - 1) (void)atomicExch(ctr, 0u);
    return claim_queue_block(c, (long)t);
}

template <class C>
__global__ void __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::WPE, 8)))
    k_fpipe(const cx<typename C::T>* __restrict__ in, cx<typename C::T>* __restrict__ out,
            long nvol, long nblocks, const Bank<typename C::T>* __restrict__ bk,
            typename C::T scl, unsigned* __restrict__ work)
{
    using K = FusedPipe<C>;
    __shared__ cx<typename C::T> sm[C::LDS];
    __shared__ long s_next;
    typename K::Regs R;
    const int tid = threadIdx.x;
    long b = blockIdx.x;
    if (b >= nblocks) return;
    if (C::ST) stagger_start();
    K::load_x(tid, b, nvol, in, R);
    for (;;) {
        long claim = 0;
        if (tid == 0) claim = claim_block(work, nblocks, b);
        K::xpass(tid, b, nvol, R, sm, bk);
        if (tid == 0) s_next = claim;
        __syncthreads();
        const long nb = s_next;
        if (nb < nblocks) K::template load_x<0, C::SPLIT>(tid, nb, nvol, in, R);
        K::zpass(tid, b, nvol, sm, bk);
        if (C::SPLIT < C::SPLIT2 && nb < nblocks)
            K::template load_x<C::SPLIT, C::SPLIT2>(tid, nb, nvol, in, R);
        __syncthreads();
        K::ypass(tid, b, nvol, sm, out, bk, scl);
        if (nb >= nblocks) break;
        if (C::SPLIT2 < C::N0) K::template load_x<C::SPLIT2, C::N0>(tid, nb, nvol, in, R);
        __syncthreads();
        b = nb;
    }
}

// ------------------------------------------------------------------------------ pencil
// PIPE = 0: pass A is k_slab (one slab group per workgroup); PIPE = P > 0: pass A is the
// persistent pipelined k_spipe with at most P resident workgroups per CU.
// ST = 1: k_spipe staggers its HBM traffic as PipeCfg ST does (y-lines of the next group
// issued in two halves, half of the workgroups starting ~1.6 us late).
// ALG: line algorithm of the codelets (dft's fourth parameter; 1 = factors in the other order).
// XH: hints of pass B, bit 0 its loads, bit 1 its stores (nontemporal when set).
template <int N0_, int N1_, int N2_, typename T_, int S_, int SPW_, int NT_, int PZ_,
          int NTB_, int CHUNK_, int PIPE_ = 0, int ST_ = 0, int ALG_ = 0, int XH_ = 3>
struct PencilCfg {
    using T = T_;
    static constexpr int N0 = N0_, N1 = N1_, N2 = N2_, S = S_, SPW = SPW_, NT = NT_;
    static constexpr int PZ = PZ_, NTB = NTB_, CHUNK = CHUNK_, PIPE = PIPE_, ST = ST_;
    static constexpr int ALG = ALG_, XH = XH_;
    static constexpr int SPLIT = ST_ ? (N1_ + 1) / 2 : N1_;  // y points in the first half
    static constexpr int V = N0 * N1 * N2;
    static constexpr int PLANE = N1 * N2;
    static constexpr int YS = N2 + PZ;
    static constexpr int SS = N1 * YS;
    static constexpr int LDS = SPW * SS;
    static constexpr int WPE = waves_per_eu(LDS * (int)sizeof(cx<T>), NT, PIPE);
};

template <class C>
struct Pencil {
    using T = typename C::T;

    static FX_HD int nslabs(long bid, long nslab)
    {
        const long left = nslab - bid * C::SPW;
        return (int)(left < C::SPW ? left : C::SPW);
    }
    static FX_HD int slot(int e)
    {
        const int s = e / C::PLANE;
        const int q = e - s * C::PLANE;
        const int y = q / C::N2;
        const int z = q - y * C::N2;
        return s * C::SS + y * C::YS + z;
    }
    // Pass A, phase 1: stage SPW contiguous (v, x) slabs.
    static FX_HD void stage(int tid, long bid, long nslab, const cx<T>* __restrict__ in, cx<T>* sm)
    {
        const int ns = nslabs(bid, nslab);
        const cx<T>* __restrict__ src = in + bid * C::SPW * (long)C::PLANE;
        if (ns == C::SPW)
            stage_full<C::SPW * C::PLANE, C::NT>(tid, src, sm, [](int e) { return slot(e); });
        else
            for (int e = tid; e < ns * C::PLANE; e += C::NT) sm[slot(e)] = ld_stream(src + e);
    }
    // Pass A, phase 2: z lines (s, y).
    static FX_HD void zpass(int tid, long bid, long nslab, cx<T>* sm, const Bank<T>* __restrict__ bk)
    {
        lds_pass<C::N2, 1, C::S, C::ALG, C::NT>(
            tid, sm, nslabs(bid, nslab) * C::N1,
            [](int l) {
                const int s = l / C::N1;
                const int y = l - s * C::N1;
                return s * C::SS + y * C::YS;
            },
            bk);
    }
    // Pass A, phase 3: y lines (s, z), LDS -> out (coalesced across z).
    static FX_HD void ypass(int tid, long bid, long nslab, const cx<T>* sm, cx<T>* __restrict__ out,
                            const Bank<T>* __restrict__ bk)
    {
        const int ns = nslabs(bid, nslab);
        cx<T>* __restrict__ dst = out + bid * C::SPW * (long)C::PLANE;
        for (int l = tid; l < ns * C::N2; l += C::NT) {
            const int s = l / C::N2;
            const int z = l - s * C::N2;
            const int b = s * C::SS + z;
            cx<T> t[C::N1];
#pragma unroll
            for (int j = 0; j < C::N1; ++j) t[j] = sm[b + j * C::YS];
            dft<C::N1, C::S, T, C::ALG>(t, bk);
            cx<T>* __restrict__ o = dst + (long)s * C::PLANE + z;
#pragma unroll
            for (int k = 0; k < C::N1; ++k) o[k * C::N2] = t[k];
        }
    }
    // Pass B: x-line l = (v, y, z) of buf, in place.
    static FX_HD void xline(long l, long nlines, cx<T>* __restrict__ buf,
                            const Bank<T>* __restrict__ bk, T scl)
    {
        if (l >= nlines) return;
        const long v = l / C::PLANE;
        const int q = (int)(l - v * C::PLANE);
        cx<T>* __restrict__ p = buf + v * C::V + q;
        cx<T> t[C::N0];
#pragma unroll
        for (int j = 0; j < C::N0; ++j) t[j] = ld_hint<C::XH & 1>(p + j * C::PLANE);
        dft<C::N0, C::S, T, C::ALG>(t, bk);
        if (scl != T(1)) {
#pragma unroll
            for (int k = 0; k < C::N0; ++k) st_hint<C::XH & 2>(p + k * C::PLANE, scale(t[k], scl));
        } else {
#pragma unroll
            for (int k = 0; k < C::N0; ++k) st_hint<C::XH & 2>(p + k * C::PLANE, t[k]);
        }
    }
};

template <class C>
__global__ void __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::WPE, 8)))
    k_slab(const cx<typename C::T>* __restrict__ in, cx<typename C::T>* __restrict__ out,
           long nslab, const Bank<typename C::T>* __restrict__ bk)
{
    using K = Pencil<C>;
    __shared__ cx<typename C::T> sm[C::LDS];
    const int tid = threadIdx.x;
    const long bid = blockIdx.x;
    K::stage(tid, bid, nslab, in, sm);
    __syncthreads();
    K::zpass(tid, bid, nslab, sm, bk);
    __syncthreads();
    K::ypass(tid, bid, nslab, sm, out, bk);
}

// Pass A, persistent and pipelined: the y-lines (s, z) of a slab group come straight from
// HBM into VGPRs (lanes hold consecutive z: contiguous runs), are transformed and written
// to LDS, and the next group's y-lines are issued at once; then z in LDS and a coalesced
// store of the group.
template <class C>
struct SlabPipe {
    using T = typename C::T;
    static constexpr int NLY = C::SPW * C::N2;             // y-lines per group
    static constexpr int LPT = (NLY + C::NT - 1) / C::NT;
    using Regs = cx<T>[LPT][C::N1];

    static FX_HD int nslabs(long bid, long nslab) { return Pencil<C>::nslabs(bid, nslab); }

    // Points J0 .. J1-1 of each y-line (all of them by default).
    template <int J0 = 0, int J1 = C::N1>
    static FX_HD void load_y(int tid, long b, long nslab, const cx<T>* __restrict__ in, Regs& R)
    {
        const int nl = nslabs(b, nslab) * C::N2;
        const cx<T>* __restrict__ src = in + b * C::SPW * (long)C::PLANE;
#pragma unroll
        for (int i = 0; i < LPT; ++i) {
            int l = tid + i * C::NT;
            l = l < nl ? l : nl - 1;
            const int s = l / C::N2;
            const int z = l - s * C::N2;
            const cx<T>* __restrict__ p = src + (long)s * C::PLANE + z;
#pragma unroll
            for (int j = J0; j < J1; ++j) R[i][j] = ld_stream(p + j * C::N2);
        }
    }
    static FX_HD void ypass(int tid, long b, long nslab, Regs& R, cx<T>* sm,
                            const Bank<T>* __restrict__ bk)
    {
        const int nl = nslabs(b, nslab) * C::N2;
#pragma unroll
        for (int i = 0; i < LPT; ++i) {
            const int l = tid + i * C::NT;
            if (l < nl) {
                dft<C::N1, C::S, T, C::ALG>(R[i], bk);
                const int s = l / C::N2;
                const int z = l - s * C::N2;
#pragma unroll
                for (int k = 0; k < C::N1; ++k) sm[s * C::SS + k * C::YS + z] = R[i][k];
            }
        }
    }
    static FX_HD void zpass(int tid, long b, long nslab, cx<T>* sm, const Bank<T>* __restrict__ bk)
    {
        Pencil<C>::zpass(tid, b, nslab, sm, bk);
    }
    static FX_HD void store(int tid, long b, long nslab, const cx<T>* sm, cx<T>* __restrict__ out)
    {
        const int n = nslabs(b, nslab) * C::PLANE;
        cx<T>* __restrict__ dst = out + b * C::SPW * (long)C::PLANE;
        for (int e = tid; e < n; e += C::NT) dst[e] = sm[Pencil<C>::slot(e)];
    }
};

template <class C>
__global__ void __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::WPE, 8)))
    k_spipe(const cx<typename C::T>* __restrict__ in, cx<typename C::T>* __restrict__ out,
            long nslab, long nblocks, const Bank<typename C::T>* __restrict__ bk,
            unsigned* __restrict__ work)
{
    using K = SlabPipe<C>;
    __shared__ cx<typename C::T> sm[C::LDS];
    __shared__ long s_next;
    typename K::Regs R;
    const int tid = threadIdx.x;
    long b = blockIdx.x;
    if (b >= nblocks) return;
    if (C::ST) stagger_start();
    K::load_y(tid, b, nslab, in, R);
    for (;;) {
        long claim = 0;
        if (tid == 0) claim = claim_block(work, nblocks, b);
        K::ypass(tid, b, nslab, R, sm, bk);
        if (tid == 0) s_next = claim;
        __syncthreads();
        const long nb = s_next;
        if (nb < nblocks) K::template load_y<0, C::SPLIT>(tid, nb, nslab, in, R);
        K::zpass(tid, b, nslab, sm, bk);
        if (C::SPLIT < C::N1 && nb < nblocks)
            K::template load_y<C::SPLIT, C::N1>(tid, nb, nslab, in, R);
        __syncthreads();
        K::store(tid, b, nslab, sm, out);
        if (nb >= nblocks) break;
        __syncthreads();
        b = nb;
    }
}

template <class C>
__global__ void __launch_bounds__(C::NTB)
    k_xline(cx<typename C::T>* __restrict__ buf, long nlines,
            const Bank<typename C::T>* __restrict__ bk, typename C::T scl)
{
    Pencil<C>::xline((long)blockIdx.x * C::NTB + threadIdx.x, nlines, buf, bk, scl);
}

}  // namespace fx

#include <fx_launch.hpp>
