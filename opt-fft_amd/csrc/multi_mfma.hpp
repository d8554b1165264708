// Mixed-geometry launch of the DFT-as-matrix variant (complex64, mfma.hpp): the items and
// task list of multi.hpp, each task one block of mfma_block with the forward matrices
// (backward items conjugate on load and store).
#pragma once
#include "mfma.hpp"
#include "multi.hpp"

namespace fx {

// Tiles of the matrix-core kernels (mfma.hpp), one workgroup size.  The LDS array is sized
// for 17^3, so every task runs at four workgroups per CU; halving the 6^3 and 8^3 tiles (the
// standalone kernels' best) only adds tasks: 200.8 against 192.9 us for the bf16x3 mix.
struct MultiMfmaPreset {
    static constexpr int NT = 256;
    using G0 = FusedCfg<6, 6, 6, float, -1, 16, NT, 1, 0, 0>;
    using G1 = FusedCfg<8, 8, 8, float, -1, 8, NT, 1, 0, 0>;
    using G2 = FusedCfg<12, 12, 12, float, -1, 2, NT, 1, 0, 0>;
    using G3 = FusedCfg<17, 17, 17, float, -1, 1, NT, 0, 0, 0>;
    using G4 = FusedCfg<8, 8, 16, float, -1, 4, NT, 1, 0, 0>;
    static constexpr int LDS =
        cmax(cmax(cmax(G0::LDS, G1::LDS), cmax(G2::LDS, G3::LDS)), G4::LDS);
    static constexpr int WPE = waves_per_eu(LDS * (int)sizeof(cx<float>), NT, 0);
    static constexpr int vpw(int code)
    {
        return code == 0 ? G0::VPW : code == 1 ? G1::VPW : code == 2 ? G2::VPW
             : code == 3 ? G3::VPW : G4::VPW;
    }
};

// One block of geometry G with its matrices built for it: 17^3 by the conjugate-pair fold,
// the K = 32 instructions on the other axes with 2N > 16; the x pass stores straight to HBM
// for the short lines and through LDS for 17 (the faster form of each).
template <class G, int SPLIT>
__device__ __forceinline__ void run_mfma_task(long bid, long nvol, const cx<float>* in,
                                              cx<float>* out, const Bank<float>* bk, float scl,
                                              bool conj, cx<float>* sm)
{
    constexpr int KW = (G::N0 == 17 && G::N1 == 17 && G::N2 == 17) ? kFold17 : 32;
    MfmaMats<G, SPLIT, KW> A;
    A.build(threadIdx.x & 63, bk);
    mfma_block<G, SPLIT, KW, (G::N0 > 12 ? 0 : 1)>(bid, nvol, in, out, scl, conj, A, sm);
}

// One task per workgroup (grid = ntasks; the hardware dispatcher hands tasks, heaviest
// first, to whichever CU frees up).  A persistent form drawing tasks from the claim queues
// was built and measured: it spilled (130 VGPRs, lane-derived offsets of five geometries held
// across the loop) until those were recomputed per task, and then took 205.7 us for the
// bf16x3 mix against 190-196 us for this form; not kept.
template <int SPLIT>
__global__ void __launch_bounds__(MultiMfmaPreset::NT)
    __attribute__((amdgpu_waves_per_eu(MultiMfmaPreset::WPE, 8)))
    k_multi_mfma(const MultiArgs<float> a, const int2* __restrict__ tasks, int ntasks,
                 const Bank<float>* __restrict__ bk)
{
    using P = MultiMfmaPreset;
    __shared__ cx<float> sm[P::LDS];
    const long t = blockIdx.x;
    if (t >= ntasks) return;
    const int2 task = tasks[t];
    const MultiItem<float>* ip = &a.item[task.x];
    const cx<float>* in = ip->in;
    cx<float>* out = ip->out;
    const long nvol = ip->nvol;
    const bool cj = ip->conj != 0;
    const float scl = ip->scale;
    switch (ip->code) {  // uniform over the workgroup
    case 0: run_mfma_task<P::G0, SPLIT>(task.y, nvol, in, out, bk, scl, cj, sm); break;
    case 1: run_mfma_task<P::G1, SPLIT>(task.y, nvol, in, out, bk, scl, cj, sm); break;
    case 2: run_mfma_task<P::G2, SPLIT>(task.y, nvol, in, out, bk, scl, cj, sm); break;
    case 3: run_mfma_task<P::G3, SPLIT>(task.y, nvol, in, out, bk, scl, cj, sm); break;
    default: run_mfma_task<P::G4, SPLIT>(task.y, nvol, in, out, bk, scl, cj, sm); break;
    }
}

}  // namespace fx

#include <fx_multi_mfma_launch.hpp>
