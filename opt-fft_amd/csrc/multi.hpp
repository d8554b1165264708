// One persistent launch over a mixed batch: several items, each a batch of volumes of one
// of the fused geometries (6^3, 8^3, 12^3, 17^3, 8x8x16) in either direction.
//
// The host cuts every item into blocks (VPW volumes, the tile the preset gives that
// geometry), sorts the blocks heaviest first into a task list, and launches one grid of
// resident workgroups that take tasks from a claim counter (claim_block).  A task runs the pipelined fused
// pass structure of its geometry (x from HBM into VGPRs, z in LDS, y from LDS to HBM) out
// of one LDS buffer sized for the largest block.  The backward direction reuses the forward
// code: conj(DFT(conj(x))) is the unnormalised backward DFT, and both conjugations are sign
// flips folded into the load and the store.
#pragma once
#include "kernels.hpp"

namespace fx {

constexpr int kMultiMaxItems = 16;
constexpr int kMultiGeoms = 5;   // 6^3, 8^3, 12^3, 17^3, 8x8x16
constexpr int kMultiShape[kMultiGeoms][3] = {{6, 6, 6}, {8, 8, 8}, {12, 12, 12}, {17, 17, 17},
                                             {8, 8, 16}};

template <typename T>
struct MultiItem {
    const cx<T>* in;
    cx<T>* out;
    long nvol;
    int code;    // index into kMultiShape
    int conj;    // 1 for the backward direction
    T scale;
};

template <typename T>
struct MultiArgs {
    MultiItem<T> item[kMultiMaxItems];
};

// Tile presets: workgroup size and volumes per block for each geometry code.
template <typename T, int PRESET>
struct MultiPreset;
// 320 threads: 90-100% of the lanes busy in every line pass.
template <typename T>
struct MultiPreset<T, 0> {
    static constexpr int NT = 320;
    static constexpr int HN = 0;  // streaming hints (FusedCfg::HN): none
    using G0 = FusedCfg<6, 6, 6, T, -1, 8, NT, 1, 0, 1>;
    using G1 = FusedCfg<8, 8, 8, T, -1, 5, NT, 1, 0, 1>;
    using G2 = FusedCfg<12, 12, 12, T, -1, 2, NT, 1, 0, 1>;
    using G3 = FusedCfg<17, 17, 17, T, -1, 1, NT, 0, 0, 1>;
    using G4 = FusedCfg<8, 8, 16, T, -1, 2, NT, 1, 0, 1>;
};
// 640 threads: twice the volumes per block.
template <typename T>
struct MultiPreset<T, 1> {
    static constexpr int NT = 640;
    static constexpr int HN = 0;
    using G0 = FusedCfg<6, 6, 6, T, -1, 16, NT, 1, 0, 1>;
    using G1 = FusedCfg<8, 8, 8, T, -1, 10, NT, 1, 0, 1>;
    using G2 = FusedCfg<12, 12, 12, T, -1, 4, NT, 1, 0, 1>;
    using G3 = FusedCfg<17, 17, 17, T, -1, 2, NT, 0, 0, 1>;
    using G4 = FusedCfg<8, 8, 16, T, -1, 5, NT, 1, 0, 1>;
};

// 640 threads, every geometry's block sized to the LDS the two 17^3 volumes need (78.6 KB
// at complex64): the small geometries move 2-4x the bytes per task of preset 1, in the same
// two resident workgroups per CU.
template <typename T>
struct MultiPreset<T, 2> {
    static constexpr int NT = 640;
    static constexpr int HN = 0;
    using G0 = FusedCfg<6, 6, 6, T, -1, 38, NT, 1, 0, 1>;
    using G1 = FusedCfg<8, 8, 8, T, -1, 17, NT, 1, 0, 1>;
    using G2 = FusedCfg<12, 12, 12, T, -1, 5, NT, 1, 0, 1>;
    using G3 = FusedCfg<17, 17, 17, T, -1, 2, NT, 0, 0, 1>;
    using G4 = FusedCfg<8, 8, 16, T, -1, 9, NT, 1, 0, 1>;
};
// 320 threads, every geometry's block sized to the LDS of one 17^3 volume (39.3 KB at
// complex64, 78.6 KB at complex128): four (complex64) or two (complex128) resident
// workgroups per CU, each with a whole block of its own.
template <typename T>
struct MultiPreset<T, 3> {
    static constexpr int NT = 320;
    static constexpr int HN = 0;
    using G0 = FusedCfg<6, 6, 6, T, -1, 19, NT, 1, 0, 1>;
    using G1 = FusedCfg<8, 8, 8, T, -1, 8, NT, 1, 0, 1>;
    using G2 = FusedCfg<12, 12, 12, T, -1, 2, NT, 1, 0, 1>;
    using G3 = FusedCfg<17, 17, 17, T, -1, 1, NT, 0, 0, 1>;
    using G4 = FusedCfg<8, 8, 16, T, -1, 4, NT, 1, 0, 1>;
};
// The staged task structure (block HBM -> LDS in one grouped read, z and y in LDS, x pass
// LDS -> HBM with whole-line stores; the per-geometry plans of 6^3, 8^3 and 12^3 complex64
// race to it) was built on presets 0 and 3 and measured slower in one launch: 213-228 us on
// the mix below against 164.5 us for preset 0, and alone 69.8 us at 6^3 x 65536 against
// 50.6 us (the per-geometry staged kernel: 40.0 us at six workgroups per CU; here four, and a
// staged task computes nothing until its whole block has landed).  Not built.
// Measured on the complex64 mix 6^3 x 65536, 8^3 x 32768, 12^3 x 8192, 17^3 x 2048 (not
// built): preset 3 with the items interleaved and nontemporal stores 168.6 us, with
// nontemporal loads, or loads and stores, 174.0 us, against 168.9 us plain on one device;
// preset 2 with nontemporal loads, or loads and stores, 207.8 us against 186.8 us plain.
// (The same blocks on 1024 threads, 32 waves per CU, spill at the 64-VGPR cap: not built.)

constexpr int cmax(int a, int b) { return a > b ? a : b; }

template <typename T, int PRESET>
struct MultiCfg {
    using P = MultiPreset<T, PRESET>;
    static constexpr int NT = P::NT;
    static constexpr int LDS =
        cmax(cmax(cmax(P::G0::LDS, P::G1::LDS), cmax(P::G2::LDS, P::G3::LDS)), P::G4::LDS);
    static constexpr int WPE = waves_per_eu(LDS * (int)sizeof(cx<T>), NT, 0);
    static constexpr int vpw(int code)
    {
        return code == 0 ? P::G0::VPW : code == 1 ? P::G1::VPW : code == 2 ? P::G2::VPW
             : code == 3 ? P::G3::VPW : P::G4::VPW;
    }
};

// Phase `ph` (0: x from HBM, 1: z, 2: y to HBM) of one block of one item.  Phases are
// separated by workgroup barriers; no register state crosses a barrier.
template <class C, int HN>
FX_HD void multi_phase(int ph, int tid, long b, const MultiItem<typename C::T>& it,
                       cx<typename C::T>* sm, const Bank<typename C::T>* __restrict__ bk)
{
    using K = FusedPipe<HintCfg<C, HN>>;
    if (ph == 0) {
        typename K::Regs R;
        K::load_x(tid, b, it.nvol, it.in, R);
        if (it.conj) {
#pragma unroll
            for (int i = 0; i < K::LPT; ++i)
#pragma unroll
                for (int j = 0; j < C::N0; ++j) R[i][j].i = -R[i][j].i;
        }
        K::xpass(tid, b, it.nvol, R, sm, bk);
    } else if (ph == 1) {
        K::zpass(tid, b, it.nvol, sm, bk);
    } else {
        K::ypass(tid, b, it.nvol, sm, it.out, bk, it.scale, it.conj != 0);
    }
}

template <typename T, int PRESET>
FX_HD void multi_task_phase(int ph, int tid, int2 task, const MultiArgs<T>& a, cx<T>* sm,
                            const Bank<T>* __restrict__ bk)
{
    using P = MultiPreset<T, PRESET>;
    const MultiItem<T>& it = a.item[task.x];
    switch (it.code) {
    case 0: multi_phase<typename P::G0, P::HN>(ph, tid, task.y, it, sm, bk); break;
    case 1: multi_phase<typename P::G1, P::HN>(ph, tid, task.y, it, sm, bk); break;
    case 2: multi_phase<typename P::G2, P::HN>(ph, tid, task.y, it, sm, bk); break;
    case 3: multi_phase<typename P::G3, P::HN>(ph, tid, task.y, it, sm, bk); break;
    default: multi_phase<typename P::G4, P::HN>(ph, tid, task.y, it, sm, bk); break;
    }
}

// Walk the task list: the first task is blockIdx.x, the next ones are claimed from `work`
// (NULL: a fixed stride of gridDim.x).  The claim is issued before phase 0 and published to
// the workgroup at the first barrier.
template <typename T, int PRESET>
__device__ __forceinline__ void multi_walk(const MultiArgs<T>& a, const int2* __restrict__ tasks,
                                           int ntasks, const Bank<T>* __restrict__ bk,
                                           unsigned* work, cx<T>* sm, long* s_next)
{
    const int tid = threadIdx.x;
    long t = blockIdx.x;
    while (t < ntasks) {
        long claim = 0;
        if (tid == 0) claim = claim_block(work, ntasks, t);
        const int2 task = tasks[t];
        multi_task_phase<T, PRESET>(0, tid, task, a, sm, bk);
        if (tid == 0) *s_next = claim;
        __syncthreads();
        const long nt = *s_next;
        multi_task_phase<T, PRESET>(1, tid, task, a, sm, bk);
        __syncthreads();
        multi_task_phase<T, PRESET>(2, tid, task, a, sm, bk);
        __syncthreads();
        t = nt;
    }
}

template <typename T, int PRESET>
__global__ void __launch_bounds__((MultiCfg<T, PRESET>::NT))
    __attribute__((amdgpu_waves_per_eu(MultiCfg<T, PRESET>::WPE, 8)))
    k_multi(const MultiArgs<T> a, const int2* __restrict__ tasks, int ntasks,
            const Bank<T>* __restrict__ bk, unsigned* __restrict__ work)
{
    using M = MultiCfg<T, PRESET>;
    __shared__ cx<T> sm[M::LDS];
    __shared__ long s_next;
    multi_walk<T, PRESET>(a, tasks, ntasks, bk, work, sm, &s_next);
}

// Same kernel with the items, the task count and the task list read from device memory, so a
// launch can be captured once in a hipGraph and replayed with new descriptors (serving).
// `work`: the launch's claim counters (kClaimWords, zero before the first launch and left
// at zero by every launch, claim_block), owned by whoever owns the descriptor.
template <typename T>
struct MultiDescriptor {
    MultiArgs<T> args;
    int ntasks;
    int2 tasks[1];  // ntasks entries follow
};

template <typename T, int PRESET>
__global__ void __launch_bounds__((MultiCfg<T, PRESET>::NT))
    __attribute__((amdgpu_waves_per_eu(MultiCfg<T, PRESET>::WPE, 8)))
    k_multi_dev(MultiDescriptor<T>* __restrict__ d, const Bank<T>* __restrict__ bk,
                unsigned* __restrict__ work)
{
    using M = MultiCfg<T, PRESET>;
    __shared__ cx<T> sm[M::LDS];
    __shared__ long s_next;
    const int ntasks = d->ntasks;
    multi_walk<T, PRESET>(d->args, d->tasks, ntasks, bk, ntasks >= 2L * gridDim.x ? work : nullptr,
                          sm, &s_next);
}

// Registry entry of a multi kernel.
// `work`: claim counter as for LaunchFn (registry.hpp).
typedef hipError_t (*MultiLaunchFn)(const void* args, const int2* tasks, int ntasks,
                                    const void* bank, unsigned* work, hipStream_t stream);
// Launch of the descriptor-reading form on `grid` workgroups with claim counters `work`.
typedef hipError_t (*MultiDevLaunchFn)(const void* descriptor, const void* bank, unsigned* work,
                                       int grid, hipStream_t stream);
struct MultiDesc {
    int prec;
    int preset;
    int nthreads;
    int vpw[kMultiGeoms];
    size_t lds_bytes;
    const char* name;
    MultiLaunchFn launch;
    MultiDevLaunchFn launch_dev;  // NULL: no descriptor form
    int (*resident_grid)();  // persistent grid of the descriptor form on the current device
    int algo = 0;            // as KernelDesc::algo: 0 FFT, 1 / 2 matrix-core bf16 / fp16
    long (*launch_grid)(long ntasks) = nullptr;  // workgroups `launch` starts for ntasks
    int order = 0;  // task order (multi_tasks): 0 heaviest first, 1 items interleaved
};

void register_multi(std::vector<MultiDesc>& r);
const std::vector<MultiDesc>& multi_registry();

}  // namespace fx

#include <fx_multi_launch.hpp>
