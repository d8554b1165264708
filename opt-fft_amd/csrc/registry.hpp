// Kernel registry: every compiled kernel variant describes itself with one KernelDesc.
// The plan step (plan.cpp) filters the registry by geometry, precision and direction and,
// when more than one variant matches, races them at create time.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <vector>

#include "claim_layout.hpp"

namespace fx {

enum Prec { kC128 = 0, kC64 = 1 };
enum Kind { kFused = 0, kPencil = 1 };

// Launch on `stream`: transform nvol contiguous volumes from `in` to `out` (device
// pointers, out-of-place), multiplying every output by `scale`.  `bank` is the device
// Bank<T> of matching precision.  `work` points to kClaimWords device claim counters
// (claim_layout.hpp) that are zero when the launch starts and that the launch leaves at
// zero (persistent kernels hand out blocks from them); launches that share counters must be
// ordered on one stream.  NULL makes persistent kernels walk their blocks in a fixed order
// instead.  Returns the launch status.
typedef hipError_t (*LaunchFn)(const void* in, void* out, long nvol, const void* bank,
                               double scale, unsigned* work, hipStream_t stream);

struct KernelDesc {
    int n0, n1, n2;
    int prec;        // Prec
    int dir;         // -1 forward, +1 backward
    int kind;        // Kind
    int vpw;         // volumes (fused) or slabs (pencil pass A) per workgroup
    int nthreads;    // workgroup size of the main kernel
    int variant;     // line algorithm variant (17: 0 Rader, 1 folded dense)
    int persist;     // 0: one block per workgroup; P > 0: persistent, pipelined, <= P per CU
    size_t lds_bytes;
    const char* name;
    LaunchFn launch;
    int algo = 0;    // 0 exact FP FFT; 1 DFT matrix on MFMA, split bf16; 2 split fp16
};

// Appended by each kernel translation unit.
void register_fused_small(std::vector<KernelDesc>& r);   // 6, 8, 12, 8x8x16
void register_fused_17(std::vector<KernelDesc>& r);      // 17
void register_pencil(std::vector<KernelDesc>& r);        // 24, 36, 24x24x48
void register_split(std::vector<KernelDesc>& r);         // 24 complex128, single pass
void register_mfma(std::vector<KernelDesc>& r);          // complex64 DFT-as-matrix
void register_f17(std::vector<KernelDesc>& r);           // 17, balanced four-wave kernel
void register_dense(std::vector<KernelDesc>& r);         // dense row-column floor, every geometry

const std::vector<KernelDesc>& registry();

}  // namespace fx
