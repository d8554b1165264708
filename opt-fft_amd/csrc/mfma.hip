// Matrix-core DFT-as-matrix kernels (mfma.hpp): complex64, L = 6, 8, 12, 17 and 8x8x16,
// both directions, split-bf16 and split-fp16 compensated products.  Selected only when a
// plan asks for them (FFT3D_MX_MFMA_BF16 / FFT3D_MX_MFMA_FP16).
#include "mfma.hpp"

namespace fx {

#define FX_M(N0, N1, N2, S, VPW, NT, PZ, SPLIT, NAME) \
    r.push_back(mfma_desc<FusedCfg<N0, N1, N2, float, S, VPW, NT, PZ, 0, 0>, SPLIT>(NAME))
#define FX_M_ALL(N0, N1, N2, VPW, NT, PZ)                                                          \
    FX_M(N0, N1, N2, -1, VPW, NT, PZ, 0, "mfma_" #N0 "x" #N1 "x" #N2 "_c64_fwd_v" #VPW "t" #NT "_bf16x3"); \
    FX_M(N0, N1, N2, +1, VPW, NT, PZ, 0, "mfma_" #N0 "x" #N1 "x" #N2 "_c64_bwd_v" #VPW "t" #NT "_bf16x3"); \
    FX_M(N0, N1, N2, -1, VPW, NT, PZ, 1, "mfma_" #N0 "x" #N1 "x" #N2 "_c64_fwd_v" #VPW "t" #NT "_fp16x3"); \
    FX_M(N0, N1, N2, +1, VPW, NT, PZ, 1, "mfma_" #N0 "x" #N1 "x" #N2 "_c64_bwd_v" #VPW "t" #NT "_fp16x3")

// The same with the gfx950 K = 32 instructions on the axes with 2N > 16 ("k32"), for
// N = 12 and 17.
#define FX_MK(N0, N1, N2, S, VPW, NT, PZ, SPLIT, NAME) \
    r.push_back(mfma_desc<FusedCfg<N0, N1, N2, float, S, VPW, NT, PZ, 0, 0>, SPLIT, 32>(NAME))
#define FX_MK_ALL(N0, N1, N2, VPW, NT, PZ)                                                          \
    FX_MK(N0, N1, N2, -1, VPW, NT, PZ, 0, "mfma_" #N0 "x" #N1 "x" #N2 "_c64_fwd_v" #VPW "t" #NT "k32_bf16x3"); \
    FX_MK(N0, N1, N2, +1, VPW, NT, PZ, 0, "mfma_" #N0 "x" #N1 "x" #N2 "_c64_bwd_v" #VPW "t" #NT "k32_bf16x3"); \
    FX_MK(N0, N1, N2, -1, VPW, NT, PZ, 1, "mfma_" #N0 "x" #N1 "x" #N2 "_c64_fwd_v" #VPW "t" #NT "k32_fp16x3"); \
    FX_MK(N0, N1, N2, +1, VPW, NT, PZ, 1, "mfma_" #N0 "x" #N1 "x" #N2 "_c64_bwd_v" #VPW "t" #NT "k32_fp16x3")

void register_mfma(std::vector<KernelDesc>& r)
{
    FX_M_ALL(6, 6, 6, 16, 256, 1);
    FX_M_ALL(8, 8, 8, 8, 256, 1);
    FX_M_ALL(12, 12, 12, 2, 256, 1);
    FX_M_ALL(17, 17, 17, 1, 256, 0);
    FX_M_ALL(8, 8, 16, 4, 256, 1);
    FX_MK_ALL(12, 12, 12, 2, 256, 1);
    FX_MK_ALL(17, 17, 17, 1, 256, 0);
}

}  // namespace fx
