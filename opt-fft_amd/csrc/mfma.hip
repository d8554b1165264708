// Matrix-core DFT-as-matrix kernels (mfma.hpp): complex64, L = 6, 8, 12, 17 and 8x8x16,
// both directions, split-bf16 and split-fp16 compensated products.  Selected only when a
// plan asks for them (FFT3D_MX_MFMA_BF16 / FFT3D_MX_MFMA_FP16); the create-time race picks
// among the forms registered here.
//
// Name: mfma_<shape>_c64_<dir>_v<VPW>t<NT>[k32][p<P>][d]_<split>x3
//   k32  the gfx950 K = 32 instructions on the axes with 2N > 16 (otherwise K = 16);
//   p<P> persistent and pipelined, at most P workgroups per CU (otherwise one block each);
//   d    the x pass stores straight to HBM (otherwise through LDS and one coalesced copy);
//   fold 17-point lines by the conjugate-pair fold on K = 32 instructions.
#include "mfma.hpp"

namespace fx {

#define FX_MG(N0, N1, N2, S, VPW, NT, PZ, SPLIT, KW, P, XD, NAME)                                \
    r.push_back(mfma_desc<FusedCfg<N0, N1, N2, float, S, VPW, NT, PZ, 0, 0>, SPLIT, KW, P, XD>( \
        NAME))
// Both directions and both splits of one form; TAG is the form's part of the name.
#define FX_MG4(N0, N1, N2, VPW, NT, PZ, KW, P, XD, TAG)                                         \
    FX_MG(N0, N1, N2, -1, VPW, NT, PZ, 0, KW, P, XD,                                            \
          "mfma_" #N0 "x" #N1 "x" #N2 "_c64_fwd_v" #VPW "t" #NT TAG "_bf16x3");                \
    FX_MG(N0, N1, N2, +1, VPW, NT, PZ, 0, KW, P, XD,                                            \
          "mfma_" #N0 "x" #N1 "x" #N2 "_c64_bwd_v" #VPW "t" #NT TAG "_bf16x3");                \
    FX_MG(N0, N1, N2, -1, VPW, NT, PZ, 1, KW, P, XD,                                            \
          "mfma_" #N0 "x" #N1 "x" #N2 "_c64_fwd_v" #VPW "t" #NT TAG "_fp16x3");                \
    FX_MG(N0, N1, N2, +1, VPW, NT, PZ, 1, KW, P, XD,                                            \
          "mfma_" #N0 "x" #N1 "x" #N2 "_c64_bwd_v" #VPW "t" #NT TAG "_fp16x3")

void register_mfma(std::vector<KernelDesc>& r)
{
    // The measured-fastest form of each geometry first (the no-race default and the
    // race's tie-break), then the others.
    FX_MG4(6, 6, 6, 8, 256, 1, 0, 0, 1, "d");
    FX_MG4(8, 8, 8, 4, 256, 1, 0, 0, 1, "d");
    FX_MG4(12, 12, 12, 2, 256, 1, 32, 4, 1, "p4k32d");
    FX_MG4(17, 17, 17, 1, 256, 0, kFold17, 3, 0, "p3fold");
    FX_MG4(8, 8, 16, 4, 256, 1, 0, 0, 0, "");
    // One block per workgroup.
    FX_MG4(17, 17, 17, 1, 256, 0, 0, 0, 0, "");
    FX_MG4(6, 6, 6, 16, 256, 1, 0, 0, 0, "");
    FX_MG4(6, 6, 6, 8, 256, 1, 0, 0, 0, "");
    FX_MG4(8, 8, 8, 8, 256, 1, 0, 0, 0, "");
    FX_MG4(8, 8, 8, 4, 256, 1, 0, 0, 0, "");
    FX_MG4(12, 12, 12, 2, 256, 1, 0, 0, 1, "d");
    FX_MG4(12, 12, 12, 1, 192, 1, 0, 0, 1, "d");
    FX_MG4(17, 17, 17, 1, 256, 0, 0, 0, 1, "d");
    FX_MG4(8, 8, 16, 4, 256, 1, 0, 0, 1, "d");
    // K = 32 on the long axes.
    FX_MG4(12, 12, 12, 2, 256, 1, 32, 0, 1, "k32d");
    FX_MG4(17, 17, 17, 1, 256, 0, 32, 0, 0, "k32");
    FX_MG4(8, 8, 16, 4, 256, 1, 32, 0, 1, "k32d");
    // 17-point fold.
    FX_MG4(17, 17, 17, 1, 256, 0, kFold17, 0, 0, "fold");
    FX_MG4(17, 17, 17, 1, 256, 0, kFold17, 3, 1, "p3dfold");
    // Persistent and pipelined.
    FX_MG4(6, 6, 6, 8, 256, 1, 0, 8, 1, "p8d");
    FX_MG4(8, 8, 8, 4, 256, 1, 0, 8, 1, "p8d");
    FX_MG4(12, 12, 12, 2, 256, 1, 0, 4, 0, "p4");
    FX_MG4(12, 12, 12, 2, 256, 1, 0, 4, 1, "p4d");
    FX_MG4(17, 17, 17, 1, 256, 0, 0, 3, 0, "p3");
    FX_MG4(8, 8, 16, 4, 256, 1, 0, 8, 1, "p8d");
    FX_MG4(8, 8, 16, 4, 256, 1, 32, 8, 1, "p8k32d");
}

}  // namespace fx
