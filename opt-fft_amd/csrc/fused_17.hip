// LDS-resident fused kernels for L = 17: one 78,608 B complex128 volume per workgroup (two
// workgroups per CU in 160 KB of LDS), or two complex64 volumes.  Odd extents make every
// strided line read conflict-free without padding.  Variants: Rader (0) or folded dense (1)
// line algorithm, workgroup size, and the pipelined persistent forms (see fused_small.hip),
// one or two volumes per workgroup.
#include "kernels.hpp"

namespace fx {

#define FX_17(T, S, VPW, NT, ALG, NAME) \
    r.push_back(fused_desc<FusedCfg<17, 17, 17, T, S, VPW, NT, 0, 0, ALG>>(NAME))
#define FX_17_BOTH(T, TAG, VPW, NT, ALG, ATAG)                                        \
    FX_17(T, -1, VPW, NT, ALG, "fused_17x17x17_" TAG "_fwd_v" #VPW "t" #NT "_" ATAG); \
    FX_17(T, +1, VPW, NT, ALG, "fused_17x17x17_" TAG "_bwd_v" #VPW "t" #NT "_" ATAG)
#define FX_17P(T, S, VPW, NT, ALG, P, NAME) \
    r.push_back(fpipe_desc<FusedCfg<17, 17, 17, T, S, VPW, NT, 0, 0, ALG>, P>(NAME))
#define FX_17P_BOTH(T, TAG, VPW, NT, ALG, ATAG, P)                                                \
    FX_17P(T, -1, VPW, NT, ALG, P, "pipe_17x17x17_" TAG "_fwd_v" #VPW "t" #NT "p" #P "_" ATAG); \
    FX_17P(T, +1, VPW, NT, ALG, P, "pipe_17x17x17_" TAG "_bwd_v" #VPW "t" #NT "p" #P "_" ATAG)
// (One volume per 320-thread workgroup with two pipelined workgroups per CU, P = 2, x-lines
// in three parts, 161 VGPRs: 40.6 us at batch 1024 against 32.5 us for the first variant
// below on the same device, with claims spread over queues; not registered.)
// Staggered HBM traffic (PipeCfg ST = 1): suffix "_st"; ST = 2, the next block's x-lines
// issued in three parts: suffix "_st3".  HN = streaming hints (1 loads, 2 stores; the
// pipelined y pass never hints its 17-element store runs).  Two volumes per 640-thread
// workgroup with hinted loads, in three parts (161 VGPRs, no spill): 36.3-37.5 us at batch
// 1024 against 39.2-40.2 us for one volume per 320-thread workgroup, staggered, on the same
// device.
#define FX_17S(T, S, VPW, NT, ALG, P, ST, HN, NAME) \
    r.push_back(fpipe_desc<FusedCfg<17, 17, 17, T, S, VPW, NT, 0, 0, ALG>, P, ST, HN>(NAME))
#define FX_17S_BOTH(T, TAG, VPW, NT, ALG, ATAG, P, ST, HN, SFX)                                \
    FX_17S(T, -1, VPW, NT, ALG, P, ST, HN,                                                     \
           "pipe_17x17x17_" TAG "_fwd_v" #VPW "t" #NT "p" #P "_" ATAG SFX);                    \
    FX_17S(T, +1, VPW, NT, ALG, P, ST, HN,                                                     \
           "pipe_17x17x17_" TAG "_bwd_v" #VPW "t" #NT "p" #P "_" ATAG SFX)

void register_fused_17(std::vector<KernelDesc>& r)
{
    // Fastest from HBM at batch 1024 first (the race keeps the earlier variant on near-ties).
    FX_17S_BOTH(double, "c128", 2, 640, 1, "fold", 1, 2, 1, "_st3");
    FX_17S_BOTH(double, "c128", 1, 320, 1, "fold", 1, 1, 0, "_st");
    FX_17P_BOTH(double, "c128", 1, 320, 1, "fold", 1);
    FX_17P_BOTH(double, "c128", 2, 640, 1, "fold", 1);
    FX_17P_BOTH(double, "c128", 1, 320, 1, "fold", 0);
    FX_17_BOTH(double, "c128", 1, 320, 1, "fold");
    FX_17_BOTH(double, "c128", 1, 320, 0, "rader");
    FX_17_BOTH(double, "c128", 1, 192, 1, "fold");
    // (complex64 with the three-part prefetch, one volume P = 4 or two volumes P = 2, and
    // three volumes per 960-thread workgroup, one per CU, with or without it (38.8 / 41.1
    // us): never faster than the first form below at 17^3 x 2048, 38.0 us; not registered.)
    FX_17P_BOTH(float, "c64", 1, 320, 1, "fold", 4);
    FX_17S_BOTH(float, "c64", 1, 320, 1, "fold", 2, 1, 0, "_st");
    FX_17P_BOTH(float, "c64", 2, 640, 1, "fold", 2);
    FX_17_BOTH(float, "c64", 2, 320, 1, "fold");
    FX_17_BOTH(float, "c64", 1, 320, 1, "fold");
}

}  // namespace fx
