// LDS-resident fused kernels for the small geometries: L = 6, 8, 12 and 8x8x16, both
// precisions and both directions, and L = 24 at complex64, at several tilings each for the create-time race.
// "pipe_pP" variants are the persistent software-pipelined form (at most P workgroups per
// CU); "pipe_p0" is the same pass structure with one block per workgroup; "fused_" is the
// staged form (whole block read into LDS first); "shfl_" keeps an 8^3 volume in one
// wavefront's registers and does the y and z axes with lane shuffles (shfl8.hpp).
#include "kernels.hpp"
#include "shfl8.hpp"

namespace fx {

#define FX_F(N0, N1, N2, T, S, VPW, NT, PZ, NAME) \
    r.push_back(fused_desc<FusedCfg<N0, N1, N2, T, S, VPW, NT, PZ, 0, 0>>(NAME))
#define FX_F_BOTH(N0, N1, N2, T, TAG, VPW, NT, PZ)                                         \
    FX_F(N0, N1, N2, T, -1, VPW, NT, PZ, "fused_" #N0 "x" #N1 "x" #N2 "_" TAG "_fwd_v" #VPW "t" #NT); \
    FX_F(N0, N1, N2, T, +1, VPW, NT, PZ, "fused_" #N0 "x" #N1 "x" #N2 "_" TAG "_bwd_v" #VPW "t" #NT)
#define FX_P(N0, N1, N2, T, S, VPW, NT, PZ, P, NAME) \
    r.push_back(fpipe_desc<FusedCfg<N0, N1, N2, T, S, VPW, NT, PZ, 0, 0>, P>(NAME))
#define FX_P_BOTH(N0, N1, N2, T, TAG, VPW, NT, PZ, P)                                        \
    FX_P(N0, N1, N2, T, -1, VPW, NT, PZ, P,                                                 \
         "pipe_" #N0 "x" #N1 "x" #N2 "_" TAG "_fwd_v" #VPW "t" #NT "p" #P);                  \
    FX_P(N0, N1, N2, T, +1, VPW, NT, PZ, P,                                                 \
         "pipe_" #N0 "x" #N1 "x" #N2 "_" TAG "_bwd_v" #VPW "t" #NT "p" #P)

// Autotuning dimensions beyond the tiling: radix order ("_r1": the line factors in the other
// order, FusedCfg V17 = 1) and LDS padding ("_z0": rows unpadded, PZ = 0).
#define FX_FV(N0, N1, N2, T, S, VPW, NT, PZ, ALG, NAME) \
    r.push_back(fused_desc<FusedCfg<N0, N1, N2, T, S, VPW, NT, PZ, 0, ALG>>(NAME))
#define FX_FV_BOTH(N0, N1, N2, T, TAG, VPW, NT, PZ, ALG, SUF)                                     \
    FX_FV(N0, N1, N2, T, -1, VPW, NT, PZ, ALG, "fused_" #N0 "x" #N1 "x" #N2 "_" TAG "_fwd_v" #VPW "t" #NT SUF); \
    FX_FV(N0, N1, N2, T, +1, VPW, NT, PZ, ALG, "fused_" #N0 "x" #N1 "x" #N2 "_" TAG "_bwd_v" #VPW "t" #NT SUF)
#define FX_PV(N0, N1, N2, T, S, VPW, NT, PZ, P, ALG, NAME) \
    r.push_back(fpipe_desc<FusedCfg<N0, N1, N2, T, S, VPW, NT, PZ, 0, ALG>, P>(NAME))
#define FX_PV_BOTH(N0, N1, N2, T, TAG, VPW, NT, PZ, P, ALG, SUF)                                  \
    FX_PV(N0, N1, N2, T, -1, VPW, NT, PZ, P, ALG,                                                 \
          "pipe_" #N0 "x" #N1 "x" #N2 "_" TAG "_fwd_v" #VPW "t" #NT "p" #P SUF);                  \
    FX_PV(N0, N1, N2, T, +1, VPW, NT, PZ, P, ALG,                                                 \
          "pipe_" #N0 "x" #N1 "x" #N2 "_" TAG "_bwd_v" #VPW "t" #NT "p" #P SUF)

#define FX_S(T, S, W, PERM, X, NAME) r.push_back(shfl8_desc<Shfl8Cfg<T, S, W, PERM, X>>(NAME))
#define FX_S_BOTH(T, TAG, W, PERM, ORD, X, XT)                                       \
    FX_S(T, -1, W, PERM, X, "shfl_8x8x8_" TAG "_fwd_w" #W ORD XT);                   \
    FX_S(T, +1, W, PERM, X, "shfl_8x8x8_" TAG "_bwd_w" #W ORD XT)

void register_fused_small(std::vector<KernelDesc>& r)
{
    // The first variant of a geometry and precision is the measured fastest at the
    // configurations' batches (taken without a race: FFT3D_MX_NO_TUNE, the allocation-free
    // entry point; and kept by the race on near-ties).
    // L = 6: 36 lines per volume per pass.  Staged blocks of 8 volumes on 288 threads (one
    // line per lane) beat 16 on 576 by 9% at complex64 and 3% at complex128 (65536 volumes);
    // at L = 12, complex64, one volume on 192 threads beats the pipelined forms by 15%.
    FX_F_BOTH(6, 6, 6, double, "c128", 8, 288, 1);
    FX_P_BOTH(6, 6, 6, double, "c128", 16, 576, 1, 4);
    FX_P_BOTH(6, 6, 6, double, "c128", 16, 576, 1, 0);
    FX_P_BOTH(6, 6, 6, double, "c128", 32, 576, 1, 4);
    FX_F_BOTH(6, 6, 6, double, "c128", 16, 576, 1);
    FX_F_BOTH(6, 6, 6, float, "c64", 8, 288, 1);
    FX_P_BOTH(6, 6, 6, float, "c64", 32, 576, 1, 4);
    FX_PV_BOTH(6, 6, 6, double, "c128", 16, 576, 1, 4, 1, "_r1");
    FX_P_BOTH(6, 6, 6, float, "c64", 16, 576, 1, 0);
    FX_F_BOTH(6, 6, 6, float, "c64", 16, 576, 1);
    FX_F_BOTH(6, 6, 6, float, "c64", 4, 144, 1);
    // L = 8: 64 lines per volume per pass.
    // 8 volumes x 512 threads first: the fastest from HBM at batch 4096 (the race keeps the
    // earlier-registered variant on near-ties).
    FX_F_BOTH(8, 8, 8, double, "c128", 8, 512, 1);
    FX_P_BOTH(8, 8, 8, double, "c128", 8, 512, 1, 4);
    FX_P_BOTH(8, 8, 8, double, "c128", 8, 512, 1, 0);
    FX_F_BOTH(8, 8, 8, double, "c128", 4, 256, 1);
    FX_P_BOTH(8, 8, 8, double, "c128", 4, 256, 1, 0);
    FX_F_BOTH(8, 8, 8, float, "c64", 4, 256, 1);
    FX_P_BOTH(8, 8, 8, float, "c64", 4, 256, 1, 8);
    FX_P_BOTH(8, 8, 8, float, "c64", 8, 512, 1, 4);
    FX_FV_BOTH(8, 8, 8, double, "c128", 8, 512, 1, 1, "_r1");
    FX_FV_BOTH(8, 8, 8, double, "c128", 8, 512, 0, 0, "_z0");
    FX_FV_BOTH(8, 8, 8, float, "c64", 4, 256, 1, 1, "_r1");
    // Register-resident, shuffles for y and z: W waves (volumes) per workgroup; "n" stores
    // in natural lane order after one more shuffle, "r" from the bit-reversed lanes.
    FX_S_BOTH(double, "c128", 4, 1, "n", 0, "");
    FX_S_BOTH(double, "c128", 4, 0, "r", 0, "");
    FX_S_BOTH(double, "c128", 8, 1, "n", 0, "");
    FX_S_BOTH(float, "c64", 4, 1, "n", 0, "");
    // The same with DPP and permlane-swap moves for the exchanges instead of ds_bpermute
    // (X = 1): 15.5 against 15.0 us at 4096 volumes (the LDS kernel: 10.9 us); the
    // exchanges do not bound this form, its radix-2 butterflies on every lane do.
    FX_S_BOTH(double, "c128", 4, 0, "r", 1, "_dpp");
    // L = 12: 144 lines per volume per pass.
    FX_F_BOTH(12, 12, 12, double, "c128", 1, 192, 1);
    FX_P_BOTH(12, 12, 12, double, "c128", 4, 576, 1, 2);
    FX_P_BOTH(12, 12, 12, double, "c128", 1, 192, 1, 8);
    FX_P_BOTH(12, 12, 12, double, "c128", 4, 576, 1, 0);
    FX_F_BOTH(12, 12, 12, double, "c128", 4, 576, 1);
    FX_F_BOTH(12, 12, 12, float, "c64", 1, 192, 1);
    FX_P_BOTH(12, 12, 12, float, "c64", 4, 576, 1, 4);
    FX_PV_BOTH(12, 12, 12, double, "c128", 4, 576, 1, 2, 1, "_r1");
    FX_PV_BOTH(12, 12, 12, float, "c64", 4, 576, 1, 4, 1, "_r1");
    FX_PV_BOTH(12, 12, 12, float, "c64", 4, 576, 0, 4, 0, "_z0");
    FX_P_BOTH(12, 12, 12, float, "c64", 2, 320, 1, 4);
    FX_F_BOTH(12, 12, 12, float, "c64", 4, 576, 1);
    FX_F_BOTH(12, 12, 12, float, "c64", 2, 384, 1);
    // 8 x 8 x 16: x lines 128 per volume.
    FX_P_BOTH(8, 8, 16, double, "c128", 2, 256, 1, 4);
    FX_P_BOTH(8, 8, 16, double, "c128", 4, 512, 1, 2);
    FX_F_BOTH(8, 8, 16, double, "c128", 2, 256, 1);
    FX_P_BOTH(8, 8, 16, float, "c64", 4, 512, 1, 4);
    FX_PV_BOTH(8, 8, 16, double, "c128", 2, 256, 1, 4, 1, "_r1");
    FX_F_BOTH(8, 8, 16, float, "c64", 4, 512, 1);
    // L = 24 at complex64: one 115 KB volume per workgroup, 576 lines per pass, staged (the
    // pipelined form spills at the 168-VGPR cap of 9 waves per CU; at complex128 the volume
    // exceeds LDS and runs as a pencil).
    FX_F_BOTH(24, 24, 24, float, "c64", 1, 576, 1);
}

}  // namespace fx
