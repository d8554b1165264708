// Two-pass pencil kernels for volumes that exceed LDS at complex128 (L = 24, 36 and
// 24x24x48) and their complex64 counterparts.  Names: s<SPW> slabs per workgroup,
// t<NT> threads, b<NTB> threads of the x-line pass, c<CHUNK> volumes per launch pair
// (0 = whole batch), p<P> persistent pipelined pass A with at most P workgroups per CU,
// "_st" pass A with staggered HBM traffic (PencilCfg ST = 1).  Also the residue-split
// single-pass kernels of the same geometries (xres.hpp): "xres_..._v1r<R>t<NT>".
#include "kernels.hpp"
#include "xres.hpp"

namespace fx {

#define FX_P(N0, N1, N2, T, S, SPW, NT, PZ, NTB, CH, PIPE, NAME) \
    r.push_back(pencil_desc<PencilCfg<N0, N1, N2, T, S, SPW, NT, PZ, NTB, CH, PIPE>>(NAME))
#define FX_P_BOTH(N0, N1, N2, T, TAG, SPW, NT, PZ, NTB, CH, PIPE)                                \
    FX_P(N0, N1, N2, T, -1, SPW, NT, PZ, NTB, CH, PIPE,                                          \
         "pencil_" #N0 "x" #N1 "x" #N2 "_" TAG "_fwd_s" #SPW "t" #NT "b" #NTB "c" #CH "p" #PIPE); \
    FX_P(N0, N1, N2, T, +1, SPW, NT, PZ, NTB, CH, PIPE,                                          \
         "pencil_" #N0 "x" #N1 "x" #N2 "_" TAG "_bwd_s" #SPW "t" #NT "b" #NTB "c" #CH "p" #PIPE)

#define FX_PS(N0, N1, N2, T, S, SPW, NT, PZ, NTB, CH, PIPE, NAME) \
    r.push_back(pencil_desc<PencilCfg<N0, N1, N2, T, S, SPW, NT, PZ, NTB, CH, PIPE, 1>>(NAME))
#define FX_PS_BOTH(N0, N1, N2, T, TAG, SPW, NT, PZ, NTB, CH, PIPE)                                \
    FX_PS(N0, N1, N2, T, -1, SPW, NT, PZ, NTB, CH, PIPE,                                          \
         "pencil_" #N0 "x" #N1 "x" #N2 "_" TAG "_fwd_s" #SPW "t" #NT "b" #NTB "c" #CH "p" #PIPE "_st"); \
    FX_PS(N0, N1, N2, T, +1, SPW, NT, PZ, NTB, CH, PIPE,                                          \
         "pencil_" #N0 "x" #N1 "x" #N2 "_" TAG "_bwd_s" #SPW "t" #NT "b" #NTB "c" #CH "p" #PIPE "_st")

// Radix order ("_r1": the line factors in the other order, PencilCfg ALG = 1).
#define FX_PR(N0, N1, N2, T, S, SPW, NT, PZ, NTB, CH, PIPE, ST, NAME) \
    r.push_back(pencil_desc<PencilCfg<N0, N1, N2, T, S, SPW, NT, PZ, NTB, CH, PIPE, ST, 1>>(NAME))
#define FX_PR_BOTH(N0, N1, N2, T, TAG, SPW, NT, PZ, NTB, CH, PIPE, ST, SUF)                      \
    FX_PR(N0, N1, N2, T, -1, SPW, NT, PZ, NTB, CH, PIPE, ST,                                      \
          "pencil_" #N0 "x" #N1 "x" #N2 "_" TAG "_fwd_s" #SPW "t" #NT "b" #NTB "c" #CH "p" #PIPE SUF "_r1"); \
    FX_PR(N0, N1, N2, T, +1, SPW, NT, PZ, NTB, CH, PIPE, ST,                                      \
          "pencil_" #N0 "x" #N1 "x" #N2 "_" TAG "_bwd_s" #SPW "t" #NT "b" #NTB "c" #CH "p" #PIPE SUF "_r1")

// Residue-split single pass (xres.hpp): "xres_<shape>_<prec>_<dir>_v1r<R>t<NT>", R workgroups
// per volume, each holding N0 / R x-planes in LDS.
#define FX_X(N0, N1, N2, T, S, R, NT, OR, NAME) \
    r.push_back(xres_desc<XResCfg<N0, N1, N2, T, S, R, NT, OR>>(NAME))
#define FX_X_BOTH(N0, N1, N2, T, TAG, R, NT)                                                       \
    FX_X(N0, N1, N2, T, -1, R, NT, 0, "xres_" #N0 "x" #N1 "x" #N2 "_" TAG "_fwd_v1r" #R "t" #NT); \
    FX_X(N0, N1, N2, T, +1, R, NT, 0, "xres_" #N0 "x" #N1 "x" #N2 "_" TAG "_bwd_v1r" #R "t" #NT)
// "o<OR>": the x pass's point order (XResCfg OR: 1 rotated thread blocks, 2 offset starts).
#define FX_XO_BOTH(N0, N1, N2, T, TAG, R, NT, OR)                                                \
    FX_X(N0, N1, N2, T, -1, R, NT, OR,                                                          \
         "xres_" #N0 "x" #N1 "x" #N2 "_" TAG "_fwd_v1r" #R "t" #NT "o" #OR);                   \
    FX_X(N0, N1, N2, T, +1, R, NT, OR,                                                          \
         "xres_" #N0 "x" #N1 "x" #N2 "_" TAG "_bwd_v1r" #R "t" #NT "o" #OR)

// (Pass B with plain stores, PencilCfg XH = 1, measured 8% slower at 36^3 x 128 and 1% at
// 24x24x48 x 64 than with nontemporal ones: not registered.)

void register_pencil(std::vector<KernelDesc>& r)
{
    // L = 36: slab 36x36 = 20.7 KB complex128.  The race keeps an earlier variant unless a
    // later one is >2% faster: at batch 128 s7t256p1 staggered (71 us), then unstaggered
    // (73 us), ahead of s3t128p2 (73-74 us).
    FX_PS_BOTH(36, 36, 36, double, "c128", 7, 256, 1, 128, 0, 1);
    FX_P_BOTH(36, 36, 36, double, "c128", 7, 256, 1, 128, 0, 1);
    FX_P_BOTH(36, 36, 36, double, "c128", 3, 128, 1, 128, 0, 2);
    FX_P_BOTH(36, 36, 36, double, "c128", 3, 128, 1, 128, 0, 0);
    FX_P_BOTH(36, 36, 36, double, "c128", 7, 256, 1, 128, 0, 0);
    FX_PR_BOTH(36, 36, 36, double, "c128", 7, 256, 1, 128, 0, 1, 1, "_st");
    FX_P_BOTH(36, 36, 36, float, "c64", 7, 256, 1, 128, 0, 1);
    FX_P_BOTH(36, 36, 36, float, "c64", 3, 128, 1, 128, 0, 4);
    // L = 24: slab 24x24 = 9.2 KB complex128.
    FX_P_BOTH(24, 24, 24, double, "c128", 8, 192, 1, 128, 0, 2);
    FX_P_BOTH(24, 24, 24, double, "c128", 8, 192, 1, 128, 0, 0);
    FX_P_BOTH(24, 24, 24, double, "c128", 16, 384, 1, 128, 0, 0);
    FX_P_BOTH(24, 24, 24, float, "c64", 16, 384, 1, 128, 0, 2);
    FX_P_BOTH(24, 24, 24, float, "c64", 16, 384, 1, 128, 0, 0);
    // 24 x 24 x 48: slab 24x48 = 18.4 KB complex128; z lines 48-point, y and x 24-point.
    // Two slabs per 128-thread block first (768 blocks at batch 64, ~3 resident per CU, so
    // blocks in different phases overlap): 26.6 us against 27.9 us for four slabs.
    FX_P_BOTH(24, 24, 48, double, "c128", 2, 128, 1, 128, 0, 0);
    FX_P_BOTH(24, 24, 48, double, "c128", 4, 128, 1, 128, 0, 0);
    FX_P_BOTH(24, 24, 48, double, "c128", 8, 256, 1, 128, 0, 0);
    FX_PR_BOTH(24, 24, 48, double, "c128", 2, 128, 1, 128, 0, 0, 0, "");
    FX_PR_BOTH(24, 24, 24, double, "c128", 8, 192, 1, 128, 0, 2, 0, "");
    FX_P_BOTH(24, 24, 48, float, "c64", 8, 384, 1, 128, 0, 2);
    FX_P_BOTH(24, 24, 48, float, "c64", 8, 256, 1, 128, 0, 0);
    // Residue-split single pass (forced by name, rotated buffers, medians of >= 25 ms
    // windows).  36^3 x 128: r6 (127.9 KB, one workgroup per CU) 64.6 us with 512 threads,
    // 65.3 with 256, against 66.9 us for the best pencil; r12 (63.9 KB, two per CU, 12 reads
    // of every element from L2) 84.7 us.  24x24x48 x 64: r4 (112.9 KB) 16.9 us with 384
    // threads, 18.2 with 256, against 22.5 us; r6 / r8 (two per CU) 23.1 / 22.6 us.  24^3 x 512:
    // r3 (76.8 KB, two per CU) 46.1 us with 256 threads, 48.5 with 192; r2 50.2, r4 55.8-56.1,
    // against 53.3 us for the split-resident kernel; at batch 64 r4 / r3 10.4 / 10.5 us
    // against 16.1 us for the pencil.  Thread counts keep every variant within its VGPR
    // budget without scratch (a 48-point complex128 line takes ~200 VGPRs).  Point order
    // of the x pass (one device): 36^3 x 128 57.1 us with offset starts (o2), 57.5 rotated
    // blocks (o1), 65.2 from point 0; 24x24x48 x 64 16.6 (o2) against 17.6 us; 24^3 x 512
    // 48.0 (o1) against 49.8 us; 36^3 complex64 x 64 19.8 (o2) against 20.6 us; 24x24x48
    // complex64 x 256 30.0 (o2) against 29.1 us, not registered.  At 36^3 a
    // workgroup's x pass (reading 746 KB from L2 at ~52 GB/s per CU) takes 14.4 us of its
    // ~20 us; a persistent form computing the next task's x pass into registers during the
    // LDS passes (256 threads, 486 VGPRs) ran at 165 us: at that register pressure its loads
    // were issued one or two at a time, and issuing them together spilled.  A second form
    // (512 threads, 169 VGPRs: two-stage 4 x 9 line passes, loads six at a time, offset
    // starts) ran without scratch but took 69.8 us at x 128 and 251.4 at x 512 against 57.0 /
    // 225.3 us for the one-task form.  Not kept.
    FX_XO_BOTH(36, 36, 36, double, "c128", 6, 512, 2);
    FX_XO_BOTH(36, 36, 36, double, "c128", 6, 512, 1);
    FX_X_BOTH(36, 36, 36, double, "c128", 6, 512);
    FX_X_BOTH(36, 36, 36, double, "c128", 6, 256);
    FX_XO_BOTH(24, 24, 48, double, "c128", 4, 384, 2);
    FX_X_BOTH(24, 24, 48, double, "c128", 4, 384);
    FX_X_BOTH(24, 24, 48, double, "c128", 4, 256);
    FX_XO_BOTH(24, 24, 24, double, "c128", 3, 256, 1);
    FX_X_BOTH(24, 24, 24, double, "c128", 3, 256);
    FX_X_BOTH(24, 24, 24, double, "c128", 3, 192);
    FX_X_BOTH(24, 24, 24, double, "c128", 4, 192);
    // complex64: 24x24x48 x 256 r2 (12 planes, 112.9 KB) 28.8 us and r3 31.6 us against 42.0 us
    // for the best pencil (x 64: 12.6 / 14.0 against 17.1 us); 36^3 x 512 r3 125.7 us against
    // 94.0 us for the split-resident kernel (r4 161.3, r6 134.6), but at x 64 r3 20.5 against
    // 21.6 us for the pencil.
    FX_X_BOTH(24, 24, 48, float, "c64", 2, 384);
    FX_X_BOTH(24, 24, 48, float, "c64", 3, 256);
    FX_XO_BOTH(36, 36, 36, float, "c64", 3, 512, 2);
    FX_X_BOTH(36, 36, 36, float, "c64", 3, 512);
}

}  // namespace fx
