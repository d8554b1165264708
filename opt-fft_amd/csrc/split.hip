// Split-resident single-pass kernels (split.hpp): 24^3 complex128 and 36^3 complex64, both
// directions, one workgroup per volume, persistent (one workgroup per CU) or one volume per
// workgroup.
#include "split.hpp"

namespace fx {

#define FX_SP(L, T, S, NT, R, P, SP1, SP2, NAME) \
    r.push_back(split_desc<SplitCfg<L, L, L, T, S, NT, R, 1, P, SP1, SP2>>(NAME))
#define FX_SP_BOTH(L, T, TAG, NT, R, P, SP1, SP2)                                              \
    FX_SP(L, T, -1, NT, R, P, SP1, SP2,                                                       \
          "split_" #L "x" #L "x" #L "_" TAG "_fwd_v1t" #NT "r" #R "p" #P);                    \
    FX_SP(L, T, +1, NT, R, P, SP1, SP2,                                                       \
          "split_" #L "x" #L "x" #L "_" TAG "_bwd_v1t" #NT "r" #R "p" #P)

void register_split(std::vector<KernelDesc>& r)
{
    // 24^3 complex128 (221 KB): one x-line per thread, two rounds of 12 x-planes (115.2 KB of
    // LDS).  Persistent: the next volume's x-line points 0-11 loaded before the last round's
    // z pass, 12-17 before its y pass, 18-23 after it.
    FX_SP_BOTH(24, double, "c128", 576, 2, 1, 12, 18);
    FX_SP_BOTH(24, double, "c128", 576, 2, 0, 24, 24);
    // 36^3 complex64 (373 KB): two x-lines per thread, three rounds of 12 x-planes (127.9 KB
    // of LDS).
    FX_SP_BOTH(36, float, "c64", 648, 3, 1, 4, 8);
    FX_SP_BOTH(36, float, "c64", 648, 3, 0, 36, 36);
}

}  // namespace fx
