// Split-resident single-pass kernels (split.hpp): 24^3 complex128, both directions, one
// workgroup of 576 threads per volume (one x-line per thread), two rounds of 12 x-planes
// (115.2 KB of LDS), persistent (one workgroup per CU) or one volume per workgroup.
#include "split.hpp"

namespace fx {

#define FX_SP(T, S, NT, R, P, SP1, SP2, NAME) \
    r.push_back(split_desc<SplitCfg<24, 24, 24, T, S, NT, R, 1, P, SP1, SP2>>(NAME))
#define FX_SP_BOTH(T, TAG, NT, R, P, SP1, SP2)                                            \
    FX_SP(T, -1, NT, R, P, SP1, SP2, "split_24x24x24_" TAG "_fwd_v1t" #NT "r" #R "p" #P); \
    FX_SP(T, +1, NT, R, P, SP1, SP2, "split_24x24x24_" TAG "_bwd_v1t" #NT "r" #R "p" #P)

void register_split(std::vector<KernelDesc>& r)
{
    // Persistent: the next volume's x-line points 0-11 loaded before the last round's z
    // pass, 12-17 before its y pass, 18-23 after it.
    FX_SP_BOTH(double, "c128", 576, 2, 1, 12, 18);
    FX_SP_BOTH(double, "c128", 576, 2, 0, 24, 24);
}

}  // namespace fx
