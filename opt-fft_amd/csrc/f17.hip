// Balanced 17^3 kernels (f17.hpp): complex128 at two workgroups per CU, complex64 at four,
// both directions.
#include "f17.hpp"

namespace fx {

#define FX_B(T, TAG, S, DIR, P) \
    r.push_back(f17_desc<F17Cfg<T, S, P>>("bal_17x17x17_" TAG "_" DIR "_v1t256p" #P "_fold"))

void register_f17(std::vector<KernelDesc>& r)
{
    FX_B(double, "c128", -1, "fwd", 2);
    FX_B(double, "c128", +1, "bwd", 2);
    FX_B(double, "c128", -1, "fwd", 1);
    FX_B(double, "c128", +1, "bwd", 1);
    FX_B(float, "c64", -1, "fwd", 4);
    FX_B(float, "c64", +1, "bwd", 4);
    FX_B(float, "c64", -1, "fwd", 2);
    FX_B(float, "c64", +1, "bwd", 2);
}

}  // namespace fx
