// Dense row-column floor kernels (dense.hpp) for every geometry, both precisions and
// directions.  Registered last, so a plan takes one only when forced by name
// (FFT3D_MX_KERNEL / fft3d_mx_create_named) or when the race finds it fastest.
#include "dense.hpp"

namespace fx {

#define FX_D(N0, N1, N2, T, TAG)                                                             \
    r.push_back(dense_desc<DenseCfg<N0, N1, N2, T, -1>>("dense_" #N0 "x" #N1 "x" #N2 "_" TAG "_fwd_l1")); \
    r.push_back(dense_desc<DenseCfg<N0, N1, N2, T, +1>>("dense_" #N0 "x" #N1 "x" #N2 "_" TAG "_bwd_l1"))

void register_dense(std::vector<KernelDesc>& r)
{
    FX_D(6, 6, 6, double, "c128");
    FX_D(8, 8, 8, double, "c128");
    FX_D(12, 12, 12, double, "c128");
    FX_D(17, 17, 17, double, "c128");
    FX_D(24, 24, 24, double, "c128");
    FX_D(36, 36, 36, double, "c128");
    FX_D(8, 8, 16, double, "c128");
    FX_D(24, 24, 48, double, "c128");
    FX_D(6, 6, 6, float, "c64");
    FX_D(8, 8, 8, float, "c64");
    FX_D(12, 12, 12, float, "c64");
    FX_D(17, 17, 17, float, "c64");
    FX_D(24, 24, 24, float, "c64");
    FX_D(36, 36, 36, float, "c64");
    FX_D(8, 8, 16, float, "c64");
    FX_D(24, 24, 48, float, "c64");
}

}  // namespace fx
