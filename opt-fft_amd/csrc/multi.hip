// Instantiations of the mixed-batch persistent kernel (multi.hpp): both precisions, several
// tile presets, raced at create time; and the complex64 matrix-core form
// (multi_mfma.hpp), selected by the plan flags FFT3D_MX_MFMA_BF16 / FFT3D_MX_MFMA_FP16.
#include "multi.hpp"
#include "multi_mfma.hpp"

namespace fx {

void register_multi(std::vector<MultiDesc>& r)
{
    r.push_back(multi_desc<float, 0>("multi_c64_t320"));
    r.push_back(multi_desc<float, 1>("multi_c64_t640"));
    r.push_back(multi_desc<float, 2>("multi_c64_t640_fill"));
    r.push_back(multi_desc<float, 2>("multi_c64_t640_fill_mix", 1));
    r.push_back(multi_desc<float, 1>("multi_c64_t640_mix", 1));
    r.push_back(multi_desc<float, 3>("multi_c64_t320_fill"));
    r.push_back(multi_desc<float, 3>("multi_c64_t320_fill_mix", 1));
    r.push_back(multi_desc<double, 0>("multi_c128_t320"));
    r.push_back(multi_desc<double, 1>("multi_c128_t640"));
    r.push_back(multi_desc<double, 1>("multi_c128_t640_mix", 1));
    r.push_back(multi_desc<double, 3>("multi_c128_t320_fill"));
    r.push_back(multi_desc<double, 3>("multi_c128_t320_fill_mix", 1));
    r.push_back(multi_mfma_desc<0>("multi_c64_mfma_t256_bf16x3"));
    r.push_back(multi_mfma_desc<1>("multi_c64_mfma_t256_fp16x3"));
}

}  // namespace fx
