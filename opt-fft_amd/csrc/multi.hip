// Instantiations of the mixed-batch persistent kernel (multi.hpp): both precisions, two
// tile presets each, raced at create time.
#include "multi.hpp"

namespace fx {

void register_multi(std::vector<MultiDesc>& r)
{
    r.push_back(multi_desc<float, 0>("multi_c64_t320"));
    r.push_back(multi_desc<float, 1>("multi_c64_t640"));
    r.push_back(multi_desc<double, 0>("multi_c128_t320"));
    r.push_back(multi_desc<double, 1>("multi_c128_t640"));
}

}  // namespace fx
