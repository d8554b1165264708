// The kernel registry: every compiled variant, in a fixed order (the first variant of a
// geometry is the one taken when the create-time race is skipped).
#include "registry.hpp"

namespace fx {

const std::vector<KernelDesc>& registry()
{
    static const std::vector<KernelDesc> r = [] {
        std::vector<KernelDesc> v;
        register_fused_small(v);
        register_fused_17(v);
        register_pencil(v);
        register_split(v);
        register_mfma(v);
        register_f17(v);
        register_dense(v);
        return v;
    }();
    return r;
}

}  // namespace fx
