// Mixed-batch task lists and the registry of mixed kernels (host only, no device calls).
#include <algorithm>
#include <vector>

#include "multi_tasks.hpp"

namespace fx {

const std::vector<MultiDesc>& multi_registry()
{
    static const std::vector<MultiDesc> r = [] {
        std::vector<MultiDesc> v;
        register_multi(v);
        return v;
    }();
    return r;
}

int multi_code(int n0, int n1, int n2)
{
    for (int g = 0; g < kMultiGeoms; ++g)
        if (kMultiShape[g][0] == n0 && kMultiShape[g][1] == n1 && kMultiShape[g][2] == n2)
            return g;
    return -1;
}

// Blocks of every item as (item, block) pairs.  Order 0: heaviest first (ties keep item
// order), so the short blocks balance the end of the launch.  Order 1: every item's blocks
// spread evenly over the list (block b of an item with n blocks at position (b + 1/2) / n,
// ties in item order), so the claims hand out the items side by side and a workgroup
// running an arithmetic-bound block (17^3) shares its CU with ones running memory-bound
// blocks instead of all CUs running one kind at a time.
void multi_tasks(const MultiDesc& d, const int* code, const int* batch, int nitems,
                 size_t elem, std::vector<int2>& tasks)
{
    struct T {
        size_t bytes;
        double pos;
        int item, block;
    };
    std::vector<T> all;
    for (int i = 0; i < nitems; ++i) {
        const int g = code[i];
        const size_t vb = (size_t)kMultiShape[g][0] * kMultiShape[g][1] * kMultiShape[g][2] * elem;
        const int vpw = d.vpw[g];
        const int nb = (batch[i] + vpw - 1) / vpw;
        for (int b = 0; b < nb; ++b) {
            const int nv = std::min(vpw, batch[i] - b * vpw);
            all.push_back(T{vb * (size_t)nv, (b + 0.5) / nb, i, b});
        }
    }
    if (d.order == 1)
        std::stable_sort(all.begin(), all.end(), [](const T& a, const T& b) { return a.pos < b.pos; });
    else
        std::stable_sort(all.begin(), all.end(), [](const T& a, const T& b) { return a.bytes > b.bytes; });
    tasks.clear();
    for (const T& t : all) tasks.push_back(make_int2(t.item, t.block));
}

}  // namespace fx
