// Host helpers of the mixed-batch launch shared by the plan step and the tests.
#pragma once
#include <stddef.h>

#include <vector>

#include "multi.hpp"

namespace fx {

// Geometry code of (n0, n1, n2) in kMultiShape, or -1.
int multi_code(int n0, int n1, int n2);
// Blocks of every item as (item, block) pairs, in the descriptor's order (heaviest first,
// or the items interleaved).
void multi_tasks(const MultiDesc& d, const int* code, const int* batch, int nitems, size_t elem,
                 std::vector<int2>& tasks);

}  // namespace fx
