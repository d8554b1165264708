// Device launcher and registry descriptor of the mixed matrix-core launch (multi_mfma.hpp).
#pragma once

namespace fx {

template <int SPLIT>
hipError_t launch_multi_mfma(const void* args, const int2* tasks, int ntasks, const void* bank,
                             unsigned*, hipStream_t stream)
{
    if (ntasks <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_multi_mfma<SPLIT>), dim3((unsigned)ntasks), dim3(MultiMfmaPreset::NT),
                       0, stream, *(const MultiArgs<float>*)args, tasks, ntasks,
                       (const Bank<float>*)bank);
    return hipGetLastError();
}

// One workgroup per task.
inline long multi_mfma_grid(long ntasks) { return ntasks; }

template <int SPLIT>
MultiDesc multi_mfma_desc(const char* name)
{
    using P = MultiMfmaPreset;
    return MultiDesc{kC64, 2 + SPLIT, P::NT, {P::vpw(0), P::vpw(1), P::vpw(2), P::vpw(3), P::vpw(4)},
                     sizeof(cx<float>) * (size_t)P::LDS, name, &launch_multi_mfma<SPLIT>,
                     nullptr, nullptr, 1 + SPLIT, &multi_mfma_grid};
}

}  // namespace fx
