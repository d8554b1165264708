// Device launcher and registry descriptor for the mixed-batch kernel (multi.hpp).
#pragma once

namespace fx {

template <typename T, int PRESET>
long multi_launch_grid(long ntasks)
{
    return persistent_grid(k_multi<T, PRESET>, MultiCfg<T, PRESET>::NT, 8, ntasks);
}

template <typename T, int PRESET>
hipError_t launch_multi(const void* args, const int2* tasks, int ntasks, const void* bank,
                        unsigned* work, hipStream_t stream)
{
    using M = MultiCfg<T, PRESET>;
    if (ntasks <= 0) return hipSuccess;
    const long grid = multi_launch_grid<T, PRESET>(ntasks);
    hipLaunchKernelGGL((k_multi<T, PRESET>), dim3((unsigned)grid), dim3(M::NT), 0, stream,
                       *(const MultiArgs<T>*)args, tasks, ntasks, (const Bank<T>*)bank,
                       claim_counter(work, grid, ntasks));
    return hipGetLastError();
}

template <typename T, int PRESET>
hipError_t launch_multi_dev(const void* d, const void* bank, unsigned* work, int grid,
                            hipStream_t stream)
{
    using M = MultiCfg<T, PRESET>;
    hipLaunchKernelGGL((k_multi_dev<T, PRESET>), dim3((unsigned)grid), dim3(M::NT), 0, stream,
                       (MultiDescriptor<T>*)d, (const Bank<T>*)bank, work);
    return hipGetLastError();
}

template <typename T, int PRESET>
int multi_dev_grid()
{
    using M = MultiCfg<T, PRESET>;
    return (int)persistent_grid(k_multi_dev<T, PRESET>, M::NT, 8, 1L << 30);
}

template <typename T, int PRESET>
MultiDesc multi_desc(const char* name, int order = 0)
{
    using M = MultiCfg<T, PRESET>;
    return MultiDesc{prec_of<T>(), PRESET, M::NT,
                     {M::vpw(0), M::vpw(1), M::vpw(2), M::vpw(3), M::vpw(4)},
                     sizeof(cx<T>) * (size_t)M::LDS, name, &launch_multi<T, PRESET>,
                     &launch_multi_dev<T, PRESET>, &multi_dev_grid<T, PRESET>, 0,
                     &multi_launch_grid<T, PRESET>, order};
}

}  // namespace fx
