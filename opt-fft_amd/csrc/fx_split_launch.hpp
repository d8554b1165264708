// Device launcher and registry descriptor of the split-resident kernel (split.hpp).
#pragma once

namespace fx {

template <class C>
hipError_t launch_split(const void* in, void* out, long nvol, const void* bank, double scl,
                        unsigned* work, hipStream_t stream)
{
    using T = typename C::T;
    if (nvol <= 0) return hipSuccess;
    if (nvol >= (1L << 31)) return hipErrorInvalidValue;
    const long grid = C::P > 0 ? persistent_grid(k_split<C>, C::NT, C::P, nvol) : nvol;
    hipLaunchKernelGGL((k_split<C>), dim3((unsigned)grid), dim3(C::NT), 0, stream,
                       (const cx<T>*)in, (cx<T>*)out, nvol, (const Bank<T>*)bank, (T)scl,
                       C::P > 0 ? claim_counter(work, grid, nvol) : nullptr);
    return hipGetLastError();
}

template <class C>
KernelDesc split_desc(const char* name)
{
    return KernelDesc{C::N0, C::N1, C::N2, prec_of<typename C::T>(), C::S, kFused, 1, C::NT,
                      C::V17, C::P, sizeof(cx<typename C::T>) * (size_t)C::LDS, name,
                      &launch_split<C>};
}

}  // namespace fx
