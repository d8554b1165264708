// Device launcher and registry descriptor of the wavefront-shuffle 8^3 kernel (shfl8.hpp).
#pragma once

namespace fx {

template <class C>
hipError_t launch_shfl8(const void* in, void* out, long nvol, const void* bank, double scl,
                        unsigned*, hipStream_t stream)
{
    using T = typename C::T;
    if (nvol <= 0) return hipSuccess;
    const long grid = (nvol + C::W - 1) / C::W;
    if (grid >= (1L << 31)) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_shfl8<C>), dim3((unsigned)grid), dim3(C::NT), 0, stream,
                       (const cx<T>*)in, (cx<T>*)out, nvol, (const Bank<T>*)bank, (T)scl);
    return hipGetLastError();
}

template <class C>
KernelDesc shfl8_desc(const char* name)
{
    return KernelDesc{8, 8, 8, prec_of<typename C::T>(), C::S, kFused, C::W, C::NT, 0, 0, 0,
                      name, &launch_shfl8<C>};
}

}  // namespace fx
