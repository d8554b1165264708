// Device launcher and registry descriptor of the residue-split kernel (xres.hpp).
#pragma once

namespace fx {

template <class C>
hipError_t launch_xres(const void* in, void* out, long nvol, const void* bank, double scl,
                       unsigned*, hipStream_t stream)
{
    using T = typename C::T;
    if (nvol <= 0) return hipSuccess;
    const long grid = XRes<C>::grid(nvol);
    if (grid >= (1L << 31)) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_xres<C>), dim3((unsigned)grid), dim3(C::NT), 0, stream,
                       (const cx<T>*)in, (cx<T>*)out, nvol, (const Bank<T>*)bank, (T)scl);
    return hipGetLastError();
}

template <class C>
KernelDesc xres_desc(const char* name)
{
    return KernelDesc{C::N0, C::N1, C::N2, prec_of<typename C::T>(), C::S, kFused, 1, C::NT,
                      C::V17, 0, sizeof(cx<typename C::T>) * (size_t)C::LDS, name,
                      &launch_xres<C>};
}

}  // namespace fx
