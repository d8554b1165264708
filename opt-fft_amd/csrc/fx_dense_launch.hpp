// Device launcher and registry descriptor of the dense floor (dense.hpp).
#pragma once

namespace fx {

template <class C>
hipError_t launch_dense(const void* in, void* out, long nvol, const void* bank, double scl,
                        unsigned*, hipStream_t stream)
{
    using T = typename C::T;
    if (nvol <= 0) return hipSuccess;
    const Bank<T>* bk = (const Bank<T>*)bank;
    auto grid = [&](int n) { return (nvol * (C::V / n) + C::NT - 1) / C::NT; };
    if (grid(C::N0) >= (1L << 31) || grid(C::N1) >= (1L << 31) || grid(C::N2) >= (1L << 31))
        return hipErrorInvalidValue;
    // z (in -> out), y and x in place on out; the scale goes with the last pass.
    hipLaunchKernelGGL((k_dense<C, 2>), dim3((unsigned)grid(C::N2)), dim3(C::NT), 0, stream,
                       (const cx<T>*)in, (cx<T>*)out, nvol, bk, T(1));
    hipLaunchKernelGGL((k_dense<C, 1>), dim3((unsigned)grid(C::N1)), dim3(C::NT), 0, stream,
                       (const cx<T>*)out, (cx<T>*)out, nvol, bk, T(1));
    hipLaunchKernelGGL((k_dense<C, 0>), dim3((unsigned)grid(C::N0)), dim3(C::NT), 0, stream,
                       (const cx<T>*)out, (cx<T>*)out, nvol, bk, (T)scl);
    return hipGetLastError();
}

template <class C>
KernelDesc dense_desc(const char* name)
{
    return KernelDesc{C::N0, C::N1, C::N2, prec_of<typename C::T>(), C::S, kFused, 1, C::NT, 0,
                      0, 0, name, &launch_dense<C>};
}

}  // namespace fx
