// Device launcher and registry descriptor of the balanced 17^3 kernel (f17.hpp).
#pragma once

namespace fx {

template <class C>
hipError_t launch_f17(const void* in, void* out, long nvol, const void* bank, double scl,
                      unsigned* work, hipStream_t stream)
{
    using T = typename C::T;
    if (nvol <= 0) return hipSuccess;
    if (nvol >= (1L << 31)) return hipErrorInvalidValue;
    const long grid = persistent_grid(k_f17<C>, C::NT, C::PERSIST, nvol);
    hipLaunchKernelGGL((k_f17<C>), dim3((unsigned)grid), dim3(C::NT), 0, stream,
                       (const cx<T>*)in, (cx<T>*)out, nvol, (const Bank<T>*)bank, (T)scl,
                       claim_counter(work, grid, nvol));
    return hipGetLastError();
}

template <class C>
KernelDesc f17_desc(const char* name)
{
    return KernelDesc{17, 17, 17, prec_of<typename C::T>(), C::S, kFused, 1, C::NT, 1,
                      C::PERSIST, (size_t)C::LDS_BYTES, name, &launch_f17<C>};
}

}  // namespace fx
