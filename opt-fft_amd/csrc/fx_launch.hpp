// Device launchers and registry descriptors for the kernel templates of kernels.hpp.
// Included at the end of kernels.hpp through the include path.
#pragma once

namespace fx {

template <class C>
hipError_t launch_fused(const void* in, void* out, long nvol, const void* bank, double scl,
                        hipStream_t stream)
{
    using T = typename C::T;
    if (nvol <= 0) return hipSuccess;
    const long grid = (nvol + C::VPW - 1) / C::VPW;
    hipLaunchKernelGGL((k_fused<C>), dim3((unsigned)grid), dim3(C::NT), 0, stream,
                       (const cx<T>*)in, (cx<T>*)out, nvol, (const Bank<T>*)bank, (T)scl);
    return hipGetLastError();
}

template <class C>
hipError_t launch_pencil(const void* in, void* out, long nvol, const void* bank, double scl,
                         hipStream_t stream)
{
    using T = typename C::T;
    if (nvol <= 0) return hipSuccess;
    const long chunk = C::CHUNK > 0 ? C::CHUNK : nvol;
    for (long v0 = 0; v0 < nvol; v0 += chunk) {
        const long nv = (nvol - v0) < chunk ? (nvol - v0) : chunk;
        const cx<T>* src = (const cx<T>*)in + v0 * C::V;
        cx<T>* dst = (cx<T>*)out + v0 * C::V;
        const long nslab = nv * C::N0;
        const long ga = (nslab + C::SPW - 1) / C::SPW;
        hipLaunchKernelGGL((k_slab<C>), dim3((unsigned)ga), dim3(C::NT), 0, stream, src, dst,
                           nslab, (const Bank<T>*)bank);
        const long nlines = nv * C::PLANE;
        const long gb = (nlines + C::NTB - 1) / C::NTB;
        hipLaunchKernelGGL((k_xline<C>), dim3((unsigned)gb), dim3(C::NTB), 0, stream, dst,
                           nlines, (const Bank<T>*)bank, (T)scl);
    }
    return hipGetLastError();
}

template <class C>
KernelDesc fused_desc(const char* name)
{
    return KernelDesc{C::N0, C::N1, C::N2, prec_of<typename C::T>(), C::S, kFused, C::VPW,
                      C::NT, C::V17, sizeof(cx<typename C::T>) * (size_t)C::LDS, name,
                      &launch_fused<C>};
}

template <class C>
KernelDesc pencil_desc(const char* name)
{
    return KernelDesc{C::N0, C::N1, C::N2, prec_of<typename C::T>(), C::S, kPencil, C::SPW,
                      C::NT, 0, sizeof(cx<typename C::T>) * (size_t)C::LDS, name,
                      &launch_pencil<C>};
}

}  // namespace fx
