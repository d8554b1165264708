// Device launcher and registry descriptor of the matrix-core variant (mfma.hpp).
#pragma once

namespace fx {

template <class C, int SPLIT, int KW = 0, int PERSIST = 0, int XD = 0>
hipError_t launch_mfma(const void* in, void* out, long nvol, const void* bank, double scl,
                       unsigned*, hipStream_t stream)
{
    if (nvol <= 0) return hipSuccess;
    const long nblocks = (nvol + C::VPW - 1) / C::VPW;
    if (nblocks >= (1L << 31)) return hipErrorInvalidValue;
    if constexpr (PERSIST > 0) {
        const long grid = persistent_grid(k_mfma_pipe<C, SPLIT, KW, PERSIST, XD>, C::NT, PERSIST, nblocks);
        hipLaunchKernelGGL((k_mfma_pipe<C, SPLIT, KW, PERSIST, XD>), dim3((unsigned)grid), dim3(C::NT),
                           0, stream, (const cx<float>*)in, (cx<float>*)out, nvol, nblocks,
                           (const Bank<float>*)bank, (float)scl);
    } else {
        hipLaunchKernelGGL((k_mfma<C, SPLIT, KW, XD>), dim3((unsigned)nblocks), dim3(C::NT), 0, stream,
                           (const cx<float>*)in, (cx<float>*)out, nvol, (const Bank<float>*)bank,
                           (float)scl);
    }
    return hipGetLastError();
}

// The descriptor's `persist` stays 0 for every form: it marks kernels that claim blocks,
// and the persistent matrix-core form walks a fixed stride.
template <class C, int SPLIT, int KW = 0, int PERSIST = 0, int XD = 0>
KernelDesc mfma_desc(const char* name)
{
    return KernelDesc{C::N0, C::N1, C::N2, kC64, C::S, kFused, C::VPW, C::NT, 0, 0,
                      sizeof(cx<float>) * (size_t)C::LDS, name,
                      &launch_mfma<C, SPLIT, KW, PERSIST, XD>, 1 + SPLIT};
}

}  // namespace fx
