// Device launcher and registry descriptor of the matrix-core variant (mfma.hpp).
#pragma once

namespace fx {

template <class C, int SPLIT, int KW = 0>
hipError_t launch_mfma(const void* in, void* out, long nvol, const void* bank, double scl,
                       unsigned*, hipStream_t stream)
{
    if (nvol <= 0) return hipSuccess;
    const long grid = (nvol + C::VPW - 1) / C::VPW;
    hipLaunchKernelGGL((k_mfma<C, SPLIT, KW>), dim3((unsigned)grid), dim3(C::NT), 0, stream,
                       (const cx<float>*)in, (cx<float>*)out, nvol, (const Bank<float>*)bank,
                       (float)scl);
    return hipGetLastError();
}

template <class C, int SPLIT, int KW = 0>
KernelDesc mfma_desc(const char* name)
{
    return KernelDesc{C::N0, C::N1, C::N2, kC64, C::S, kFused, C::VPW, C::NT, 0, 0,
                      sizeof(cx<float>) * (size_t)C::LDS, name, &launch_mfma<C, SPLIT, KW>,
                      1 + SPLIT};
}

}  // namespace fx
