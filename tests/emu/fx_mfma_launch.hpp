// TEST INFRASTRUCTURE: the matrix-core kernels are wave-level (MFMA) and are not emulated;
// their launcher reports "not supported" and the emulator tests skip them (they are checked
// on the GPU against the oracle).
#pragma once

namespace fx {

template <class C, int SPLIT, int KW = 0, int PERSIST = 0, int XD = 0>
hipError_t launch_mfma(const void*, void*, long, const void*, double, unsigned*, hipStream_t)
{
    return hipErrorNotSupported;
}

template <class C, int SPLIT, int KW = 0, int PERSIST = 0, int XD = 0>
KernelDesc mfma_desc(const char* name)
{
    return KernelDesc{C::N0, C::N1, C::N2, kC64, C::S, kFused, C::VPW, C::NT, 0, 0,
                      sizeof(cx<float>) * (size_t)C::LDS, name,
                      &launch_mfma<C, SPLIT, KW, PERSIST, XD>, 1 + SPLIT};
}

}  // namespace fx
