// TEST INFRASTRUCTURE: the mixed matrix-core launch is wave-level (MFMA) and is not emulated;
// its launcher reports "not supported" and the emulator tests skip it (it is checked on the
// GPU against the oracle).
#pragma once

namespace fx {

template <int SPLIT>
hipError_t launch_multi_mfma(const void*, const int2*, int, const void*, unsigned*, hipStream_t)
{
    return hipErrorNotSupported;
}

template <int SPLIT>
MultiDesc multi_mfma_desc(const char* name)
{
    using P = MultiMfmaPreset;
    return MultiDesc{kC64, 2 + SPLIT, P::NT, {P::vpw(0), P::vpw(1), P::vpw(2), P::vpw(3), P::vpw(4)},
                     sizeof(cx<float>) * (size_t)P::LDS, name, &launch_multi_mfma<SPLIT>,
                     nullptr, nullptr, 1 + SPLIT};
}

}  // namespace fx
