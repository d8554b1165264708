// TEST INFRASTRUCTURE: host emulation of the wavefront-shuffle 8^3 kernel (shfl8.hpp): the
// 64 lanes of a wave run each phase in turn, and every shuffle reads a snapshot of all
// lanes' registers taken before the phase, as the hardware exchange does.
#pragma once
#include <vector>

namespace fx {

template <class C>
hipError_t launch_shfl8(const void* in, void* out, long nvol, const void* bank, double scl,
                        unsigned*, hipStream_t)
{
    using T = typename C::T;
    using K = Shfl8<C>;
    const Bank<T>* bk = (const Bank<T>*)bank;
    std::vector<cx<T>> r(64 * 8), p(64 * 8);
    for (long v = 0; v < nvol; ++v) {
        const cx<T>* src = (const cx<T>*)in + v * 512;
        cx<T>* dst = (cx<T>*)out + v * 512;
        for (int l = 0; l < 64; ++l) K::load_x(l, src, &r[8 * l], bk);
        for (int s = 0; s < 6; ++s) {
            const int m = s < 3 ? (4 >> s) : (32 >> (s - 3));
            for (int l = 0; l < 64; ++l)
                for (int x = 0; x < 8; ++x) p[8 * l + x] = r[8 * (l ^ m) + x];
            for (int l = 0; l < 64; ++l) K::stage(l, m, s < 3 ? 0 : 3, &r[8 * l], &p[8 * l], bk);
        }
        if (C::PERM) {
            p = r;
            for (int l = 0; l < 64; ++l)
                for (int x = 0; x < 8; ++x) r[8 * l + x] = p[8 * K::natural_src(l) + x];
        }
        for (int l = 0; l < 64; ++l) K::store(l, dst, &r[8 * l], (T)scl);
    }
    return hipSuccess;
}

template <class C>
KernelDesc shfl8_desc(const char* name)
{
    return KernelDesc{8, 8, 8, prec_of<typename C::T>(), C::S, kFused, C::W, C::NT, 0, 0, 0,
                      name, &launch_shfl8<C>};
}

}  // namespace fx
