// TEST INFRASTRUCTURE: host emulation of the residue-split kernel (xres.hpp) with the same
// phase functions and the same workgroup -> (volume, residue) map; workgroups whose volume
// lies past the batch return before any phase, as on the device.
#pragma once

namespace fx {

template <class C>
hipError_t launch_xres(const void* in, void* out, long nvol, const void* bank, double scl,
                       unsigned*, hipStream_t)
{
    using T = typename C::T;
    using K = XRes<C>;
    if (nvol <= 0) return hipSuccess;
    const Bank<T>* bk = (const Bank<T>*)bank;
    const cx<T>* src = (const cx<T>*)in;
    cx<T>* dst = (cx<T>*)out;
    EmuLds<T> lds(C::LDS);
    const long grid = K::grid(nvol);
    for (long g = 0; g < grid; ++g) {
        long v;
        int h;
        K::task(g, v, h);
        if (v >= nvol) continue;
        lds.poison();
        cx<T>* sm = lds.data();
        emu_threads(C::NT, [&](int t) { K::xpass(t, v, h, src, sm, bk); });
        emu_threads(C::NT, [&](int t) { K::zpass(t, sm, bk); });
        emu_threads(C::NT, [&](int t) { K::ypass(t, v, h, sm, dst, bk, (T)scl); });
        if (!lds.guards_ok() && g_emu_bad_block < 0) g_emu_bad_block = g;
    }
    return hipSuccess;
}

template <class C>
KernelDesc xres_desc(const char* name)
{
    return KernelDesc{C::N0, C::N1, C::N2, prec_of<typename C::T>(), C::S, kFused, 1, C::NT,
                      C::V17, 0, sizeof(cx<typename C::T>) * (size_t)C::LDS, name,
                      &launch_xres<C>};
}

}  // namespace fx
