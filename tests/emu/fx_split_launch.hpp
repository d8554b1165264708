// TEST INFRASTRUCTURE: host emulation of the split-resident kernel (split.hpp) with the same
// phase functions: a grid of kEmuPersistentGrid workgroups (persistent form) drawing volumes
// from the launch's claim queues, or one workgroup per volume; each workgroup keeps its own
// registers, the next volume's x-lines are loaded after the last round's LDS write, and the
// LDS is re-poisoned before every round.
#pragma once
#include <type_traits>
#include <utility>
#include <vector>

namespace fx {

// f(std::integral_constant<int, RD>, RD == R - 1) for RD = 0 .. R-1 in order.
template <int R, typename F, int... RD>
inline void emu_rounds(F& f, std::integer_sequence<int, RD...>)
{
    (f(std::integral_constant<int, RD>{}, RD == R - 1), ...);
}

template <class C>
hipError_t launch_split(const void* in, void* out, long nvol, const void* bank, double scl,
                        unsigned* work, hipStream_t)
{
    using T = typename C::T;
    using K = Split<C>;
    struct Box {
        typename K::Regs r;
    };
    if (nvol <= 0) return hipSuccess;
    const long grid = C::P > 0 ? (nvol < kEmuPersistentGrid ? nvol : kEmuPersistentGrid) : nvol;
    const Bank<T>* bk = (const Bank<T>*)bank;
    const cx<T>* src = (const cx<T>*)in;
    cx<T>* dst = (cx<T>*)out;
    EmuLds<T> lds(C::LDS);
    std::vector<std::vector<Box>> R(grid, std::vector<Box>(C::NT));
    emu_persistent(
        nvol, grid, C::P > 0 && nvol >= 2 * grid ? work : nullptr,
        [&](long w, long b) { emu_threads(C::NT, [&](int t) { K::load(t, b, src, R[w][t].r); }); },
        [&](long w, long b, long nb) {
            lds.poison();
            emu_threads(C::NT, [&](int t) { K::xdft_put0(t, R[w][t].r, lds.data(), bk); });
            auto round = [&](auto rd, bool last) {
                constexpr int RD = decltype(rd)::value;
                cx<T>* sm = lds.data();
                const bool pre = last && C::P > 0 && nb < nvol;
                if constexpr (RD > 0) {
                    lds.poison();
                    emu_threads(C::NT, [&](int t) { K::template put<RD>(t, R[w][t].r, sm); });
                }
                if (pre)
                    emu_threads(C::NT, [&](int t) { K::template load<0, C::SP1>(t, nb, src, R[w][t].r); });
                emu_threads(C::NT, [&](int t) { K::zpass(t, sm, bk); });
                if (pre)
                    emu_threads(C::NT, [&](int t) { K::template load<C::SP1, C::SP2>(t, nb, src, R[w][t].r); });
                emu_threads(C::NT, [&](int t) { K::ypass(t, RD, b, sm, dst, bk, (T)scl); });
                if (pre)
                    emu_threads(C::NT, [&](int t) { K::template load<C::SP2, C::N0>(t, nb, src, R[w][t].r); });
                if (!lds.guards_ok() && g_emu_bad_block < 0) g_emu_bad_block = b;
            };
            emu_rounds<C::R>(round, std::make_integer_sequence<int, C::R>{});
        });
    return hipSuccess;
}

template <class C>
KernelDesc split_desc(const char* name)
{
    return KernelDesc{C::N0, C::N1, C::N2, prec_of<typename C::T>(), C::S, kFused, 1, C::NT,
                      C::V17, C::P, sizeof(cx<typename C::T>) * (size_t)C::LDS, name,
                      &launch_split<C>};
}

}  // namespace fx
