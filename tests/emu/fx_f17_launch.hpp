// TEST INFRASTRUCTURE: host emulation of the balanced 17^3 kernel (f17.hpp) with the same
// phase functions, a grid of kEmuPersistentGrid workgroups that draw volumes from the
// launch's claim counter in round-robin order, and split lines emulated as the device runs
// them: every lane of a wave reads its line before any lane writes it.
#pragma once
#include <type_traits>
#include <vector>

namespace fx {

template <class C>
hipError_t launch_f17(const void* in, void* out, long nvol, const void* bank, double scl,
                      unsigned* work, hipStream_t)
{
    using T = typename C::T;
    using K = F17<C>;
    if (nvol <= 0) return hipSuccess;
    const long grid = nvol < kEmuPersistentGrid ? nvol : kEmuPersistentGrid;
    const Bank<T>* bk = (const Bank<T>*)bank;
    const cx<T>* src = (const cx<T>*)in;
    cx<T>* dst = (cx<T>*)out;
    EmuLds<T> lds(C::LDS);
    std::vector<T> pct(C::PCT);
    for (int t = 0; t < C::NT; ++t) K::fill_pct(t, bk, pct.data());
    std::vector<std::vector<typename K::Regs>> R(grid, std::vector<typename K::Regs>(C::NT));
    std::vector<typename K::Pair> P(C::NT);
    auto pass = [&](auto d, cx<T>* sm, cx<T>* o) {
        constexpr int D = decltype(d)::value;
        emu_threads(C::NT, [&](int t) { K::template main_task<D>(t, sm, bk, o, (T)scl); });
        for (int round = 0; round < 2; ++round) {
            emu_threads(C::NT, [&](int t) { K::template pair_read<D>(t, round, sm, pct.data(), P[t]); });
            emu_threads(C::NT, [&](int t) { K::template pair_write<D>(P[t], sm, o, (T)scl); });
        }
    };
    emu_persistent(
        nvol, grid, nvol >= 2 * grid ? work : nullptr,
        [&](long w, long b) { emu_threads(C::NT, [&](int t) { K::load(t, b, src, R[w][t]); }); },
        [&](long w, long b, long nb) {
            lds.poison();
            cx<T>* sm = lds.data();
            emu_threads(C::NT, [&](int t) { K::stage(t, R[w][t], sm); });
            if (nb < nvol) emu_threads(C::NT, [&](int t) { K::load(t, nb, src, R[w][t]); });
            pass(std::integral_constant<int, 0>{}, sm, nullptr);
            pass(std::integral_constant<int, 1>{}, sm, nullptr);
            pass(std::integral_constant<int, 2>{}, sm, dst + b * (long)C::V);
            if (!lds.guards_ok() && g_emu_bad_block < 0) g_emu_bad_block = b;
        });
    return hipSuccess;
}

template <class C>
KernelDesc f17_desc(const char* name)
{
    return KernelDesc{17, 17, 17, prec_of<typename C::T>(), C::S, kFused, 1, C::NT, 1,
                      C::PERSIST, (size_t)C::LDS_BYTES, name, &launch_f17<C>};
}

}  // namespace fx
