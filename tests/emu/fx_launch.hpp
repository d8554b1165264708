// TEST INFRASTRUCTURE: host emulation of the kernel launchers.
//
// Found on the include path before opt-fft_amd/csrc when the emulator library is built, so
// the kernel translation units register these host loops instead of device launches.  Each
// block runs its phases in barrier order; within a phase the threads run one after another
// in ascending order, or descending when fx::g_emu_reverse is set -- a phase whose result
// depends on that order has an intra-phase race.  LDS starts out NaN-filled (any read of a
// slot no thread wrote poisons the output) and is flanked by guard elements that must
// survive every block (an out-of-range LDS write is reported as -1000 - block).
#pragma once
#include <cmath>
#include <cstring>
#include <vector>

namespace fx {

extern int g_emu_reverse;
extern long g_emu_bad_block;

template <typename T>
struct EmuLds {
    static constexpr int kGuard = 64;
    std::vector<cx<T>> buf;
    explicit EmuLds(int n) : buf(n + 2 * kGuard) {}
    cx<T>* data() { return buf.data() + kGuard; }
    void poison()
    {
        for (auto& e : buf) e = {T(NAN), T(NAN)};
        for (int g = 0; g < kGuard; ++g) {
            buf[g] = {T(-7777), T(7777)};
            buf[buf.size() - 1 - g] = {T(-7777), T(7777)};
        }
    }
    bool guards_ok() const
    {
        for (int g = 0; g < kGuard; ++g) {
            const cx<T>& a = buf[g];
            const cx<T>& b = buf[buf.size() - 1 - g];
            if (a.r != T(-7777) || a.i != T(7777) || b.r != T(-7777) || b.i != T(7777))
                return false;
        }
        return true;
    }
};

template <typename F>
inline void emu_threads(int nt, F f)
{
    if (g_emu_reverse)
        for (int t = nt - 1; t >= 0; --t) f(t);
    else
        for (int t = 0; t < nt; ++t) f(t);
}

// Host mirror of the device claim_block (kernels.hpp, same queue arithmetic): persistent
// workgroup w at block b draws its next block from its queue's counter; the holder of the
// queue's last ticket resets it.
inline long emu_claim(unsigned* work, long nblocks, long grid, long w, long b)
{
    if (!work) return b + grid;
    const ClaimQueue c = claim_queue(w, grid, nblocks);
    unsigned* ctr = work + c.q * kClaimLine;
    const unsigned t = (*ctr)++;
    if ((long)t == c.nq - 1) *ctr = 0u;
    return claim_queue_block(c, (long)t);
}

// Emulated persistent launch: `grid` workgroups, each starting at its own index, advanced
// one iteration at a time in round-robin order (or reversed), so claims interleave as they
// do on the device.  `start(w, b)` loads the first block of workgroup w, `iter(w, b, nb)`
// runs one iteration on block b while prefetching nb (if nb < nblocks).
template <typename Start, typename Iter>
inline void emu_persistent(long nblocks, long grid, unsigned* work, Start start, Iter iter)
{
    std::vector<long> cur(grid);
    std::vector<char> live(grid, 1);
    for (long w = 0; w < grid; ++w) {
        cur[w] = w;
        start(w, w);
    }
    long left = grid;
    while (left > 0)
        for (long i = 0; i < grid; ++i) {
            const long w = g_emu_reverse ? grid - 1 - i : i;
            if (!live[w]) continue;
            const long nb = emu_claim(work, nblocks, grid, w, cur[w]);
            iter(w, cur[w], nb);
            if (nb >= nblocks) {
                live[w] = 0;
                --left;
            } else {
                cur[w] = nb;
            }
        }
}

template <class C>
hipError_t launch_fused(const void* in, void* out, long nvol, const void* bank, double scl,
                        unsigned*, hipStream_t)
{
    using T = typename C::T;
    using K = Fused<C>;
    if (nvol <= 0) return hipSuccess;
    const long grid = (nvol + C::VPW - 1) / C::VPW;
    const Bank<T>* bk = (const Bank<T>*)bank;
    EmuLds<T> lds(C::LDS);
    for (long b = 0; b < grid; ++b) {
        lds.poison();
        cx<T>* sm = lds.data();
        emu_threads(C::NT, [&](int t) { K::stage(t, b, nvol, (const cx<T>*)in, sm); });
        emu_threads(C::NT, [&](int t) { K::zpass(t, b, nvol, sm, bk); });
        emu_threads(C::NT, [&](int t) { K::ypass(t, b, nvol, sm, bk); });
        emu_threads(C::NT, [&](int t) { K::xpass(t, b, nvol, sm, (cx<T>*)out, bk, (T)scl); });
        if (!lds.guards_ok() && g_emu_bad_block < 0) g_emu_bad_block = b;
    }
    return hipSuccess;
}

// Persistent kernels are emulated with a grid of at most 3 workgroups, so every workgroup
// walks several blocks; LDS is re-poisoned at each block (a block must not depend on the
// previous block's LDS contents).
constexpr long kEmuPersistentGrid = 3;

template <class C, int PERSIST, int ST = 0, int HN = 3>
hipError_t launch_fpipe(const void* in, void* out, long nvol, const void* bank, double scl,
                        unsigned* work, hipStream_t)
{
    using T = typename C::T;
    using K = FusedPipe<C>;
    struct Box { typename K::Regs r; };
    if (nvol <= 0) return hipSuccess;
    const long nblocks = (nvol + C::VPW - 1) / C::VPW;
    const long grid = PERSIST > 0 ? (nblocks < kEmuPersistentGrid ? nblocks : kEmuPersistentGrid)
                                  : nblocks;
    const Bank<T>* bk = (const Bank<T>*)bank;
    const cx<T>* src = (const cx<T>*)in;
    cx<T>* dst = (cx<T>*)out;
    EmuLds<T> lds(C::LDS);
    std::vector<std::vector<Box>> R(grid, std::vector<Box>(C::NT));
    emu_persistent(
        nblocks, grid, PERSIST > 0 && nblocks >= 2 * grid ? work : nullptr,
        [&](long w, long b) { emu_threads(C::NT, [&](int t) { K::load_x(t, b, nvol, src, R[w][t].r); }); },
        [&](long w, long b, long nb) {
            lds.poison();
            cx<T>* sm = lds.data();
            constexpr int SPLIT = PipeCfg<C, PERSIST, ST>::SPLIT;
            constexpr int SPLIT2 = PipeCfg<C, PERSIST, ST>::SPLIT2;
            emu_threads(C::NT, [&](int t) { K::xpass(t, b, nvol, R[w][t].r, sm, bk); });
            if (nb < nblocks)
                emu_threads(C::NT, [&](int t) {
                    K::template load_x<0, SPLIT>(t, nb, nvol, src, R[w][t].r);
                });
            emu_threads(C::NT, [&](int t) { K::zpass(t, b, nvol, sm, bk); });
            if (SPLIT < SPLIT2 && nb < nblocks)
                emu_threads(C::NT, [&](int t) {
                    K::template load_x<SPLIT, SPLIT2>(t, nb, nvol, src, R[w][t].r);
                });
            emu_threads(C::NT, [&](int t) { K::ypass(t, b, nvol, sm, dst, bk, (T)scl); });
            if (SPLIT2 < C::N0 && nb < nblocks)
                emu_threads(C::NT, [&](int t) {
                    K::template load_x<SPLIT2, C::N0>(t, nb, nvol, src, R[w][t].r);
                });
            if (!lds.guards_ok() && g_emu_bad_block < 0) g_emu_bad_block = b;
        });
    return hipSuccess;
}

template <class C>
void emu_spipe(const cx<typename C::T>* src, cx<typename C::T>* dst, long nslab,
               const Bank<typename C::T>* bk, unsigned* work)
{
    using T = typename C::T;
    using K = SlabPipe<C>;
    struct Box { typename K::Regs r; };
    const long nblocks = (nslab + C::SPW - 1) / C::SPW;
    const long grid = nblocks < kEmuPersistentGrid ? nblocks : kEmuPersistentGrid;
    EmuLds<T> lds(C::LDS);
    std::vector<std::vector<Box>> R(grid, std::vector<Box>(C::NT));
    emu_persistent(
        nblocks, grid, nblocks >= 2 * grid ? work : nullptr,
        [&](long w, long b) { emu_threads(C::NT, [&](int t) { K::load_y(t, b, nslab, src, R[w][t].r); }); },
        [&](long w, long b, long nb) {
            lds.poison();
            cx<T>* sm = lds.data();
            emu_threads(C::NT, [&](int t) { K::ypass(t, b, nslab, R[w][t].r, sm, bk); });
            if (nb < nblocks)
                emu_threads(C::NT, [&](int t) {
                    K::template load_y<0, C::SPLIT>(t, nb, nslab, src, R[w][t].r);
                });
            emu_threads(C::NT, [&](int t) { K::zpass(t, b, nslab, sm, bk); });
            if (C::SPLIT < C::N1 && nb < nblocks)
                emu_threads(C::NT, [&](int t) {
                    K::template load_y<C::SPLIT, C::N1>(t, nb, nslab, src, R[w][t].r);
                });
            emu_threads(C::NT, [&](int t) { K::store(t, b, nslab, sm, dst); });
            if (!lds.guards_ok() && g_emu_bad_block < 0) g_emu_bad_block = b;
        });
}

template <class C>
hipError_t launch_pencil(const void* in, void* out, long nvol, const void* bank, double scl,
                         unsigned* work, hipStream_t)
{
    using T = typename C::T;
    using K = Pencil<C>;
    if (nvol <= 0) return hipSuccess;
    const Bank<T>* bk = (const Bank<T>*)bank;
    EmuLds<T> lds(C::LDS);
    const long chunk = C::CHUNK > 0 ? C::CHUNK : nvol;
    for (long v0 = 0; v0 < nvol; v0 += chunk) {
        const long nv = (nvol - v0) < chunk ? (nvol - v0) : chunk;
        const cx<T>* src = (const cx<T>*)in + v0 * C::V;
        cx<T>* dst = (cx<T>*)out + v0 * C::V;
        const long nslab = nv * C::N0;
        const long ga = (nslab + C::SPW - 1) / C::SPW;
        if constexpr (C::PIPE > 0) emu_spipe<C>(src, dst, nslab, bk, work);
        else for (long b = 0; b < ga; ++b) {
            lds.poison();
            cx<T>* sm = lds.data();
            emu_threads(C::NT, [&](int t) { K::stage(t, b, nslab, src, sm); });
            emu_threads(C::NT, [&](int t) { K::zpass(t, b, nslab, sm, bk); });
            emu_threads(C::NT, [&](int t) { K::ypass(t, b, nslab, sm, dst, bk); });
            if (!lds.guards_ok() && g_emu_bad_block < 0) g_emu_bad_block = b;
        }
        const long nlines = nv * C::PLANE;
        const long gb = (nlines + C::NTB - 1) / C::NTB;
        for (long b = 0; b < gb; ++b)
            emu_threads(C::NTB, [&](int t) { K::xline(b * C::NTB + t, nlines, dst, bk, (T)scl); });
    }
    return hipSuccess;
}

template <class C>
KernelDesc fused_desc(const char* name)
{
    return KernelDesc{C::N0, C::N1, C::N2, prec_of<typename C::T>(), C::S, kFused, C::VPW,
                      C::NT, C::V17, 0, sizeof(cx<typename C::T>) * (size_t)C::LDS, name,
                      &launch_fused<C>};
}

template <class C, int PERSIST, int ST = 0, int HN = 3>
KernelDesc fpipe_desc(const char* name)
{
    return KernelDesc{C::N0, C::N1, C::N2, prec_of<typename C::T>(), C::S, kFused, C::VPW,
                      C::NT, C::V17, PERSIST, sizeof(cx<typename C::T>) * (size_t)C::LDS, name,
                      &launch_fpipe<C, PERSIST, ST, HN>};
}

template <class C>
KernelDesc pencil_desc(const char* name)
{
    return KernelDesc{C::N0, C::N1, C::N2, prec_of<typename C::T>(), C::S, kPencil, C::SPW,
                      C::NT, 0, C::PIPE, sizeof(cx<typename C::T>) * (size_t)C::LDS, name,
                      &launch_pencil<C>};
}

}  // namespace fx
