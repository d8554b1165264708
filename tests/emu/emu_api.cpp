// TEST INFRASTRUCTURE: C entry points of the host emulator library (tests only).
#include <cstring>
#include <vector>

#include "cx.hpp"
#include "plan_internal.hpp"
#include "registry.hpp"
#include "multi_tasks.hpp"

namespace fx {
int g_emu_reverse = 0;
long g_emu_bad_block = -1;
}  // namespace fx

using namespace fx;

extern "C" int emu_kernel_count(void) { return (int)registry().size(); }

extern "C" const char* emu_kernel_name(int i)
{
    return (i >= 0 && i < (int)registry().size()) ? registry()[i].name : nullptr;
}

// Run kernel `name` on host buffers.  Returns 0, -1 for an unknown name, -2 when the
// launcher refuses, -3 when the claim counter is not back at zero, or -1000 - block when a
// block wrote outside its LDS allocation.
extern "C" int emu_transform(const char* name, const void* in, void* out, long nvol,
                             double scale, int reverse)
{
    for (const KernelDesc& k : registry()) {
        if (strcmp(k.name, name)) continue;
        std::vector<unsigned char> img;
        bank_host_image(k.prec, img);
        g_emu_reverse = reverse;
        g_emu_bad_block = -1;
        std::vector<unsigned> work(kClaimWords, 0u);  // the launch must hand them back at zero
        if (k.launch(in, out, nvol, img.data(), scale, work.data(), nullptr) != hipSuccess)
            return -2;
        if (g_emu_bad_block >= 0) return (int)(-1000 - g_emu_bad_block);
        for (unsigned c : work)
            if (c != 0) return -3;
        return 0;
    }
    return -1;
}

// Claim queues at a real grid: `grid` workgroups, each starting at its own block and then
// claiming (emu_claim, the device arithmetic) until refused, interleaved in a random order.
// Returns 0 when every block ran exactly once and every counter is back at zero, -1 for a
// block run twice or out of range, -2 for a block never run, -3 for a counter left non-zero.
extern "C" int emu_claim_sim(long grid, long nblocks, unsigned seed)
{
    if (grid < 1 || nblocks < grid) return -1;
    std::vector<unsigned> work(kClaimWords, 0u);
    std::vector<int> runs(nblocks, 0);
    std::vector<long> cur(grid);
    std::vector<long> live;
    for (long w = 0; w < grid; ++w) {
        cur[w] = w;
        live.push_back(w);
    }
    unsigned r = seed * 2654435761u + 1u;
    while (!live.empty()) {
        r = r * 1664525u + 1013904223u;
        const size_t i = (r >> 8) % live.size();
        const long w = live[i];
        if (cur[w] < 0 || cur[w] >= nblocks || runs[cur[w]]++) return -1;
        const long nb = emu_claim(work.data(), nblocks, grid, w, cur[w]);
        if (nb >= nblocks) {
            live[i] = live.back();
            live.pop_back();
        } else {
            cur[w] = nb;
        }
    }
    for (long b = 0; b < nblocks; ++b)
        if (runs[b] != 1) return -2;
    for (unsigned c : work)
        if (c != 0) return -3;
    return 0;
}

// Thread and workgroup order of the next emu_multi calls (0 ascending, 1 descending).
static int g_multi_reverse = 0;
extern "C" void emu_set_multi_reverse(int r) { g_multi_reverse = r; }

extern "C" int emu_multi_count(void) { return (int)multi_registry().size(); }
extern "C" const char* emu_multi_name(int i)
{
    return (i >= 0 && i < (int)multi_registry().size()) ? multi_registry()[i].name : nullptr;
}

// Run mixed kernel `name` over nitems items (shape[3*i..], batch[i], dir[i], scale[i]) on
// host buffers.  Returns 0, -1 for an unknown name or geometry, -3 when the claim counter is
// not back at zero, -1000 - task on an LDS guard violation.
template <typename T>
static int emu_multi_t(const MultiDesc& d, int n, const int* shape, const int* batch,
                       const int* dir, const double* scale, const void* const* in,
                       void* const* out)
{
    MultiArgs<T> a{};
    int code[kMultiMaxItems];
    for (int i = 0; i < n; ++i) {
        code[i] = multi_code(shape[3 * i], shape[3 * i + 1], shape[3 * i + 2]);
        if (code[i] < 0) return -1;
        a.item[i] = MultiItem<T>{(const cx<T>*)in[i], (cx<T>*)out[i], batch[i], code[i],
                                 dir[i] > 0 ? 1 : 0, (T)scale[i]};
    }
    std::vector<int2> tasks;
    multi_tasks(d, code, batch, n, sizeof(cx<T>), tasks);
    std::vector<unsigned char> img;
    bank_host_image(d.prec, img);
    g_emu_reverse = g_multi_reverse;
    g_emu_bad_block = -1;
    std::vector<unsigned> work(kClaimWords, 0u);  // must all be back at zero afterwards
    (void)d.launch(&a, tasks.data(), (int)tasks.size(), img.data(), work.data(), nullptr);
    if (g_emu_bad_block >= 0) return (int)(-1000 - g_emu_bad_block);
    for (unsigned c : work)
        if (c != 0) return -3;
    return 0;
}

extern "C" int emu_multi(const char* name, int n, const int* shape, const int* batch,
                         const int* dir, const double* scale, const void* const* in,
                         void* const* out)
{
    if (n < 1 || n > kMultiMaxItems) return -1;
    for (const MultiDesc& d : multi_registry()) {
        if (strcmp(d.name, name)) continue;
        return d.prec == kC128 ? emu_multi_t<double>(d, n, shape, batch, dir, scale, in, out)
                               : emu_multi_t<float>(d, n, shape, batch, dir, scale, in, out);
    }
    return -1;
}
