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
        unsigned work = 0;  // the launch must hand it back at zero
        if (k.launch(in, out, nvol, img.data(), scale, &work, nullptr) != hipSuccess) return -2;
        if (g_emu_bad_block >= 0) return (int)(-1000 - g_emu_bad_block);
        return work == 0 ? 0 : -3;
    }
    return -1;
}

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
    g_emu_reverse = 0;
    g_emu_bad_block = -1;
    unsigned work = 0;  // claim counter; the launch must hand it back at zero
    (void)d.launch(&a, tasks.data(), (int)tasks.size(), img.data(), &work, nullptr);
    if (g_emu_bad_block >= 0) return (int)(-1000 - g_emu_bad_block);
    return work == 0 ? 0 : -3;
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
