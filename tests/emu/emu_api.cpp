// TEST INFRASTRUCTURE: C entry points of the host emulator library (tests only).
#include <cstring>
#include <vector>

#include "cx.hpp"
#include "plan_internal.hpp"
#include "registry.hpp"

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

// Run kernel `name` on host buffers.  Returns 0, -1 for an unknown name, or
// -1000 - block when a block wrote outside its LDS allocation.
extern "C" int emu_transform(const char* name, const void* in, void* out, long nvol,
                             double scale, int reverse)
{
    for (const KernelDesc& k : registry()) {
        if (strcmp(k.name, name)) continue;
        std::vector<unsigned char> img;
        bank_host_image(k.prec, img);
        g_emu_reverse = reverse;
        g_emu_bad_block = -1;
        (void)k.launch(in, out, nvol, img.data(), scale, nullptr);
        return g_emu_bad_block >= 0 ? (int)(-1000 - g_emu_bad_block) : 0;
    }
    return -1;
}
