// TEST INFRASTRUCTURE: AddressSanitizer sweep of every kernel variant on the host.
// Each run gets exactly-sized heap buffers for `in` and `out`, so any global access outside
// them is reported by ASan; exit status 0 means every variant and batch size was clean.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cx.hpp"
#include "registry.hpp"

extern "C" int emu_transform(const char* name, const void* in, void* out, long nvol,
                             double scale, int reverse);

int main()
{
    using namespace fx;
    int bad = 0;
    for (const KernelDesc& k : registry()) {
        if (k.algo != 0) continue;  // matrix-core kernels are not emulated
        const size_t V = (size_t)k.n0 * k.n1 * k.n2;
        const size_t eb = k.prec == kC128 ? 16 : 8;
        const long batches[] = {1, 2, (long)k.vpw + 1, 2L * k.vpw + 3, 7L * k.vpw + 1};
        for (long B : batches) {
            const size_t bytes = V * (size_t)B * eb;
            unsigned char* in = (unsigned char*)malloc(bytes);
            unsigned char* out = (unsigned char*)malloc(bytes);
            for (size_t i = 0; i < bytes; ++i) in[i] = (unsigned char)(i * 2654435761u >> 24) & 0x3f;
            const int rc = emu_transform(k.name, in, out, B, 0.5, 0);
            if (rc != 0) {
                printf("FAIL %s B=%ld rc=%d\n", k.name, B, rc);
                ++bad;
            }
            free(in);
            free(out);
        }
    }
    printf("%s: %zu variants swept\n", bad ? "FAILED" : "ok", registry().size());
    return bad ? 1 : 0;
}
