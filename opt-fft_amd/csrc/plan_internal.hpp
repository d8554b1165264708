// Internal helpers shared by the plan step and the ABI shims.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace fx {

// Record an error message for fft3d_mx_last_error() and return `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
const char* last_error();
void clear_error();
size_t elem_bytes(int prec);
// True when p is a multiple of the element size of `prec` (kernels load whole elements).
inline bool elem_aligned(const void* p, int prec)
{
    return (uintptr_t)p % elem_bytes(prec) == 0;
}
void bank_host_image(int prec, std::vector<unsigned char>& img);

}  // namespace fx
