// Device launchers and registry descriptors for the kernel templates of kernels.hpp.
// Included at the end of kernels.hpp through the include path.
#pragma once
#include <mutex>
#include <unordered_map>

namespace fx {

// Compute units of the current device (cached per device ordinal).
inline int cu_count()
{
    static int cache[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (!cache[dev]) {
        int c = 0;
        if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            c < 1)
            c = 1;
        cache[dev] = c;
    }
    return cache[dev];
}

// Workgroups of kernel `f` (nt threads) resident per CU, cached per kernel (by address:
// kernels of one signature share a function pointer type, not an occupancy).
template <typename F>
int resident_per_cu(F f, int nt)
{
    static std::mutex mu;
    static std::unordered_map<const void*, int> cache;
    const void* key = reinterpret_cast<const void*>(f);
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int o = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, key, nt, 0) != hipSuccess || o < 1)
        o = 1;
    cache.emplace(key, o);
    return o;
}

// Grid of a persistent kernel: at most `cap` workgroups per CU, never more than nblocks.
template <typename F>
long persistent_grid(F f, int nt, int cap, long nblocks)
{
    int per = resident_per_cu(f, nt);
    if (per > cap) per = cap;
    const long g = (long)per * cu_count();
    return g < nblocks ? g : nblocks;
}

// Dynamic block claiming pays when workgroups walk several blocks each; when the grid
// covers (nearly) every block, the one-per-workgroup failing claims would only queue on a
// single address, so those launches walk a fixed stride instead.
inline unsigned* claim_counter(unsigned* work, long grid, long nblocks)
{
    return nblocks >= 2 * grid ? work : nullptr;
}

template <class C>
hipError_t launch_fused(const void* in, void* out, long nvol, const void* bank, double scl,
                        unsigned*, hipStream_t stream)
{
    using T = typename C::T;
    if (nvol <= 0) return hipSuccess;
    const long grid = (nvol + C::VPW - 1) / C::VPW;
    hipLaunchKernelGGL((k_fused<C>), dim3((unsigned)grid), dim3(C::NT), 0, stream,
                       (const cx<T>*)in, (cx<T>*)out, nvol, (const Bank<T>*)bank, (T)scl);
    return hipGetLastError();
}

template <class C, int PERSIST, int ST = 0, int HN = 3>
hipError_t launch_fpipe(const void* in, void* out, long nvol, const void* bank, double scl,
                        unsigned* work, hipStream_t stream)
{
    using T = typename C::T;
    if (nvol <= 0) return hipSuccess;
    const long nblocks = (nvol + C::VPW - 1) / C::VPW;
    if (nblocks >= (1L << 31)) return hipErrorInvalidValue;
    using P = PipeCfg<C, PERSIST, ST, HN>;
    const long grid = PERSIST > 0 ? persistent_grid(k_fpipe<P>, C::NT, PERSIST, nblocks) : nblocks;
    hipLaunchKernelGGL((k_fpipe<P>), dim3((unsigned)grid), dim3(C::NT), 0, stream,
                       (const cx<T>*)in, (cx<T>*)out, nvol, nblocks, (const Bank<T>*)bank,
                       (T)scl, claim_counter(work, grid, nblocks));
    return hipGetLastError();
}

template <class C>
hipError_t launch_pencil(const void* in, void* out, long nvol, const void* bank, double scl,
                         unsigned* work, hipStream_t stream)
{
    using T = typename C::T;
    if (nvol <= 0) return hipSuccess;
    const long chunk = C::CHUNK > 0 ? C::CHUNK : nvol;
    for (long v0 = 0; v0 < nvol; v0 += chunk) {
        const long nv = (nvol - v0) < chunk ? (nvol - v0) : chunk;
        const cx<T>* src = (const cx<T>*)in + v0 * C::V;
        cx<T>* dst = (cx<T>*)out + v0 * C::V;
        const long nslab = nv * C::N0;
        const long nblocks = (nslab + C::SPW - 1) / C::SPW;
        if constexpr (C::PIPE > 0) {
            const long ga = persistent_grid(k_spipe<C>, C::NT, C::PIPE, nblocks);
            if (nblocks >= (1L << 31)) return hipErrorInvalidValue;
            hipLaunchKernelGGL((k_spipe<C>), dim3((unsigned)ga), dim3(C::NT), 0, stream, src,
                               dst, nslab, nblocks, (const Bank<T>*)bank,
                               claim_counter(work, ga, nblocks));
        } else {
            hipLaunchKernelGGL((k_slab<C>), dim3((unsigned)nblocks), dim3(C::NT), 0, stream,
                               src, dst, nslab, (const Bank<T>*)bank);
        }
        const long nlines = nv * C::PLANE;
        const long gb = (nlines + C::NTB - 1) / C::NTB;
        hipLaunchKernelGGL((k_xline<C>), dim3((unsigned)gb), dim3(C::NTB), 0, stream, dst,
                           nlines, (const Bank<T>*)bank, (T)scl);
    }
    return hipGetLastError();
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
