// TEST INFRASTRUCTURE: host emulation of the dense floor (dense.hpp): every line of each
// pass through the same line function, passes in launch order.
#pragma once

namespace fx {

template <class C>
hipError_t launch_dense(const void* in, void* out, long nvol, const void* bank, double scl,
                        unsigned*, hipStream_t)
{
    using T = typename C::T;
    using K = Dense<C>;
    const Bank<T>* bk = (const Bank<T>*)bank;
    const cx<T>* src = (const cx<T>*)in;
    cx<T>* dst = (cx<T>*)out;
    for (long l = 0; l < nvol * (C::V / C::N2); ++l) K::template line<2>(l, nvol, src, dst, bk, T(1));
    for (long l = 0; l < nvol * (C::V / C::N1); ++l) K::template line<1>(l, nvol, dst, dst, bk, T(1));
    for (long l = 0; l < nvol * (C::V / C::N0); ++l) K::template line<0>(l, nvol, dst, dst, bk, (T)scl);
    return hipSuccess;
}

template <class C>
KernelDesc dense_desc(const char* name)
{
    return KernelDesc{C::N0, C::N1, C::N2, prec_of<typename C::T>(), C::S, kFused, 1, C::NT, 0,
                      0, 0, name, &launch_dense<C>};
}

}  // namespace fx
