// walk.hpp -- the sample loop of the level walkers (kernels_walk.hip,
// kernels_walkseg.hip): one workgroup walks a range of samples in order,
// with the operand loads of the next NB - 1 samples in flight while the
// current one is formed.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace magnus {

template <int U, int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (U < N) {
        f(std::integral_constant<int, U>{});
        static_for<U + 1, N>(f);
    }
}

// Samples lo..hi in order (UP) or hi..lo.  Sample k of the walk (k = 0 first)
// sits in buffer k % NB, a named register set: the loop is unrolled UNROLL
// times (a multiple of NB), so buffers rotate by renaming, never by copying a
// register whose load is in flight.  step(i, ops, u) gets u = k % UNROLL as a
// std::integral_constant, so a step can specialise on its position (with
// UNROLL a multiple of 4 and a walk that starts at a known position in its
// Boole block, u fixes the sample's position in its block).
//
// The unrolled loop has no branch around a load: near the end it reloads the
// walk's last sample into a buffer whose sample was already formed.  The
// compiler then counts, before each step, exactly the memory operations issued
// after that step's loads and waits for those loads only; a load under a
// branch would make it wait for every load in flight at the join.
template <int NB, int UNROLL, bool UP, typename Ops, typename Load, typename Step>
__device__ __forceinline__ void walk_range(long long lo, long long hi, Ops* o, Load load, Step step) {
    static_assert(NB >= 1 && UNROLL % NB == 0, "the unroll must be a multiple of the buffer count");
    if (lo > hi) return;
    const long long n = hi - lo + 1;
    auto at = [&](long long k) {
        const long long c = k < n ? k : n - 1;
        return UP ? lo + c : hi - c;
    };
    static_for<0, NB - 1>([&](auto pc) { load(at(decltype(pc)::value), o[decltype(pc)::value]); });
    const long long full = n - n % UNROLL;
    for (long long k0 = 0; k0 < full; k0 += UNROLL) {
        static_for<0, UNROLL>([&](auto uc) {
            constexpr int u = decltype(uc)::value;
            load(at(k0 + u + NB - 1), o[(u + NB - 1) % NB]);
            step(at(k0 + u), o[u % NB], uc);
        });
    }
    static_for<0, UNROLL - 1>([&](auto uc) {
        constexpr int u = decltype(uc)::value;
        if (full + u < n) {
            load(at(full + u + NB - 1), o[(u + NB - 1) % NB]);
            step(at(full + u), o[u % NB], uc);
        }
    });
}

}  // namespace magnus
