// Compile-time loop: body(std::integral_constant<int, I>) for I in [B, E).
// Used for the 19 lattice links so that every lattice factor is a constant
// expression and exact zeros are resolved at compile time.
#pragma once

#include <type_traits>
#include <utility>

namespace acp {

template <int B, typename F, int... Is>
__host__ __device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, B + Is>{}), ...);
}

template <int B, int E, typename F>
__host__ __device__ __forceinline__ void static_for(F&& f) {
    static_for_impl<B>(static_cast<F&&>(f), std::make_integer_sequence<int, E - B>{});
}

}  // namespace acp
