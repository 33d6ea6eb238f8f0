// gg_devrng.hpp -- counter-based random streams for draws made on the GPU.
//
// A stream is identified by (seed, step, lane) and is a splitmix64 sequence
// (Steele, Lea, Flood 2014, "Fast splittable pseudorandom number
// generators"): state += 0x9E3779B97F4A7C15, output = mix(state), where mix is
// the splitmix64 finaliser.  The starting state is
//     mix(mix(seed ^ 0xA0761D6478BD642F) ^ step * 0xE7037ED1A0B428DB)
//       ^ lane * 0x8EBC6AF09C88C6E3        (all arithmetic mod 2^64)
// so every (step, lane) pair draws independently of every other and of the
// order the GPU runs them in.  Bounded integers use Lemire's multiply-shift
// with rejection, as gg::Rng does for mt19937_64.
#pragma once
#include <cstdint>
#include <hip/hip_runtime.h>

namespace gg {

__host__ __device__ inline uint64_t sm64_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t sm64_stream(uint64_t seed, uint64_t step, uint64_t lane) {
    return sm64_mix(sm64_mix(seed ^ 0xA0761D6478BD642Full) ^ (step * 0xE7037ED1A0B428DBull)) ^
           (lane * 0x8EBC6AF09C88C6E3ull);
}

__host__ __device__ inline uint64_t sm64_next(uint64_t &s) {
    s += 0x9E3779B97F4A7C15ull;
    return sm64_mix(s);
}

__host__ __device__ inline uint64_t mul_hi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// uniform integer in [0, range), range >= 1
__host__ __device__ inline uint64_t sm64_below(uint64_t &s, uint64_t range) {
    uint64_t x = sm64_next(s);
    uint64_t l = x * range;
    if (l < range) {
        const uint64_t t = (0 - range) % range;
        while (l < t) {
            x = sm64_next(s);
            l = x * range;
        }
    }
    return mul_hi64(x, range);
}

}  // namespace gg
