// gg_rng.hpp -- the one random-number mapping shared by generators and search
// drivers, so GPU and CPU runs of a driver draw identical sequences.
//
// std::mt19937_64 is fully specified by the C++ standard; the bounded-integer
// and real mappings below are explicit (Lemire's multiply-shift with rejection,
// and 53-bit mantissa doubles), unlike std::uniform_int_distribution /
// uniform_real_distribution whose algorithms are implementation-defined.
#pragma once
#include <cstdint>
#include <random>

namespace gg {

struct Rng {
    std::mt19937_64 eng;
    explicit Rng(uint64_t seed) : eng(seed) {}
    uint64_t next() { return eng(); }
    // uniform integer in [0, range), range >= 1 (Lemire 2019, "Fast random
    // integer generation in an interval", Algorithm 5)
    uint64_t below(uint64_t range) {
        uint64_t x = eng();
        unsigned __int128 m = (unsigned __int128)x * range;
        uint64_t l = (uint64_t)m;
        if (l < range) {
            uint64_t t = (0 - range) % range;
            while (l < t) {
                x = eng();
                m = (unsigned __int128)x * range;
                l = (uint64_t)m;
            }
        }
        return (uint64_t)(m >> 64);
    }
    // uniform double in [0, 1): top 53 bits
    double uniform() { return (double)(eng() >> 11) * (1.0 / 9007199254740992.0); }
};

}  // namespace gg
