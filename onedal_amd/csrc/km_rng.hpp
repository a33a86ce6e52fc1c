// km_rng.hpp -- counter-based random numbers of the initialisation methods.
//
// Draw i of stream s under seed q is splitmix64(splitmix64(q ^ s * C1) ^ i * C2),
// splitmix64 being the finaliser of Steele, Lea and Flood's SplitMix64 (OOPSLA
// 2014).  Uniform doubles take the top 53 bits; uniform indices in [0, m) are
// the high 64 bits of draw * m.  The reference draws from DAAL's mt19937 engine
// (backend/primitives/rng/host_engine.hpp:160-185); its streams are not
// reproduced, so these methods match the reference in distribution only.
#pragma once

#include <cstdint>

namespace km {

inline uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

inline uint64_t rng_draw(uint64_t seed, uint64_t stream, uint64_t i) {
    return splitmix64(splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull)) ^ (i * 0x9E3779B97F4A7C15ull));
}

inline double rng_uniform01(uint64_t seed, uint64_t stream, uint64_t i) {
    return double(rng_draw(seed, stream, i) >> 11) * (1.0 / 9007199254740992.0);
}

inline uint64_t rng_index(uint64_t seed, uint64_t stream, uint64_t i, uint64_t m) {
    return uint64_t((static_cast<unsigned __int128>(rng_draw(seed, stream, i)) * m) >> 64);
}

constexpr uint64_t RNG_STREAM_PLUS_PLUS = 1;
constexpr uint64_t RNG_STREAM_RANDOM = 2;

}  // namespace km
