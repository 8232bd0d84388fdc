// tg_rand.h -- the counter-based generator of the seeded initial field and
// of the random band forcing: U(seed, ctr) in [0, 1) from SplitMix64
// (Steele, Lea and Flood 2014), the same definition as the oracle's.
#pragma once
#include <stdint.h>

namespace tg {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t ctr)
{
    const uint64_t h = splitmix64(splitmix64(seed) ^ ctr);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace tg
