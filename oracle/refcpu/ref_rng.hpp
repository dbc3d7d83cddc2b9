// refcpu — TEST INFRASTRUCTURE ONLY (see ref_math.hpp header).
//
// PCG32 (O'Neill, "PCG: A Family of Simple Fast Space-Efficient Statistically
// Good Algorithms for Random Number Generation", 2014): 64-bit LCG state,
// XSH-RR 32-bit output of the *previous* state.  The reference uses
// pcg_engines::setseq_xsh_rr_64_32 from its vendored dep/pcg_random.hpp:1862
// (called at src/core/sampler.hpp:32).  pcg32(seed) in that header equals
// seed(seed, default stream) below; both are pinned by the KAT fixture that
// oracle/pcg_kat.cpp produces from the vendored header.
//
// The reference maps outputs to floats through libstdc++'s
// std::uniform_real_distribution<float> (sampler.hpp:85), which no reference
// test pins; this build uses its own mapping, the top 24 bits times 2^-24.
#pragma once
#include <cstdint>

namespace refcpu {

struct Pcg32 {
    static constexpr uint64_t kMult = 6364136223846793005ULL;
    static constexpr uint64_t kDefaultInc = 1442695040888963407ULL;
    uint64_t state = 0x853c49e6748fea9bULL;
    uint64_t inc = 0xda3e39cb94b95bdbULL;

    // pcg32_srandom_r(initstate, initseq)
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0;
        inc = (initseq << 1u) | 1u;
        next_u32();
        state += initstate;
        next_u32();
    }
    // C++ pcg32(seed): default increment, state = bump(seed + inc)
    void seed_default_stream(uint64_t s) {
        inc = kDefaultInc;
        state = (s + inc) * kMult + inc;
    }
    uint32_t next_u32() {
        uint64_t old = state;
        state = old * kMult + inc;
        uint32_t xorshifted = uint32_t(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = uint32_t(old >> 59u);
        return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
    }
    float next_float() { return float(next_u32() >> 8) * (1.0f / 16777216.0f); }
};

// Per-(pixel, sample) stream derivation shared by the fixed-mode renderer:
// a 64-bit avalanche hash (SplitMix64 finaliser) of the key.
inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ULL;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

}  // namespace refcpu
