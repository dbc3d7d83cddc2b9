// TEST INFRASTRUCTURE ONLY.  Known-answer harness for PCG32.
//
// Built by oracle/Makefile (target _ref/pcg_kat) against the reference's own
// vendored third-party header dep/pcg_random.hpp, which is the generator the
// reference instantiates at src/core/sampler.hpp:32 (pcg32 =
// pcg_engines::setseq_xsh_rr_64_32, pcg_random.hpp:1862).  It prints the
// first outputs of pcg32(seed) for seeds 0..4 and of pcg32(state, stream) for
// a few explicit streams, one JSON object per line, so the restated generator
// in refcpu/ref_rng.hpp can be compared with it.
#include <cstdio>
#include <cstdint>
#include "pcg_random.hpp"

int main() {
    for (uint64_t s = 0; s < 5; ++s) {
        pcg32 g(s);
        std::printf("{\"seed\": %llu, \"stream\": null, \"out\": [", (unsigned long long)s);
        for (int i = 0; i < 16; ++i) std::printf("%s%u", i ? ", " : "", (unsigned)g());
        std::printf("]}\n");
    }
    const uint64_t streams[3][2] = {{42u, 54u}, {0x853c49e6748fea9bULL, 0xda3e39cb94b95bdbULL}, {7u, 1u}};
    for (auto& st : streams) {
        pcg32 g(st[0], st[1]);
        std::printf("{\"seed\": %llu, \"stream\": %llu, \"out\": [", (unsigned long long)st[0], (unsigned long long)st[1]);
        for (int i = 0; i < 16; ++i) std::printf("%s%u", i ? ", " : "", (unsigned)g());
        std::printf("]}\n");
    }
    return 0;
}
