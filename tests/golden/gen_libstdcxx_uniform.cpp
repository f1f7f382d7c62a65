// Writes golden vectors of std::mt19937_64 + std::uniform_int_distribution<uint64_t>
// as compiled by this image's g++ (libstdc++ 11), the generator the reference's
// generate_random_vector uses (cuda/matrices.h:44-56).  Run by make_golden.sh.
#include <cstdint>
#include <cstdio>
#include <random>

int main() {
    struct Case { uint64_t seed, lo, hi; int count; };
    const Case cases[] = {
        {1, 1, 108, 64}, {12345, 0, 65520, 64}, {987654321, 1, 2147483646, 64},
        {0, 0, 2147483646, 64}, {42, 0, 6, 64}, {7, 1, 4294967294ULL, 64},
    };
    std::printf("{\n  \"cases\": [\n");
    for (size_t c = 0; c < sizeof(cases) / sizeof(cases[0]); ++c) {
        std::mt19937_64 gen(cases[c].seed);
        std::uniform_int_distribution<uint64_t> dis(cases[c].lo, cases[c].hi);
        std::printf("    {\"seed\": %llu, \"lo\": %llu, \"hi\": %llu, \"values\": [",
                    (unsigned long long)cases[c].seed, (unsigned long long)cases[c].lo,
                    (unsigned long long)cases[c].hi);
        for (int i = 0; i < cases[c].count; ++i)
            std::printf("%s%llu", i ? ", " : "", (unsigned long long)dis(gen));
        std::printf("]}%s\n", c + 1 < sizeof(cases) / sizeof(cases[0]) ? "," : "");
    }
    std::mt19937_64 def;  // default seed 5489
    unsigned long long v = 0;
    for (int i = 0; i < 10000; ++i) v = def();
    std::printf("  ],\n  \"mt19937_64_default_10000th\": %llu\n}\n", v);
    return 0;
}
