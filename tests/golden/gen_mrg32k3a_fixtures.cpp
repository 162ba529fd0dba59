// Writes MRG32k3a golden vectors with AMD rocRAND's own host implementation
// (MIT licence, /opt/rocm/include/rocrand/rocrand_mrg32k3a.h), which defines
// the (seed, subsequence, offset) semantics the build's generator follows.
// Build and run with gen_mrg32k3a_fixtures.sh; output is JSON on stdout.
#include <cstdio>
#include <cstring>
#include <cstdint>
#include <rocrand/rocrand_kernel.h>

static void dump(unsigned long long seed, unsigned long long sub, unsigned long long off, int first)
{
    rocrand_state_mrg32k3a st;
    rocrand_init(seed, sub, off, &st);
    // raw state via a copy of the public engine object
    struct raw { double bmd; float bmf; unsigned g1[3]; unsigned g2[3]; } r;
    static_assert(sizeof(raw) == sizeof(st), "state layout");
    std::memcpy(&r, &st, sizeof r);
    printf("%s{\"seed\": %llu, \"subsequence\": %llu, \"offset\": %llu, \"state\": [%u, %u, %u, %u, %u, %u], ",
           first ? "" : ",\n", seed, sub, off, r.g1[0], r.g1[1], r.g1[2], r.g2[0], r.g2[1], r.g2[2]);
    rocrand_state_mrg32k3a a = st, b = st, c = st;
    printf("\"next\": [");
    for (int i = 0; i < 8; ++i) printf("%s%u", i ? ", " : "", a.next());
    printf("], \"uniform_float_bits\": [");
    for (int i = 0; i < 8; ++i) {
        float f = rocrand_uniform(&b);
        uint32_t u; std::memcpy(&u, &f, 4);
        printf("%s%u", i ? ", " : "", u);
    }
    printf("], \"uniform_double_bits\": [");
    for (int i = 0; i < 8; ++i) {
        double d = rocrand_uniform_double(&c);
        uint64_t u; std::memcpy(&u, &d, 8);
        printf("%s%llu", i ? ", " : "", (unsigned long long)u);
    }
    printf("]}");
}

int main()
{
    const unsigned long long seeds[] = {1234ull, 7ull, 1000ull, 1063ull, 0ull, 1ull, 12345ull,
                                        4294967295ull, 4294967301ull, (1ull << 40) + 12345ull,
                                        (1ull << 63) + 17ull, 0xFFFFFFFFFFFFFFFFull, 0x123456789ABCDEFull};
    const unsigned long long subs[] = {0ull, 1ull, 2ull, 3ull, 255ull, 1000ull, 1048575ull,
                                       8388615ull, 2147483647ull};
    printf("{\"generator\": \"rocrand MRG32k3a host engine\", \"cases\": [\n");
    int first = 1;
    for (unsigned long long s : seeds)
        for (unsigned long long q : subs) { dump(s, q, 0, first); first = 0; }
    // a few nonzero offsets (skip-ahead within a subsequence)
    dump(1234ull, 5ull, 1ull, 0);
    dump(1234ull, 5ull, 1000ull, 0);
    dump(7ull, 0ull, (1ull << 20), 0);
    printf("\n]}\n");
    return 0;
}
