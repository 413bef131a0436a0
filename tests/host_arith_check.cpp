// Host build of the device inverse in damascus000_amd/csrc/dmx_arith.hpp, for the CPU tests.
//   host_arith_check inv            reads canonical Z_q values (hex, one per line) and prints
//                                   zq_inv_divsteps of each
//   host_arith_check random N SEED  compares zq_inv_divsteps with x^(q-2) (zq_inv) and checks
//                                   x * x^-1 = 1 on N seeded inputs plus edge cases
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>

#include "dmx_arith.hpp"

using namespace dmx;

static zq make(uint64_t lo, uint64_t hi) {
    return zq{{(uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32)}};
}

static void print(const zq& a) { printf("%04x%08x%08x%08x\n", a.w[3], a.w[2], a.w[1], a.w[0]); }

static bool parse(const char* s, zq& out) {
    uint32_t w[4] = {0, 0, 0, 0};
    size_t n = strlen(s);
    while (n && (s[n - 1] == '\n' || s[n - 1] == '\r')) --n;
    if (n == 0 || n > 28) return false;
    for (size_t i = 0; i < n; ++i) {
        const char c = s[n - 1 - i];
        uint32_t v;
        if (c >= '0' && c <= '9') v = c - '0';
        else if (c >= 'a' && c <= 'f') v = c - 'a' + 10;
        else if (c >= 'A' && c <= 'F') v = c - 'A' + 10;
        else return false;
        w[i / 8] |= v << (4 * (i % 8));
    }
    out = zq{{w[0], w[1], w[2], w[3]}};
    return !zq_ge_q(out);
}

int main(int argc, char** argv) {
    if (argc >= 2 && !strcmp(argv[1], "inv")) {
        char line[128];
        while (fgets(line, sizeof line, stdin)) {
            zq x;
            if (!parse(line, x)) return 2;
            print(zq_inv_divsteps(x));
        }
        return 0;
    }
    if (argc < 4 || strcmp(argv[1], "random")) return 2;
    const long n = atol(argv[2]);
    std::mt19937_64 rng(strtoull(argv[3], nullptr, 10));
    long tested = 0, bad = 0;
    auto check = [&](const zq& x) {
        if (zq_ge_q(x)) return;
        ++tested;
        const zq a = zq_inv_divsteps(x);
        bool ok = zq_eq(a, zq_inv(x)) && !zq_ge_q(a);
        if (!zq_is_zero(x) && !zq_eq(zq_mul(a, x), zq_from_u64(1))) ok = false;
        if (!ok && bad++ < 5) {
            printf("mismatch at ");
            print(x);
        }
    };
    const zq one = zq_from_u64(1);
    for (uint64_t v = 0; v < 2000; ++v) {
        check(make(v, 0));
        check(zq_neg(make(v, 0)));
    }
    for (int s = 0; s < 112; ++s) {  // powers of two, their negatives and neighbours
        const zq p = s < 64 ? make(1ull << s, 0) : make(0, 1ull << (s - 64));
        check(p);
        check(zq_neg(p));
        check(zq_add(p, one));
        check(zq_sub(p, one));
    }
    for (long i = 0; i < n; ++i) {
        uint64_t lo = rng(), hi = rng() & 0xFFFFFFFFFFFFull;
        if (i % 4 == 1) hi >>= rng() % 48;                      // shorter values
        if (i % 4 == 2) lo &= ~((1ull << (rng() % 64)) - 1);  // long runs of low zero bits
        check(make(lo, hi));
    }
    printf("tested %ld mismatches %ld\n", tested, bad);
    return bad ? 1 : 0;
}
