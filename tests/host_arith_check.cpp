// Host builds of device routines shared through the headers in damascus000_amd/csrc, for the
// CPU tests.
//   host_arith_check inv            reads canonical Z_q values (hex, one per line) and prints
//                                   zq_inv_divsteps of each
//   host_arith_check random N SEED  compares zq_inv_divsteps with x^(q-2) (zq_inv) and checks
//                                   x * x^-1 = 1 on N seeded inputs plus edge cases
//   host_arith_check tx             reads "label round epoch_hex file_id_hex n0 n1 n2", then the
//                                   K (n0 + n1 + n2) coefficients (hex, one per line), and prints
//                                   the challenge message assembled from tx_words4 (hex)
//   host_arith_check modp N SEED    prints N seeded cases per CRT prime (plus edge cases) of the
//                                   mod-p routines, one per line: "op prime_index operands result"
//                                   in hex, for the caller to check with exact integers
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "dmx_arith.hpp"
#include "dmx_transcript.hpp"

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

static bool hex_bytes(const char* s, uint8_t* out, size_t n) {
    if (strlen(s) != 2 * n) return false;
    for (size_t i = 0; i < n; ++i) {
        unsigned v;
        if (sscanf(s + 2 * i, "%2x", &v) != 1) return false;
        out[i] = (uint8_t)v;
    }
    return true;
}

static int run_tx() {
    char label[8], ep_hex[80], fid_hex[80];
    unsigned long long rnd;
    unsigned n[3];
    if (scanf("%7s %llu %79s %79s %u %u %u", label, &rnd, ep_hex, fid_hex, &n[0], &n[1], &n[2]) != 7) return 2;
    alignas(8) uint8_t ep[32];
    alignas(8) uint8_t fid[32];
    if (!hex_bytes(ep_hex, ep, 32) || !hex_bytes(fid_hex, fid, 32)) return 2;
    std::vector<zq> mods[3];
    char line[128];
    for (int i = 0; i < 3; ++i) {
        mods[i].resize((size_t)K * n[i]);
        for (auto& c : mods[i]) {
            if (scanf("%127s", line) != 1 || !parse(line, c)) return 2;
        }
    }
    TranscriptMsg t;
    t.epoch_seed = ep;
    t.file_id = fid;
    t.label_len = (uint32_t)strlen(label);
    memset(t.label, 0, 4);
    memcpy(t.label, label, t.label_len);
    t.round = rnd;
    for (int i = 0; i < 3; ++i) {
        t.mods[i] = mods[i].data();
        t.lens[i] = n[i];
    }
    const size_t len = tx_message_len(t.label_len, t.lens);
    std::vector<uint64_t> words((len + 31) / 32 * 4);
    for (size_t w = 0; w < words.size(); w += 4) tx_words4(t, (uint32_t)w, &words[w]);
    std::string hex;
    char b[3];
    for (size_t o = 0; o < len; ++o) {
        snprintf(b, sizeof b, "%02x", (unsigned)((words[o >> 3] >> (8 * (o & 7))) & 0xFF));
        hex += b;
    }
    for (size_t o = len; o < 8 * words.size(); ++o)  // past the message every byte reads zero
        if ((words[o >> 3] >> (8 * (o & 7))) & 0xFF) return 3;
    printf("%s\n", hex.c_str());
    return 0;
}

// constants of prime i as the runtime builds them
static PrimeConsts prime_consts(int i) {
    const uint64_t p = PRIMES[i];
    PrimeConsts k;
    memset(&k, 0, sizeof k);
    k.p = (uint32_t)p;
    k.pinv = host_inv32((uint32_t)p);
    const uint64_t R = (1ull << 32) % p;
    k.r4_m = (uint32_t)host_powmod(R, 4, p);
    k.r4_mp = k.r4_m * k.pinv;
    k.r3_m = (uint32_t)host_powmod(R, 3, p);
    k.r3_mp = k.r3_m * k.pinv;
    const unsigned __int128 q = ((unsigned __int128)1 << 112) - 75;
    k.q_mod_p = (uint32_t)(q % p);
    return k;
}

static int run_modp(long n, uint64_t seed) {
    std::mt19937_64 rng(seed);
    const zq qm1 = zq{{Q0 - 1, Q1, Q2, Q3}};
    for (int i = 0; i < NPRIMES; ++i) {
        const PrimeConsts c = prime_consts(i);
        const uint32_t p = c.p;
        const uint32_t edge[] = {0u, 1u, 2u, p - 2, p - 1, p / 2, p / 2 + 1};
        const uint32_t edge32[] = {0u, 1u, p - 1, p, p + 1, 0x7FFFFFFFu, 0x80000000u, 0xFFFFFFFEu, 0xFFFFFFFFu};
        auto residue = [&] { return (uint32_t)(rng() % p); };
        auto word = [&] { return (uint32_t)rng(); };
        auto pair = [&](uint32_t a, uint32_t b) {
            printf("add %d %x %x %x\n", i, a, b, addp(a, b, p));
            printf("sub %d %x %x %x\n", i, a, b, subp(a, b, p));
        };
        auto mul = [&](uint32_t b, uint32_t w) {
            const uint32_t wm = (uint32_t)(((uint64_t)w << 32) % p);
            printf("mulc %d %x %x %x\n", i, b, w, mulc(b, wm, wm * c.pinv, p));
        };
        auto red = [&](uint32_t hi, uint32_t lo) { printf("redc %d %x %x %x\n", i, hi, lo, redc64(hi, lo, p, c.pinv)); };
        auto zqc = [&](const zq& x) {
            printf("modp %d %x%08x%08x%08x %x %x\n", i, x.w[3], x.w[2], x.w[1], x.w[0], zq_mod_p(x, c),
                   zq_mod_p_centred(x, c));
        };
        auto u96 = [&](uint64_t lo, uint32_t hi) {
            printf("u96 %d %llx %x %x\n", i, (unsigned long long)lo, hi, reduce_u96(lo, hi, c));
        };
        for (uint32_t a : edge)
            for (uint32_t b : edge) pair(a, b);
        for (uint32_t b : edge32)
            for (uint32_t w : edge) mul(b, w);
        for (uint32_t h : edge32)
            for (uint32_t l : edge32) red(h, l);
        const zq xs[] = {zq_zero(), zq_from_u64(1), qm1, zq{{0xFFFFFFDAu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0x7FFFu}},
                         zq{{0xFFFFFFDBu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0x7FFFu}}, zq{{0u, 0u, 0u, 0x8000u}},
                         zq{{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0x7FFEu}}};
        for (const zq& x : xs) zqc(x);
        for (uint32_t h : edge32) {
            u96(~0ull, h);
            u96(0, h);
            u96(((uint64_t)p << 32) | (p - 1), h);
        }
        for (long t = 0; t < n; ++t) {
            pair(residue(), residue());
            mul(word(), residue());
            red(word(), word());
            uint64_t lo = rng(), hi = rng() & 0xFFFFFFFFFFFFull;
            zq x = zq{{(uint32_t)lo, (uint32_t)(lo >> 32), (uint32_t)hi, (uint32_t)(hi >> 32) & 0xFFFFu}};
            if (zq_ge_q(x)) x = zq_sub_q(x);
            zqc(x);
            u96(rng(), word());
        }
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 2 && !strcmp(argv[1], "tx")) return run_tx();
    if (argc >= 4 && !strcmp(argv[1], "modp")) return run_modp(atol(argv[2]), strtoull(argv[3], nullptr, 10));
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
