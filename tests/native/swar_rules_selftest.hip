// The byte-parallel rules of the band walks (csrc/kernels_life.hpp life_rule,
// gen_rule, gen_rule_small on cells 0 to 3, life_intervals; csrc/gate.hpp checker_open, open_select), run on the
// host over every cell value and live count against the rules stated one cell
// at a time:
//   Life        (src/rules.jl:385-389)  c == 1 ? survive bit n : birth bit n
//   Generations (src/rules.jl:446-456)  0 -> birth bit n; 1 -> 1 on the survive
//               bit n, else 2 (C > 2) or 0; any other s -> s + 1 while
//               s + 1 < C, else 0 (Int8 arithmetic)
//   Checkerboard (src/schemes.jl:84-89) open when the 1-based index sum has the
//               gate's parity (0 even, 1 odd)
// with bit n of a mask read as 0 for n >= 64.  Prints "ok" and exits 0 when all
// agree.
#include <cstdio>
#include <random>
#include <vector>

#include "kernels_life.hpp"

static int bad = 0;

static int life_cell(uint64_t birth, uint64_t survive, int c, int n) {
    const uint64_t m = c == 1 ? survive : birth;
    return n < 64 ? (int)((m >> n) & 1u) : 0;
}

static int gen_cell(uint64_t birth, uint64_t survive, int C, int c, int n) {
    if (c == 0) return n < 64 ? (int)((birth >> n) & 1u) : 0;
    if (c == 1) return (n < 64 && ((survive >> n) & 1u)) ? 1 : (C > 2 ? 2 : 0);
    return c + 1 < C ? (int)(int8_t)(c + 1) : 0;
}

// Every (cell value, live count <= 120) pair, four to a dword, through life_rule
// and gen_rule with the given masks and state count.
static void check_rules(uint64_t birth, uint64_t survive, int C) {
    LifeArgs a{};
    const int nb = life_intervals(birth, a.b_lo, a.b_hi), ns = life_intervals(survive, a.s_lo, a.s_hi);
    if (nb < 0 || ns < 0) {
        std::printf("masks %llx / %llx: more than %d runs\n", (unsigned long long)birth,
                    (unsigned long long)survive, LIFE_MAX_IV);
        ++bad;
        return;
    }
    a.gen_C = C;
    a.gen_k = (uint32_t)(C < 129 ? 129 - C : 0) * 0x01010101u;
    a.gen_dying = (C > 2 ? 2u : 0u) * 0x01010101u;
    std::vector<std::pair<int, int>> cases;
    for (int c = -128; c < 128; ++c)
        for (int n = 0; n <= 120; ++n) cases.push_back({c, n});
    for (size_t k = 0; k < cases.size(); k += 4) {
        uint32_t cw[1] = {0}, nw[1] = {0}, c1[1], life[1], gen[1];
        for (int b = 0; b < 4; ++b) {
            const auto& e = cases[(k + b) % cases.size()];
            cw[0] |= (uint32_t)(uint8_t)(int8_t)e.first << (8 * b);
            nw[0] |= (uint32_t)e.second << (8 * b);
        }
        c1[0] = bytes_eq(cw[0], 0x01010101u);
        life_rule<1, LIFE_MAX_IV>(a, nw, c1, life);
        gen_rule<1, LIFE_MAX_IV>(a, nw, cw, gen);
        for (int b = 0; b < 4; ++b) {
            const auto& e = cases[(k + b) % cases.size()];
            const int wl = life_cell(birth, survive, e.first, e.second);
            const int wg = gen_cell(birth, survive, C, e.first, e.second);
            const int gl = (int)(int8_t)(uint8_t)(life[0] >> (8 * b)), gg = (int)(int8_t)(uint8_t)(gen[0] >> (8 * b));
            if (gl != wl || gg != wg) {
                if (bad < 20)
                    std::printf("B %llx S %llx C %d: cell %d, count %d: life %d (want %d), generations %d (want %d)\n",
                                (unsigned long long)birth, (unsigned long long)survive, C, e.first, e.second, gl, wl,
                                gg, wg);
                ++bad;
            }
        }
    }
}

// gen_rule_small against the cell rule on every (cell in 0..3, live count <= 120)
// pair, for the state counts it is used with (C <= 4); the live count's source,
// bit 0 without bit 1 of each byte, against cell == 1.
static void check_small(uint64_t birth, uint64_t survive, int C) {
    LifeArgs a{};
    if (life_intervals(birth, a.b_lo, a.b_hi) < 0 || life_intervals(survive, a.s_lo, a.s_hi) < 0) {
        ++bad;
        return;
    }
    a.gen_C = C;
    a.gen_k = (uint32_t)(129 - C) * 0x01010101u;
    a.gen_dying = (C > 2 ? 2u : 0u) * 0x01010101u;
    std::vector<std::pair<int, int>> cases;
    for (int c = 0; c < 4; ++c)
        for (int n = 0; n <= 120; ++n) cases.push_back({c, n});
    for (size_t k = 0; k < cases.size(); ++k) {
        uint32_t cw[1] = {0}, nw[1] = {0}, gen[1];
        for (int b = 0; b < 4; ++b) {
            const auto& e = cases[(k + 37 * b) % cases.size()];
            cw[0] |= (uint32_t)e.first << (8 * b);
            nw[0] |= (uint32_t)e.second << (8 * b);
        }
        gen_rule_small<1, LIFE_MAX_IV>(a, nw, cw, gen);
        if ((cw[0] & ~(cw[0] >> 1) & 0x01010101u) != bytes_eq(cw[0], 0x01010101u)) {
            std::printf("cells %08x: bit 0 without bit 1 is not cell == 1\n", cw[0]);
            ++bad;
        }
        for (int b = 0; b < 4; ++b) {
            const auto& e = cases[(k + 37 * b) % cases.size()];
            const int want = gen_cell(birth, survive, C, e.first, e.second);
            const int got = (int)(int8_t)(uint8_t)(gen[0] >> (8 * b));
            if (got != want) {
                if (bad < 20)
                    std::printf("B %llx S %llx C %d: cell %d, count %d: small-cell generations %d (want %d)\n",
                                (unsigned long long)birth, (unsigned long long)survive, C, e.first, e.second, got,
                                want);
                ++bad;
            }
        }
    }
}

// A mask of at most two runs of set bits below 64.
static uint64_t two_runs(std::mt19937_64& g) {
    uint64_t m = 0;
    const int runs = (int)(g() % 3);
    int at = 0;
    for (int r = 0; r < runs && at < 64; ++r) {
        const int lo = at + (int)(g() % 30), len = 1 + (int)(g() % 40);
        for (int c = lo; c < lo + len && c < 64; ++c) m |= 1ull << c;
        at = lo + len + 1;
    }
    return m;
}

int main() {
    // the presets Brian's Brain (B2/S/C3), Star Wars (B2/S345/C4), Conway as C = 2,
    // and state counts at and around the signed-byte limits
    check_rules(1ull << 2, 0, 3);
    check_rules(1ull << 2, (1ull << 3) | (1ull << 4) | (1ull << 5), 4);
    check_rules(1ull << 3, (1ull << 2) | (1ull << 3), 2);
    for (int C : {2, 3, 5, 25, 126, 127, 128, 129, 130, 255, 1000}) check_rules(0x0F0ull, 0x3Cull << 40, C);
    std::mt19937_64 g(2026);
    for (int k = 0; k < 40; ++k) check_rules(two_runs(g), two_runs(g), 2 + (int)(g() % 140));
    // masks with bits up to 63 (counts of 64 and more match nothing)
    check_rules(~0ull << 50, ~0ull, 7);
    // the rule of lattices holding only cells below 4: Brian's Brain, Star Wars,
    // C = 2, 3 and 4 with other masks, random masks
    check_small(1ull << 2, 0, 3);
    check_small(1ull << 2, (1ull << 3) | (1ull << 4) | (1ull << 5), 4);
    check_small(1ull << 3, (1ull << 2) | (1ull << 3), 2);
    check_small(0x0F0ull, 0x3Cull << 40, 3);
    check_small(~0ull << 50, ~0ull, 4);
    for (int k = 0; k < 60; ++k) check_small(two_runs(g), two_runs(g), 2 + (int)(g() % 3));

    // a third run does not fit
    uint32_t lo[LIFE_MAX_IV], hi[LIFE_MAX_IV];
    if (life_intervals(0b10101ull, lo, hi) != -1) {
        std::printf("three runs were accepted\n");
        ++bad;
    }

    // Checkerboard over the cells x + b (x a multiple of 4) of rows y, both gates
    for (int gate = 0; gate < 2; ++gate)
        for (int y = 0; y < 6; ++y)
            for (int x = 0; x < 16; x += 4) {
                const uint32_t o = checker_open(y, gate);
                for (int b = 0; b < 4; ++b) {
                    const int want = (((x + b + 1) + (y + 1)) & 1) == gate;
                    if ((int)((o >> (8 * b)) & 0xFFu) != want) {
                        std::printf("checkerboard gate %d, cell (%d, %d): %u\n", gate, x + b, y,
                                    (unsigned)((o >> (8 * b)) & 0xFFu));
                        ++bad;
                    }
                }
            }

    // open_select: bytes of `out` where open, of `keep` elsewhere
    for (uint32_t open = 0; open < 16; ++open) {
        uint32_t ob = 0;
        for (int b = 0; b < 4; ++b) ob |= ((open >> b) & 1u) << (8 * b);
        const uint32_t out = 0xA1B2C3D4u, keep = 0x5E6F7081u, r = open_select(ob, out, keep);
        for (int b = 0; b < 4; ++b) {
            const uint32_t want = (((open >> b) & 1u) ? out : keep) >> (8 * b) & 0xFFu;
            if (((r >> (8 * b)) & 0xFFu) != want) {
                std::printf("open_select %x byte %d\n", open, b);
                ++bad;
            }
        }
    }
    if (bad) {
        std::printf("%d mismatches\n", bad);
        return 1;
    }
    std::printf("ok\n");
    return 0;
}
