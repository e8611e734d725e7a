// Concurrent callers of hostcopy::copy at sizes around its thresholds, each
// copy checked byte for byte, and its scan (whether a byte copied is above 1)
// against a direct count on contents of 0s and 1s with at most one byte 2..255
// planted at a random place; prints "ok" and exits 0 when all match.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

#include "host_copy.hpp"

int main(int argc, char** argv) {
    const int callers = argc > 1 ? std::atoi(argv[1]) : 4;
    const int rounds = argc > 2 ? std::atoi(argv[2]) : 6;
    std::vector<int> bad(callers, 0);
    std::vector<std::thread> th;
    for (int c = 0; c < callers; ++c)
        th.emplace_back([&, c] {
            std::mt19937_64 rng(1234 + c);
            const size_t sizes[] = {0, 1, 63, hostcopy::PAR_COPY_MIN - 1, hostcopy::PAR_COPY_MIN,
                                    hostcopy::PAR_COPY_MIN + 4097, (size_t)9 << 20, ((size_t)21 << 20) + 13};
            for (int r = 0; r < rounds; ++r)
                for (size_t n : sizes) {
                    const size_t off = rng() % 61;  // unaligned source and destination
                    std::vector<uint8_t> src(n + off), dst(n + off + 8, 0xA5);
                    for (size_t i = 0; i < src.size(); ++i) src[i] = (uint8_t)(rng() >> 13);
                    hostcopy::copy(dst.data() + 3, src.data() + off, n);
                    for (size_t i = 0; i < n; ++i)
                        if (dst[3 + i] != src[off + i]) { ++bad[c]; break; }
                    // the scan: 0/1 contents, then one byte above 1 planted (or none)
                    for (size_t i = 0; i < src.size(); ++i) src[i] &= 1;
                    if (n && (r & 1)) src[off + rng() % n] = (uint8_t)(2 + rng() % 254);
                    bool want = false;
                    for (size_t i = 0; i < n; ++i) want |= src[off + i] > 1;
                    bool above = !want;
                    hostcopy::copy(dst.data() + 3, src.data() + off, n, &above);
                    if (above != want) ++bad[c];
                    // bytes outside the destination range stay untouched
                    if (dst[0] != 0xA5 || dst[1] != 0xA5 || dst[2] != 0xA5 || dst[3 + n] != 0xA5) ++bad[c];
                }
        });
    for (auto& t : th) t.join();
    int total = 0;
    for (int b : bad) total += b;
    std::printf(total ? "mismatch %d\n" : "ok\n", total);
    return total ? 1 : 0;
}
