// CPU check of onefft::CopyPool (onefft_amd/csrc/copy_pool.hpp): several threads copy buffers of
// random sizes and alignments through pools of 0, 1, 3 and 8 helpers at once; every copy must match
// its source byte for byte and leave the guard bytes around its destination untouched. Built and
// run by tests/test_copy_pool.py; exits 0 on success.
#include <cstdio>
#include <random>
#include <thread>
#include <vector>

#include "../onefft_amd/csrc/copy_pool.hpp"

namespace {

constexpr size_t kGuard = 64;
constexpr unsigned char kCanary = 0xA5;

bool caller(onefft::CopyPool &pool, unsigned seed, int copies) {
    std::mt19937_64 rng(seed);
    for (int c = 0; c < copies; c++) {
        // sizes from empty to ~9 MiB, biased towards the pool's 1 MiB pieces and their edges
        size_t bytes;
        switch (rng() % 4) {
        case 0: bytes = rng() % 4096; break;
        case 1: bytes = (size_t(1) << 20) * (1 + rng() % 8) + (rng() % 3) - 1; break;
        default: bytes = rng() % (size_t(9) << 20); break;
        }
        const size_t src_off = rng() % 16, dst_off = rng() % 16;
        std::vector<unsigned char> src(bytes + src_off), dst(bytes + dst_off + 2 * kGuard, kCanary);
        for (size_t i = 0; i < src.size(); i += 8) src[i] = (unsigned char)(rng() >> 56);
        for (size_t i = 1; i < src.size(); i += 8) src[i] = (unsigned char)(i * 131 + c);
        pool.copy(dst.data() + kGuard + dst_off, src.data() + src_off, bytes);
        if (std::memcmp(dst.data() + kGuard + dst_off, src.data() + src_off, bytes) != 0) {
            std::printf("mismatch: seed %u copy %d bytes %zu\n", seed, c, bytes);
            return false;
        }
        for (size_t i = 0; i < kGuard + dst_off; i++)
            if (dst[i] != kCanary) {
                std::printf("wrote before the destination: seed %u copy %d\n", seed, c);
                return false;
            }
        for (size_t i = kGuard + dst_off + bytes; i < dst.size(); i++)
            if (dst[i] != kCanary) {
                std::printf("wrote after the destination: seed %u copy %d\n", seed, c);
                return false;
            }
    }
    return true;
}

}  // namespace

int main(int argc, char **argv) {
    const int copies = argc > 1 ? std::atoi(argv[1]) : 40;  // per calling thread and pool
    for (int helpers : {0, 1, 3, 8}) {
        onefft::CopyPool pool(helpers);
        std::vector<std::thread> threads;
        std::vector<int> ok(4, 0);
        for (int t = 0; t < 4; t++)
            threads.emplace_back([&, t] { ok[t] = caller(pool, 1000u * helpers + t, copies); });
        for (auto &th : threads) th.join();
        for (int t = 0; t < 4; t++)
            if (!ok[t]) {
                std::printf("failed with %d helpers\n", helpers);
                return 1;
            }
        std::printf("%d helpers: ok\n", helpers);
    }
    return 0;
}
