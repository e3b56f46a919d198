// HostCopy (vqm_amd/csrc/ss2_hostcopy.cc) against std::memcpy: sizes around
// the chunking thresholds, unaligned ends, several callers at once, and a
// forked child (which copies alone). Exits 1 on any difference.
#include <sys/wait.h>
#include <unistd.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "ss2_hostcopy.h"

namespace {

std::atomic<int> g_failures{0};

void CheckCopy(uint64_t bytes, uint64_t src_off, uint64_t dst_off, uint32_t seed) {
    std::vector<uint8_t> src(bytes + src_off + 64), dst(bytes + dst_off + 64, 0xA5), want;
    std::mt19937 rng(seed);
    for (auto& v : src) v = static_cast<uint8_t>(rng());
    want = dst;
    std::memcpy(want.data() + dst_off, src.data() + src_off, bytes);
    vqm::ss2::HostCopy(dst.data() + dst_off, src.data() + src_off, bytes);
    if (dst != want) {
        std::printf("mismatch: %llu bytes, offsets %llu/%llu\n", static_cast<unsigned long long>(bytes),
                    static_cast<unsigned long long>(src_off), static_cast<unsigned long long>(dst_off));
        ++g_failures;
    }
}

}  // namespace

int main() {
    const uint64_t MiB = 1ull << 20;
    std::printf("threads per copy: %u\n", vqm::ss2::HostCopyThreads());
    const uint64_t sizes[] = {0, 1, 4095, MiB, 2 * MiB - 1, 2 * MiB, 2 * MiB + 1, 3 * MiB + 7,
                              4 * MiB, 6220800, 12441600 + 13, 24883200};
    uint32_t seed = 1;
    for (uint64_t n : sizes)
        for (uint64_t off : {0ull, 3ull})
            CheckCopy(n, off, off == 0 ? 0 : 5, seed++);
    // Several callers at once share the helpers.
    std::vector<std::thread> callers;
    for (int t = 0; t < 6; ++t)
        callers.emplace_back([t] {
            for (int k = 0; k < 8; ++k) CheckCopy(6220800 + 4096 * k + t, t, 2 * t, 100 + 10 * t + k);
        });
    for (auto& c : callers) c.join();
    // A forked child copies on its own thread.
    const pid_t pid = fork();
    if (pid == 0) {
        CheckCopy(12441600, 1, 2, 7);
        _exit(g_failures.load() == 0 ? 0 : 1);
    }
    int status = 0;
    if (pid < 0 || waitpid(pid, &status, 0) != pid || !WIFEXITED(status) || WEXITSTATUS(status) != 0) {
        std::printf("forked child failed\n");
        ++g_failures;
    }
    std::printf("failures: %d\n", g_failures.load());
    return g_failures.load() == 0 ? 0 : 1;
}
