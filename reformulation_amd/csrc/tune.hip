// The shape-tuning table of include/reform_tune.h.
#include "reform_tune.h"
#include "praq_host.hpp"

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace reform {

namespace {

constexpr int kLaneBuckets = 8;
std::atomic<int64_t> lane_table[kLaneBuckets];

// bucket of "praq.lanes.<b>", or -1
int lane_key(const char* key) {
    static const char prefix[] = "praq.lanes.";
    if (!key || std::strncmp(key, prefix, sizeof(prefix) - 1) != 0) return -1;
    const char* tail = key + sizeof(prefix) - 1;
    if (tail[0] < '0' || tail[0] > '7' || tail[1] != '\0') return -1;
    return tail[0] - '0';
}

}  // namespace

int praq_tuned_lanes(double mean_degree) {
    int b = 0;
    while (b + 1 < kLaneBuckets && mean_degree >= double(int64_t(1) << (b + 1))) ++b;
    return int(lane_table[b].load(std::memory_order_relaxed));
}

}  // namespace reform

using namespace reform;

extern "C" int reform_tune_set(const char* key, int64_t value) {
    const int b = lane_key(key);
    if (b < 0) return -1;
    if (value != 0 && (value < 1 || value > 64 || (value & (value - 1)) != 0)) return -1;
    lane_table[b].store(value, std::memory_order_relaxed);
    return 0;
}

extern "C" int64_t reform_tune_get(const char* key) {
    const int b = lane_key(key);
    return b < 0 ? -1 : lane_table[b].load(std::memory_order_relaxed);
}

extern "C" void reform_tune_clear(void) {
    for (auto& v : lane_table) v.store(0, std::memory_order_relaxed);
}

extern "C" int reform_praq_lanes_for(int64_t n, int64_t m) {
    if (n <= 0 || m < 0) return -1;
    return praq_device_lanes(int(n), int(m));
}
