// The shape-tuning table of include/reform_tune.h.
#include "reform_tune.h"
#include "praq_host.hpp"
#include "tune.hpp"

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>

namespace reform {

namespace {

constexpr int kLaneBuckets = 8;
std::atomic<int64_t> lane_table[kLaneBuckets];        // "praq.lanes.<b>"
std::atomic<int64_t> batch_lane_table[kLaneBuckets];  // "batch.praq.lanes.<b>"
std::atomic<int64_t> pria_knobs[kPriaKnobs];          // "batch.pria.*"

const char* const kPriaKnobNames[kPriaKnobs] = {"batch.pria.x_threads", "batch.pria.x_hold",
                                                "batch.pria.yz_shape"};

// bucket b of "<prefix><b>", or -1
int bucket_key(const char* key, const char* prefix) {
    const size_t np = std::strlen(prefix);
    if (!key || std::strncmp(key, prefix, np) != 0) return -1;
    const char* tail = key + np;
    if (tail[0] < '0' || tail[0] > '7' || tail[1] != '\0') return -1;
    return tail[0] - '0';
}

int knob_key(const char* key) {
    if (!key) return -1;
    for (int k = 0; k < kPriaKnobs; ++k)
        if (std::strcmp(key, kPriaKnobNames[k]) == 0) return k;
    return -1;
}

bool lanes_ok(int64_t v) { return v == 0 || (v >= 1 && v <= 64 && (v & (v - 1)) == 0); }

bool knob_ok(int k, int64_t v) {
    if (v == 0) return true;
    switch (k) {
        case kPriaXThreads: return v == 256 || v == 512;
        case kPriaXHold: return v == 1 || v == 2;
        case kPriaYZShape: return v == 1 || v == 2;
    }
    return false;
}

// the slot a key names, or null
std::atomic<int64_t>* slot(const char* key, int* knob) {
    *knob = -1;
    int b = bucket_key(key, "praq.lanes.");
    if (b >= 0) return &lane_table[b];
    b = bucket_key(key, "batch.praq.lanes.");
    if (b >= 0) return &batch_lane_table[b];
    const int k = knob_key(key);
    if (k >= 0) {
        *knob = k;
        return &pria_knobs[k];
    }
    return nullptr;
}

int degree_bucket(double mean_degree) {
    int b = 0;
    while (b + 1 < kLaneBuckets && mean_degree >= double(int64_t(1) << (b + 1))) ++b;
    return b;
}

}  // namespace

int praq_tuned_lanes(double mean_degree) {
    return int(lane_table[degree_bucket(mean_degree)].load(std::memory_order_relaxed));
}

int batch_praq_lanes(int n, int m) {
    const int b = degree_bucket(2.0 * double(m) / double(n > 0 ? n : 1));
    const int tuned = int(batch_lane_table[b].load(std::memory_order_relaxed));
    return tuned > 0 ? tuned : praq_lanes_for(n, m);
}

int pria_knob(int which) {
    return int(pria_knobs[which].load(std::memory_order_relaxed));
}

}  // namespace reform

using namespace reform;

extern "C" int reform_tune_set(const char* key, int64_t value) {
    int k;
    std::atomic<int64_t>* s = slot(key, &k);
    if (!s) return -1;
    if (k >= 0 ? !knob_ok(k, value) : !lanes_ok(value)) return -1;
    s->store(value, std::memory_order_relaxed);
    return 0;
}

extern "C" int64_t reform_tune_get(const char* key) {
    int k;
    std::atomic<int64_t>* s = slot(key, &k);
    return s ? s->load(std::memory_order_relaxed) : -1;
}

extern "C" void reform_tune_clear(void) {
    for (auto& v : lane_table) v.store(0, std::memory_order_relaxed);
    for (auto& v : batch_lane_table) v.store(0, std::memory_order_relaxed);
    for (auto& v : pria_knobs) v.store(0, std::memory_order_relaxed);
}

extern "C" int reform_praq_lanes_for(int64_t n, int64_t m) {
    if (n <= 0 || m < 0) return -1;
    return praq_device_lanes(int(n), int(m));
}

extern "C" int reform_batch_praq_lanes_for(int64_t n, int64_t m) {
    if (n <= 0 || m < 0) return -1;
    return batch_praq_lanes(int(n), int(m));
}
