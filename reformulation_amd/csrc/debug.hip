// The allocation wrapper's fault injector (include/reform_debug.h).
#include <atomic>

#include "common.hpp"
#include "reform_debug.h"

namespace reform {

namespace {
std::atomic<int64_t> g_fault_after{-1};  // allocations to let through first; -1: disarmed
std::atomic<int64_t> g_allocs{0};
}  // namespace

bool device_alloc_fault() {
    g_allocs.fetch_add(1, std::memory_order_relaxed);
    int64_t k = g_fault_after.load();
    while (k >= 0)
        if (g_fault_after.compare_exchange_weak(k, k - 1)) return k == 0;
    return false;
}

}  // namespace reform

extern "C" int reform_debug_fail_device_alloc(int64_t after) {
    reform::g_fault_after = after < 0 ? -1 : after;
    return 0;
}

extern "C" int64_t reform_debug_device_allocs(void) { return reform::g_allocs.load(); }
