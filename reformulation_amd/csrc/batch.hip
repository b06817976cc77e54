// Ragged batches of PRAQ and PRIA instances: include/reform_batch.h.
//
// Device memory: one allocation holds every instance's image (inputs first,
// iteration state after, 256-byte aligned) -- in FP64 always and, for an FP32
// batch, in FP32 too -- the job records the kernels read and the scalar
// states.  Three ways to solve the resident instances:
//   batched     the wide engine (wide.hpp): every kernel launch covers the
//               work of every instance of a family through a block table,
//               instances keep their own Halpern state on the device, and
//               the launches of one check interval form a hipGraph replayed
//               until every instance has finished; PRAQ and PRIA run on two
//               streams driven from two host threads;
//   workgroups  one persistent launch per family, one workgroup per
//               instance running its whole solve, largest first;
//   sequential  the plain solver: per instance the tier and precision path
//               praq_solve_ex / pria_solve_ex would take.
// With FP32 iterates, an instance whose gap stalls stops with its Halpern
// state intact; its state is converted to FP64 and it continues there (a
// second wide engine over the stalled subset, or one more launch).
#include "reform_batch.h"
#include "praq_host.hpp"
#include "pria_host.hpp"
#include "tune.hpp"
#include "wide.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <thread>
#include <vector>

using namespace reform;

namespace {

constexpr double kWgLimitSeconds = 600.0;

double clock_rate_hz() {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0.0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess) return 0.0;
    return double(khz) * 1000.0;
}

int64_t env_int(const char* name, int64_t fallback) {
    const char* v = std::getenv(name);
    return v ? std::atoll(v) : fallback;
}

// parallel for over [0, count) on host threads
template <class F>
void parallel_for(size_t count, F f) {
    unsigned workers = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (count < 2 || workers == 1) {
        for (size_t q = 0; q < count; ++q) f(q);
        return;
    }
    std::vector<std::thread> threads;
    std::atomic<size_t> next(0);
    for (unsigned w = 0; w < workers; ++w)
        threads.emplace_back([&] {
            for (size_t q; (q = next.fetch_add(1)) < count;) f(q);
        });
    for (auto& t : threads) t.join();
}

struct PraqEntry {
    PraqInstance copy;  // pointers into the owned storage below
    std::vector<double> diag, vals;
    std::vector<int64_t> rows, cols;
    LogicParams lp;
    PraqOptions o;
    PraqHostInstance host;
    size_t off32 = 0, off64 = 0;  // image offsets in the device allocation
    size_t bytes = 0;             // the images' device bytes
    int tier = 0;
};

struct PriaEntry {
    int64_t n = 0, t = 0;
    std::vector<double> k, h;
    double k0 = 0.0;
    LogicParams lp;
    pria_options o;
    PriaHostInstance host;
    size_t off32 = 0, off64 = 0;
    size_t bytes = 0;
    int tier = 0;
};

// The job records of one family in one precision, in launch order, host and
// device copies; `sub` is device scratch for the stalled subset.
template <class Job>
struct JobSet {
    std::vector<Job> host;
    Job* dev = nullptr;
    Job* sub = nullptr;
};

template <class Real>
int read_array(std::vector<double>& out, const Real* src, size_t count) {
    std::vector<Real> v(count);
    if (count && hipMemcpy(v.data(), src, sizeof(Real) * count, hipMemcpyDeviceToHost) != hipSuccess)
        return kErrorDevice;
    out.assign(v.begin(), v.end());
    return 0;
}

struct BatchImpl {
    int use_float32 = 0;
    int praq_block = 256;
    std::vector<PraqEntry> praq;
    std::vector<PriaEntry> pria;
    void* dev = nullptr;
    size_t dev_bytes = 0;
    hipStream_t s0 = nullptr, s1 = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double praq_ms = 0.0, pria_ms = 0.0;
    bool uploaded = false;

    std::vector<int64_t> praq_order, pria_order, praq_slot, pria_slot;
    JobSet<PraqJob<float>> praq32;
    JobSet<PraqJob<double>> praq64;
    JobSet<PriaJob<float>> pria32;
    JobSet<PriaJob<double>> pria64;
    LogicState* praq_states_dev = nullptr;
    LogicState* pria_states_dev = nullptr;
    std::vector<LogicState> praq_states, pria_states;
    std::vector<char> praq_in64, pria_in64;  // final point lives in the FP64 image
    std::vector<double> praq_seconds, pria_seconds;  // per instance, last sequential solve
    double* work = nullptr;
    int pria_max_T = 2;
    int device = 0;
    double time_limit = 0.0;

    // The wide engines of the batched mode: per family, `groups[f]` engines
    // over contiguous slot ranges of balanced work (upload deals the
    // size-sorted instances round the groups), each on its own stream, so
    // the launches of one group can overlap another's (REFORM_PRAQ_GROUPS,
    // REFORM_PRIA_GROUPS; one each by default: more measured within 3% on
    // the 1024-instance mixture).  Instances switch precision inside their
    // engine.
    int groups[2] = {1, 1};
    std::vector<std::unique_ptr<WideControl>> wide[2];
    std::vector<int> group_first[2];           // slot range of group g: [first[g], first[g+1])
    std::vector<hipStream_t> group_streams;    // streams beyond s0 / s1
    std::vector<hipEvent_t> group_events;      // (start, end) per engine
    bool timing = false;
    bool compaction = true;

    hipStream_t stream_for(int family, int g) {
        if (g == 0) return family == 0 ? s0 : s1;
        const size_t k = size_t(family) * 8 + size_t(g);
        if (group_streams.size() <= k) group_streams.resize(k + 1, nullptr);
        if (!group_streams[k] &&
            hipStreamCreateWithFlags(&group_streams[k], hipStreamNonBlocking) != hipSuccess)
            group_streams[k] = nullptr;
        return group_streams[k];
    }
    void set_timing(bool on) {
        timing = on;
        for (auto& family : wide)
            for (auto& e : family) e->set_timing(on);
    }

    // every stream of the batch idle, so that nothing of a solve still runs
    // when its memory goes
    void drain() {
        for (hipStream_t gs : group_streams)
            if (gs) (void)hipStreamSynchronize(gs);
        if (s0) (void)hipStreamSynchronize(s0);
        if (s1) (void)hipStreamSynchronize(s1);
    }

    ~BatchImpl() {
        drain();
        for (auto& family : wide) family.clear();
        for (hipStream_t gs : group_streams)
            if (gs) (void)hipStreamDestroy(gs);
        for (hipEvent_t e : group_events)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (dev) (void)hipFree(dev);
        if (s0) (void)hipStreamDestroy(s0);
        if (s1) (void)hipStreamDestroy(s1);
    }

    int upload();
    int solve(int mode);
    int solve_wide();
    std::vector<double> deadlines(int family, int lo, int hi) const;
    int solve_workgroups();
    int solve_sequential();
    int praq_result(int64_t q, double* x, double* y, PraqInfo* info);
    int pria_result(int64_t q, double* x, double* y, double* z, pria_info* info);

    // one plain PRAQ solve of launch slot s from job `job` (its precision),
    // with the lanes of the instance's host build (`lanes`): the plain
    // solver's lane rule, and the lanes its hub lists were built for, which
    // the device tier's hub threshold must match (the job records carry the
    // batched tiers' lanes)
    template <class Real>
    int praq_plain(const PraqJob<Real>& job, PraqJob<Real>* devslot, int tier, bool resume,
                   double time_limit, int lanes) {
        PraqJob<Real> j = job;
        j.lanes = lanes;
        if (tier == PRAQ_TIER_WORKGROUP) {
            j.resume = resume ? 1 : 0;
            REFORM_HIP_CHECK(hipMemcpyAsync(devslot, &j, sizeof(j), hipMemcpyHostToDevice, s0));
            if (praq_launch_wg<Real>(devslot, 1, 1024, s0)) return kErrorDevice;
            return hipStreamSynchronize(s0) == hipSuccess ? 0 : kErrorDevice;
        }
        LogicState st;
        if (resume && hipMemcpy(&st, job.state, sizeof(st), hipMemcpyDeviceToHost) != hipSuccess)
            return kErrorDevice;
        if (praq_device_solve<Real>(j, st, work, time_limit, s0, resume)) return kErrorDevice;
        REFORM_HIP_CHECK(hipMemcpyAsync(job.state, &st, sizeof(st), hipMemcpyHostToDevice, s0));
        return hipStreamSynchronize(s0) == hipSuccess ? 0 : kErrorDevice;
    }

    template <class Real>
    int pria_plain(const PriaJob<Real>& job, PriaJob<Real>* devslot, int tier, bool resume, int T,
                   double time_limit) {
        if (tier == PRIA_TIER_WORKGROUP) {
            PriaJob<Real> j = job;
            j.resume = resume ? 1 : 0;
            REFORM_HIP_CHECK(hipMemcpyAsync(devslot, &j, sizeof(j), hipMemcpyHostToDevice, s1));
            if (pria_launch_wg<Real>(devslot, 1, T, s1)) return kErrorDevice;
            return hipStreamSynchronize(s1) == hipSuccess ? 0 : kErrorDevice;
        }
        LogicState st;
        if (resume && hipMemcpy(&st, job.state, sizeof(st), hipMemcpyDeviceToHost) != hipSuccess)
            return kErrorDevice;
        if (pria_device_solve<Real>(job, st, work, time_limit, s1, resume)) return kErrorDevice;
        REFORM_HIP_CHECK(hipMemcpyAsync(job.state, &st, sizeof(st), hipMemcpyHostToDevice, s1));
        return hipStreamSynchronize(s1) == hipSuccess ? 0 : kErrorDevice;
    }
};

int BatchImpl::upload() {
    if (uploaded) return kErrorArgument;
    std::vector<int> fail(praq.size() + pria.size(), 0);
    parallel_for(praq.size(), [&](size_t q) {
        PraqEntry& e = praq[q];
        fail[q] = praq_build_host(e.copy, e.o.dense_degree, nullptr, nullptr, e.host);
    });
    parallel_for(pria.size(), [&](size_t q) {
        PriaEntry& e = pria[q];
        fail[praq.size() + q] = pria_build_host(e.n, e.t, e.k.data(), e.k0, e.h.data(), nullptr,
                                                nullptr, nullptr, e.host);
    });
    for (int f : fail)
        if (f) return kErrorInstance;

    // launch order: largest first
    praq_order.resize(praq.size());
    std::iota(praq_order.begin(), praq_order.end(), 0);
    std::stable_sort(praq_order.begin(), praq_order.end(), [&](int64_t a, int64_t b) {
        return int64_t(praq[a].host.n) + 2 * int64_t(praq[a].host.m) >
               int64_t(praq[b].host.n) + 2 * int64_t(praq[b].host.m);
    });
    pria_order.resize(pria.size());
    std::iota(pria_order.begin(), pria_order.end(), 0);
    std::stable_sort(pria_order.begin(), pria_order.end(),
                     [&](int64_t a, int64_t b) { return pria[a].host.ND > pria[b].host.ND; });
    // the groups of the batched mode: size-sorted instances dealt round the
    // groups, each group a contiguous slot range
    groups[0] = int(std::max<int64_t>(1, std::min<int64_t>(8, env_int("REFORM_PRAQ_GROUPS", 1))));
    groups[1] = int(std::max<int64_t>(1, std::min<int64_t>(8, env_int("REFORM_PRIA_GROUPS", 1))));
    auto deal = [](std::vector<int64_t>& order, int g, std::vector<int>& first) {
        std::vector<int64_t> dealt;
        first.assign(1, 0);
        for (int k = 0; k < g; ++k) {
            for (size_t s = size_t(k); s < order.size(); s += size_t(g)) dealt.push_back(order[s]);
            first.push_back(int(dealt.size()));
        }
        order.swap(dealt);
    };
    deal(praq_order, groups[0], group_first[0]);
    deal(pria_order, groups[1], group_first[1]);
    praq_slot.resize(praq.size());
    pria_slot.resize(pria.size());
    for (size_t s = 0; s < praq_order.size(); ++s) praq_slot[praq_order[s]] = int64_t(s);
    for (size_t s = 0; s < pria_order.size(); ++s) pria_slot[pria_order[s]] = int64_t(s);

    // one allocation: FP64 images, FP32 images (FP32 batch), job records,
    // states, one device-tier work area
    const bool f32 = use_float32 != 0;
    int64_t max_n = 0, max_m = 0;   // the work area serves every instance in turn
    for (const auto& e : praq) {
        max_n = std::max<int64_t>(max_n, e.host.n);
        max_m = std::max<int64_t>(max_m, e.host.m);
    }
    size_t at = 0, work_bytes = praq_device_work_bytes(max_n, max_m);
    std::vector<PraqLayout> pl64(praq.size()), pl32(praq.size());
    std::vector<PriaLayout> rl64(pria.size()), rl32(pria.size());
    for (size_t q = 0; q < praq.size(); ++q) {
        pl64[q] = praq_layout(praq[q].host, sizeof(double));
        praq[q].off64 = at;
        at += align256(pl64[q].total_bytes);
        if (f32) {
            pl32[q] = praq_layout(praq[q].host, sizeof(float));
            praq[q].off32 = at;
            at += align256(pl32[q].total_bytes);
        }
        praq[q].bytes = at - praq[q].off64;
    }
    for (size_t q = 0; q < pria.size(); ++q) {
        rl64[q] = pria_layout(pria[q].host.N, pria[q].host.T, pria[q].host.ND, sizeof(double));
        pria[q].off64 = at;
        at += align256(rl64[q].total_bytes);
        if (f32) {
            rl32[q] = pria_layout(pria[q].host.N, pria[q].host.T, pria[q].host.ND, sizeof(float));
            pria[q].off32 = at;
            at += align256(rl32[q].total_bytes);
        }
        pria[q].bytes = at - pria[q].off64;
        work_bytes = std::max(work_bytes, pria_device_work_bytes(pria[q].host.N, pria[q].host.T));
        pria_max_T = std::max(pria_max_T, pria[q].host.T);
    }
    const size_t np = praq.size() + 1, nr = pria.size() + 1;
    const size_t off_p32 = at;
    at = align256(at + 2 * sizeof(PraqJob<float>) * np);
    const size_t off_p64 = at;
    at = align256(at + 2 * sizeof(PraqJob<double>) * np);
    const size_t off_r32 = at;
    at = align256(at + 2 * sizeof(PriaJob<float>) * nr);
    const size_t off_r64 = at;
    at = align256(at + 2 * sizeof(PriaJob<double>) * nr);
    const size_t off_ps = at;
    at = align256(at + sizeof(LogicState) * np);
    const size_t off_rs = at;
    at = align256(at + sizeof(LogicState) * nr);
    const size_t off_work = at;
    at = align256(at + work_bytes);
    dev_bytes = at;
    if (device_malloc(&dev, dev_bytes) != hipSuccess) {
        dev = nullptr;
        return kErrorDevice;
    }
    REFORM_HIP_CHECK(hipGetDevice(&device));
    if (hipStreamCreateWithFlags(&s0, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s1, hipStreamNonBlocking) != hipSuccess)
        return kErrorDevice;
    for (auto& e : ev)
        REFORM_HIP_CHECK(hipEventCreate(&e));
    char* base = static_cast<char*>(dev);
    praq32.dev = reinterpret_cast<PraqJob<float>*>(base + off_p32);
    praq32.sub = praq32.dev + np;
    praq64.dev = reinterpret_cast<PraqJob<double>*>(base + off_p64);
    praq64.sub = praq64.dev + np;
    pria32.dev = reinterpret_cast<PriaJob<float>*>(base + off_r32);
    pria32.sub = pria32.dev + nr;
    pria64.dev = reinterpret_cast<PriaJob<double>*>(base + off_r64);
    pria64.sub = pria64.dev + nr;
    praq_states_dev = reinterpret_cast<LogicState*>(base + off_ps);
    pria_states_dev = reinterpret_cast<LogicState*>(base + off_rs);
    work = reinterpret_cast<double*>(base + off_work);

    // inputs: each image built into a host staging area and copied
    size_t stage = 0;
    for (size_t q = 0; q < praq.size(); ++q) stage = std::max(stage, pl64[q].input_bytes);
    for (size_t q = 0; q < pria.size(); ++q)
        stage = std::max(stage, pria_upload_bytes(pria[q].host, rl64[q]));
    std::vector<char> image(stage);
    // the workgroup tier's own clock limit: the instance's, at most kWgLimitSeconds
    const double rate = clock_rate_hz();
    auto ticks_for = [&](double limit) {
        const double sec = limit > 0.0 ? std::min(limit, kWgLimitSeconds) : kWgLimitSeconds;
        return (long long)(sec * rate);
    };
    const int64_t praq_wg = env_int("REFORM_PRAQ_WG_MAX", int64_t(1) << 17);
    const int64_t pria_wg = env_int("REFORM_PRIA_WG_MAX", int64_t(1) << 15);

    praq64.host.assign(praq.size(), PraqJob<double>());
    if (f32) praq32.host.assign(praq.size(), PraqJob<float>());
    for (size_t s = 0; s < praq_order.size(); ++s) {
        const int64_t q = praq_order[s];
        PraqEntry& e = praq[q];
        e.tier = e.o.tier != PRAQ_TIER_AUTO
                     ? e.o.tier
                     : (int64_t(e.host.n) + e.host.m <= praq_wg ? PRAQ_TIER_WORKGROUP
                                                                : PRAQ_TIER_DEVICE);
        auto fill = [&](auto& job, const PraqLayout& L, size_t off, auto tag) -> int {
            using Real = decltype(tag);
            std::fill(image.begin(), image.begin() + L.input_bytes, 0);
            praq_fill_image<Real>(e.host, L, image.data());
            if (hipMemcpy(base + off, image.data(), L.input_bytes, hipMemcpyHostToDevice) !=
                hipSuccess)
                return kErrorDevice;
            std::memset(&job, 0, sizeof(job));
            job.a = praq_bind<Real>(L, base + off, e.host);
            job.p = e.lp;
            job.scale = e.host.scale;
            job.lanes = batch_praq_lanes(e.host.n, e.host.m);
            job.time_limit_ticks = ticks_for(e.o.time_limit);
            job.state = praq_states_dev + s;
            return 0;
        };
        if (fill(praq64.host[s], pl64[q], e.off64, double())) return kErrorDevice;
        if (f32) {
            if (fill(praq32.host[s], pl32[q], e.off32, float())) return kErrorDevice;
            apply_fp32_policy(praq32.host[s].p);
        }
    }
    pria64.host.assign(pria.size(), PriaJob<double>());
    if (f32) pria32.host.assign(pria.size(), PriaJob<float>());
    for (size_t s = 0; s < pria_order.size(); ++s) {
        const int64_t q = pria_order[s];
        PriaEntry& e = pria[q];
        e.tier = e.o.tier != PRIA_TIER_AUTO
                     ? e.o.tier
                     : (e.host.ND <= pria_wg && e.host.T <= kPriaWgMaxT ? PRIA_TIER_WORKGROUP
                                                                         : PRIA_TIER_DEVICE);
        auto fill = [&](auto& job, const PriaLayout& L, size_t off, auto tag) -> int {
            using Real = decltype(tag);
            size_t up = pria_upload_bytes(e.host, L);
            std::fill(image.begin(), image.begin() + up, 0);
            pria_fill_image<Real>(e.host, L, image.data());
            REFORM_HIP_CHECK(hipMemcpy(base + off, image.data(), up, hipMemcpyHostToDevice));
            std::memset(&job, 0, sizeof(job));
            job.a = pria_bind<Real>(L, base + off, e.host);
            if (e.host.cold && pria_cold_start<Real>(job.a, s1)) return kErrorDevice;
            job.p = e.lp;
            job.scale = e.host.scale;
            job.time_limit_ticks = ticks_for(e.o.time_limit);
            job.state = pria_states_dev + s;
            return 0;
        };
        if (fill(pria64.host[s], rl64[q], e.off64, double())) return kErrorDevice;
        if (f32) {
            if (fill(pria32.host[s], rl32[q], e.off32, float())) return kErrorDevice;
            apply_fp32_policy(pria32.host[s].p);
        }
    }
    auto put = [&](auto& set) -> int {
        if (set.host.empty()) return 0;
        return hipMemcpy(set.dev, set.host.data(), sizeof(set.host[0]) * set.host.size(),
                         hipMemcpyHostToDevice) == hipSuccess
                   ? 0
                   : kErrorDevice;
    };
    if (put(praq64) || put(pria64)) return kErrorDevice;
    if (f32 && (put(praq32) || put(pria32))) return kErrorDevice;
    REFORM_HIP_CHECK(hipStreamSynchronize(s1));
    praq_states.assign(praq.size(), LogicState());
    pria_states.assign(pria.size(), LogicState());
    praq_in64.assign(praq.size(), 1);
    pria_in64.assign(pria.size(), 1);
    uploaded = true;
    return 0;
}

int BatchImpl::solve(int mode) {
    if (!uploaded) return kErrorArgument;
    praq_ms = pria_ms = 0.0;
    int st = kErrorArgument;
    if (mode == REFORM_BATCHED) st = solve_wide();
    else if (mode == REFORM_SEQUENTIAL) st = solve_sequential();
    else if (mode == REFORM_WORKGROUPS) st = solve_workgroups();
    if (st) {
        // a failed solve (a device call or allocation refused part way)
        // leaves no work running, and the next solve sets its engines up anew
        drain();
        for (auto& family : wide) family.clear();
        return st;
    }
    if (!praq.empty() &&
        hipMemcpyAsync(praq_states.data(), praq_states_dev, sizeof(LogicState) * praq.size(),
                       hipMemcpyDeviceToHost, s0) != hipSuccess)
        return kErrorDevice;
    if (!pria.empty() &&
        hipMemcpyAsync(pria_states.data(), pria_states_dev, sizeof(LogicState) * pria.size(),
                       hipMemcpyDeviceToHost, s1) != hipSuccess)
        return kErrorDevice;
    if (hipStreamSynchronize(s0) != hipSuccess || hipStreamSynchronize(s1) != hipSuccess)
        return kErrorDevice;
    return 0;
}

// The instances' own time limits (options.time_limit, seconds from the start
// of the solve) for the slots [lo, hi) of a family; empty when none has one.
std::vector<double> BatchImpl::deadlines(int family, int lo, int hi) const {
    std::vector<double> d(size_t(hi - lo), std::numeric_limits<double>::infinity());
    bool any = false;
    for (int s = lo; s < hi; ++s) {
        const double limit = family == 0 ? praq[praq_order[s]].o.time_limit
                                         : pria[pria_order[s]].o.time_limit;
        if (limit > 0.0) {
            d[s - lo] = limit;
            any = true;
        }
    }
    if (!any) d.clear();
    return d;
}

// Both families at once from this thread: PRAQ on s0, PRIA on s1, each one
// wide engine whose instances switch to FP64 on their own.  The runs are
// advanced without blocking, so neither stream waits for the other's host
// work.
int BatchImpl::solve_wide() {
    const bool f32 = use_float32 != 0;
    const int np = int(praq.size()), nr = int(pria.size());
    const int counts[2] = {np, nr};
    for (int f = 0; f < 2; ++f) {
        if (!counts[f] || !wide[f].empty()) continue;
        for (int g = 0; g < groups[f]; ++g) {
            const int lo = group_first[f][g], hi = group_first[f][g + 1];
            if (lo == hi) continue;
            if (f == 0) {
                auto e = std::make_unique<WidePraq>();
                e->set_timing(timing);
                std::vector<PraqJob<double>> sub(praq64.host.begin() + lo, praq64.host.begin() + hi);
                if (e->setup(std::move(sub), f32 ? praq32.dev + lo : nullptr, praq64.dev + lo))
                    return kErrorDevice;
                e->set_deadlines(deadlines(0, lo, hi));
                wide[0].push_back(std::move(e));
            } else {
                auto e = std::make_unique<WidePria>();
                e->set_timing(timing);
                std::vector<PriaJob<double>> sub(pria64.host.begin() + lo, pria64.host.begin() + hi);
                if (e->setup(std::move(sub), f32 ? pria32.dev + lo : nullptr, pria64.dev + lo))
                    return kErrorDevice;
                e->set_deadlines(deadlines(1, lo, hi));
                wide[1].push_back(std::move(e));
            }
        }
    }
    struct Lane {
        WideControl* engine;
        hipStream_t stream;
        int family;
        hipEvent_t begin, end;
        WideRun run;
        bool done;
    };
    std::vector<Lane> lanes;
    for (int f = 0; f < 2; ++f)
        for (size_t g = 0; g < wide[f].size(); ++g) {
            hipStream_t st = stream_for(f, int(g));
            if (!st) return kErrorDevice;
            lanes.push_back(Lane{wide[f][g].get(), st, f, nullptr, nullptr, WideRun(), false});
        }
    while (group_events.size() < 2 * lanes.size()) {
        hipEvent_t e;
        REFORM_HIP_CHECK(hipEventCreate(&e));
        group_events.push_back(e);
    }
    for (size_t k = 0; k < lanes.size(); ++k) {
        Lane& l = lanes[k];
        l.begin = group_events[2 * k];
        l.end = group_events[2 * k + 1];
        l.engine->clear_timing();
        REFORM_HIP_CHECK(hipEventRecord(l.begin, l.stream));
        l.run.shrink = compaction;
        if (l.run.begin(l.engine, l.stream, time_limit)) return kErrorDevice;
    }
    for (;;) {
        int changed = 0, running = 0;
        for (Lane& l : lanes) {
            if (l.done) continue;
            const int st = l.run.advance();
            if (st < 0) return st;
            changed |= st;
            if (l.run.finished) {
                REFORM_HIP_CHECK(hipEventRecord(l.end, l.stream));
                l.done = true;
                changed = 1;
            } else {
                ++running;
            }
        }
        if (!running) break;
        if (!changed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    for (Lane& l : lanes) REFORM_HIP_CHECK(hipStreamSynchronize(l.stream));
    // per family: the longest group, start to end
    praq_ms = pria_ms = 0.0;
    for (Lane& l : lanes) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, l.begin, l.end) != hipSuccess) continue;
        double& out = l.family == 0 ? praq_ms : pria_ms;
        out = std::max(out, double(ms));
    }
    // where each answer lives: the precision the instance finished in
    if (np) {
        REFORM_HIP_CHECK(hipMemcpy(praq_states.data(), praq_states_dev, sizeof(LogicState) * np,
                                   hipMemcpyDeviceToHost));
        for (int s = 0; s < np; ++s) praq_in64[praq_order[s]] = char(praq_states[s].in64 != 0);
    }
    if (nr) {
        REFORM_HIP_CHECK(hipMemcpy(pria_states.data(), pria_states_dev, sizeof(LogicState) * nr,
                                   hipMemcpyDeviceToHost));
        for (int s = 0; s < nr; ++s) pria_in64[pria_order[s]] = char(pria_states[s].in64 != 0);
    }
    return 0;
}

int BatchImpl::solve_workgroups() {
    const bool f32 = use_float32 != 0;
    const int np = int(praq.size()), nr = int(pria.size());
    if (np) {
        REFORM_HIP_CHECK(hipEventRecord(ev[0], s0));
        int e = f32 ? praq_launch_wg<float>(praq32.dev, np, praq_block, s0)
                    : praq_launch_wg<double>(praq64.dev, np, praq_block, s0);
        if (e) return kErrorDevice;
    }
    if (nr) {
        REFORM_HIP_CHECK(hipEventRecord(ev[2], s1));
        int e = f32 ? pria_launch_wg<float>(pria32.dev, nr, pria_max_T, s1)
                    : pria_launch_wg<double>(pria64.dev, nr, pria_max_T, s1);
        if (e) return kErrorDevice;
    }
    std::fill(praq_in64.begin(), praq_in64.end(), f32 ? 0 : 1);
    std::fill(pria_in64.begin(), pria_in64.end(), f32 ? 0 : 1);
    if (f32) {
        // the stalled instances: FP32 state -> FP64, then one more launch
        if (np && hipMemcpyAsync(praq_states.data(), praq_states_dev, sizeof(LogicState) * np,
                                 hipMemcpyDeviceToHost, s0) != hipSuccess)
            return kErrorDevice;
        if (nr && hipMemcpyAsync(pria_states.data(), pria_states_dev, sizeof(LogicState) * nr,
                                 hipMemcpyDeviceToHost, s1) != hipSuccess)
            return kErrorDevice;
        if (hipStreamSynchronize(s0) != hipSuccess || hipStreamSynchronize(s1) != hipSuccess)
            return kErrorDevice;
        std::vector<PraqJob<float>> p32;
        std::vector<PraqJob<double>> p64;
        for (int s = 0; s < np; ++s) {
            if (!praq_states[s].stalled) continue;
            p32.push_back(praq32.host[s]);
            p64.push_back(praq64.host[s]);
            p64.back().resume = 1;
            praq_in64[praq_order[s]] = 1;
            praq_states[s].switch_iteration = praq_states[s].iteration;
            REFORM_HIP_CHECK(hipMemcpyAsync(praq_states_dev + s, &praq_states[s], sizeof(LogicState),
                               hipMemcpyHostToDevice, s0));
        }
        std::vector<PriaJob<float>> r32;
        std::vector<PriaJob<double>> r64;
        for (int s = 0; s < nr; ++s) {
            if (!pria_states[s].stalled) continue;
            r32.push_back(pria32.host[s]);
            r64.push_back(pria64.host[s]);
            r64.back().resume = 1;
            pria_in64[pria_order[s]] = 1;
            pria_states[s].switch_iteration = pria_states[s].iteration;
            REFORM_HIP_CHECK(hipMemcpyAsync(pria_states_dev + s, &pria_states[s], sizeof(LogicState),
                               hipMemcpyHostToDevice, s1));
        }
        if (!p32.empty()) {
            if (hipMemcpyAsync(praq32.sub, p32.data(), sizeof(p32[0]) * p32.size(),
                               hipMemcpyHostToDevice, s0) != hipSuccess ||
                hipMemcpyAsync(praq64.sub, p64.data(), sizeof(p64[0]) * p64.size(),
                               hipMemcpyHostToDevice, s0) != hipSuccess)
                return kErrorDevice;
            if (praq_convert_jobs(praq32.sub, praq64.sub, int(p32.size()), s0) ||
                praq_launch_wg<double>(praq64.sub, int(p64.size()), praq_block, s0))
                return kErrorDevice;
        }
        if (!r32.empty()) {
            if (hipMemcpyAsync(pria32.sub, r32.data(), sizeof(r32[0]) * r32.size(),
                               hipMemcpyHostToDevice, s1) != hipSuccess ||
                hipMemcpyAsync(pria64.sub, r64.data(), sizeof(r64[0]) * r64.size(),
                               hipMemcpyHostToDevice, s1) != hipSuccess)
                return kErrorDevice;
            if (pria_convert_jobs(pria32.sub, pria64.sub, int(r32.size()), s1) ||
                pria_launch_wg<double>(pria64.sub, int(r64.size()), pria_max_T, s1))
                return kErrorDevice;
        }
    }
    if (np && hipEventRecord(ev[1], s0) != hipSuccess) return kErrorDevice;
    if (nr && hipEventRecord(ev[3], s1) != hipSuccess) return kErrorDevice;
    if (hipStreamSynchronize(s0) != hipSuccess || hipStreamSynchronize(s1) != hipSuccess)
        return kErrorDevice;
    float ms = 0.0f;
    if (np && hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) praq_ms = ms;
    if (nr && hipEventElapsedTime(&ms, ev[2], ev[3]) == hipSuccess) pria_ms = ms;
    return 0;
}

// The plain solver, one instance at a time: its tier, and for FP32 the
// handover to FP64 if it stalls.
int BatchImpl::solve_sequential() {
    const bool f32 = use_float32 != 0;
    using clock = std::chrono::steady_clock;
    praq_seconds.assign(praq.size(), -1.0);
    pria_seconds.assign(pria.size(), -1.0);
    auto since = [](clock::time_point t) {
        return std::chrono::duration<double>(clock::now() - t).count();
    };
    for (size_t q = 0; q < praq.size(); ++q) {
        const int64_t s = praq_slot[q];
        const int tier = praq[q].tier, lanes = praq[q].host.lanes;
        const double limit = praq[q].o.time_limit;
        const clock::time_point t0 = clock::now();
        if (!f32) {
            if (praq_plain<double>(praq64.host[s], praq64.sub, tier, false, limit, lanes))
                return kErrorDevice;
            praq_in64[q] = 1;
            praq_seconds[q] = since(t0);
            continue;
        }
        if (praq_plain<float>(praq32.host[s], praq32.sub, tier, false, limit, lanes))
            return kErrorDevice;
        LogicState st;
        REFORM_HIP_CHECK(hipMemcpy(&st, praq_states_dev + s, sizeof(st), hipMemcpyDeviceToHost));
        praq_in64[q] = 0;
        if (st.stalled) {
            st.switch_iteration = st.iteration;
            REFORM_HIP_CHECK(hipMemcpy(praq_states_dev + s, &st, sizeof(st), hipMemcpyHostToDevice));
            if (praq_convert_state(praq32.host[s].a, praq64.host[s].a, s0)) return kErrorDevice;
            if (praq_plain<double>(praq64.host[s], praq64.sub, tier, true, limit, lanes))
                return kErrorDevice;
            praq_in64[q] = 1;
        }
        praq_seconds[q] = since(t0);
    }
    for (size_t q = 0; q < pria.size(); ++q) {
        const int64_t s = pria_slot[q];
        const int tier = pria[q].tier, T = pria[q].host.T;
        const double limit = pria[q].o.time_limit;
        const clock::time_point t0 = clock::now();
        if (!f32) {
            if (pria_plain<double>(pria64.host[s], pria64.sub, tier, false, T, limit)) return kErrorDevice;
            pria_in64[q] = 1;
            pria_seconds[q] = since(t0);
            continue;
        }
        if (pria_plain<float>(pria32.host[s], pria32.sub, tier, false, T, limit)) return kErrorDevice;
        LogicState st;
        REFORM_HIP_CHECK(hipMemcpy(&st, pria_states_dev + s, sizeof(st), hipMemcpyDeviceToHost));
        pria_in64[q] = 0;
        if (st.stalled) {
            st.switch_iteration = st.iteration;
            REFORM_HIP_CHECK(hipMemcpy(pria_states_dev + s, &st, sizeof(st), hipMemcpyHostToDevice));
            if (pria_convert_state(pria32.host[s].a, pria64.host[s].a, s1)) return kErrorDevice;
            if (pria_plain<double>(pria64.host[s], pria64.sub, tier, true, T, limit)) return kErrorDevice;
            pria_in64[q] = 1;
        }
        pria_seconds[q] = since(t0);
    }
    return 0;
}

int BatchImpl::praq_result(int64_t q, double* x, double* y, PraqInfo* info) {
    if (q < 0 || q >= int64_t(praq.size()) || !uploaded) return kErrorArgument;
    const PraqEntry& e = praq[q];
    const int64_t s = praq_slot[q];
    const LogicState& st = praq_states[s];
    const bool in64 = praq_in64[q] != 0;
    std::vector<double> hx, hy;
    if (x) {
        int r = in64 ? read_array(hx, praq64.host[s].a.x, e.host.n)
                     : read_array(hx, praq32.host[s].a.x, e.host.n);
        if (r) return r;
        std::memcpy(x, hx.data(), sizeof(double) * e.host.n);
    }
    if (y && e.host.m > 0) {
        const size_t m2 = 2 * size_t(e.host.m);
        int r = in64 ? read_array(hy, praq64.host[s].a.y, m2) : read_array(hy, praq32.host[s].a.y, m2);
        if (r) return r;
        for (int k = 0; k < e.host.m; ++k) {
            y[2 * size_t(k)] = hy[k] * e.host.scale;
            y[2 * size_t(k) + 1] = hy[e.host.m + k] * e.host.scale;
        }
    }
    if (info) {
        info->lower = st.lower;
        info->upper = st.upper;
        info->relative_gap = st.gap;
        info->iterations = st.iteration;
        info->restarts = st.restarts;
        info->checks = st.checks;
        info->omega = st.omega;
        info->scale = e.host.scale;
        info->status = st.status;
        info->tier = e.tier;
        info->switch_iteration = st.switch_iteration;
        info->device_bytes = int64_t(e.bytes);
        info->bracket = info->upper - info->lower;
    }
    return 0;
}

int BatchImpl::pria_result(int64_t q, double* x, double* y, double* z, pria_info* info) {
    if (q < 0 || q >= int64_t(pria.size()) || !uploaded) return kErrorArgument;
    const PriaEntry& e = pria[q];
    const int64_t s = pria_slot[q];
    const LogicState& st = pria_states[s];
    const bool in64 = pria_in64[q] != 0;
    std::vector<double> xd(size_t(e.n) * e.t * e.t), v;
    double* outs[3] = {xd.data(), y, z};
    for (int w = 0; w < 3; ++w) {
        if (!outs[w]) continue;
        int r;
        if (in64) {
            const double* src[3] = {pria64.host[s].a.x, pria64.host[s].a.y, pria64.host[s].a.z};
            r = read_array(v, src[w], e.host.ND);
        } else {
            const float* src[3] = {pria32.host[s].a.x, pria32.host[s].a.y, pria32.host[s].a.z};
            r = read_array(v, src[w], e.host.ND);
        }
        if (r) return r;
        pria_unpack(e.host, v, w, outs[w]);
    }
    if (x) std::memcpy(x, xd.data(), sizeof(double) * xd.size());
    if (info) {
        info->iterations = st.iteration;
        info->restarts = st.restarts;
        info->checks = st.checks;
        info->lower = st.lower;
        info->upper_iterate = st.upper;
        info->upper = pria_objective_host(e.n, e.t, e.k.data(), e.k0, e.h.data(), xd.data());
        info->relative_gap = st.gap;
        info->omega = st.omega;
        info->scale = e.host.scale;
        info->status = st.status;
        info->tier = e.tier;
        info->switch_iteration = st.switch_iteration;
        info->device_bytes = int64_t(e.bytes);
        info->bracket = info->upper - info->lower;
    }
    return 0;
}

}  // namespace

struct reform_batch {
    std::unique_ptr<BatchImpl> impl;
};

extern "C" reform_batch* reform_batch_create(int use_float32) {
    auto* b = new reform_batch;
    b->impl.reset(new BatchImpl());
    b->impl->use_float32 = use_float32 ? 1 : 0;
    return b;
}

extern "C" void reform_batch_destroy(reform_batch* batch) { delete batch; }

extern "C" int64_t reform_batch_add_praq(reform_batch* batch, const PraqInstance* instance,
                                         const PraqParameters* parameters,
                                         const PraqOptions* options) {
    if (!batch || !instance || !parameters || batch->impl->uploaded) return kErrorArgument;
    if (instance->n <= 0 || instance->m < 0 || !instance->diag ||
        (instance->m > 0 && (!instance->rows || !instance->cols || !instance->vals)))
        return kErrorInstance;
    PraqEntry e;
    e.diag.assign(instance->diag, instance->diag + instance->n);
    if (instance->m > 0) {
        e.rows.assign(instance->rows, instance->rows + instance->m);
        e.cols.assign(instance->cols, instance->cols + instance->m);
        e.vals.assign(instance->vals, instance->vals + instance->m);
    }
    PraqParameters p = *parameters;
    if (p.max_iter <= 0) p.max_iter = -1;
    if (p.op_steps <= 0) p.op_steps = praq_check_interval(instance->n, &p);
    if (p.op_steps < 2) p.op_steps = 2;
    e.o = options ? *options : praq_default_options();
    e.lp = praq_logic_params(p, e.o);
    auto& v = batch->impl->praq;
    v.push_back(std::move(e));
    PraqEntry& b = v.back();
    b.copy.n = instance->n;
    b.copy.m = instance->m;
    b.copy.diag = b.diag.data();
    b.copy.rows = b.rows.data();
    b.copy.cols = b.cols.data();
    b.copy.vals = b.vals.data();
    return int64_t(v.size()) - 1;
}

extern "C" int64_t reform_batch_add_pria(reform_batch* batch, int64_t n, int64_t t,
                                         const double* order_costs, double joint_cost,
                                         const double* holding_costs, const pria_params* params,
                                         const pria_options* options) {
    if (!batch || !params || batch->impl->uploaded || n <= 0 || t <= 0 || !order_costs ||
        !holding_costs)
        return kErrorArgument;
    PriaEntry e;
    e.n = n;
    e.t = t;
    e.k.assign(order_costs, order_costs + n);
    e.h.assign(holding_costs, holding_costs + size_t(n) * t * t);
    e.k0 = joint_cost;
    e.lp = pria_logic_params(*params);
    e.o = options ? *options : pria_default_options();
    batch->impl->pria.push_back(std::move(e));
    return int64_t(batch->impl->pria.size()) - 1;
}

extern "C" const char* reform_last_error(void) {
    static thread_local std::string copy;
    std::lock_guard<std::mutex> lock(last_error_mutex());
    copy = last_error_text();
    return copy.c_str();
}

extern "C" int reform_batch_upload(reform_batch* batch) {
    return batch ? batch->impl->upload() : kErrorArgument;
}

extern "C" int reform_batch_solve(reform_batch* batch, int mode) {
    return batch ? batch->impl->solve(mode) : kErrorArgument;
}

extern "C" int reform_batch_set_shape(reform_batch* batch, int praq_block, int pria_block) {
    if (!batch) return kErrorArgument;
    auto ok = [](int b) { return b == 0 || b == 64 || b == 128 || b == 256 || b == 512 || b == 1024; };
    if (!ok(praq_block) || !(pria_block == 0 || pria_block == 1024)) return kErrorArgument;
    if (praq_block) batch->impl->praq_block = praq_block;
    return 0;
}

extern "C" int64_t reform_batch_count(const reform_batch* batch, int family) {
    if (!batch) return kErrorArgument;
    return family == 0 ? int64_t(batch->impl->praq.size()) : int64_t(batch->impl->pria.size());
}

extern "C" int reform_batch_kernel_ms(const reform_batch* batch, double* praq_ms, double* pria_ms) {
    if (!batch) return kErrorArgument;
    if (praq_ms) *praq_ms = batch->impl->praq_ms;
    if (pria_ms) *pria_ms = batch->impl->pria_ms;
    return 0;
}

extern "C" int reform_batch_set_compaction(reform_batch* batch, int on) {
    if (!batch) return kErrorArgument;
    batch->impl->compaction = on != 0;
    return 0;
}

extern "C" int reform_batch_set_timing(reform_batch* batch, int on) {
    if (!batch) return kErrorArgument;
    batch->impl->set_timing(on != 0);
    return 0;
}

extern "C" int reform_batch_kernel_stats(const reform_batch* batch, int family,
                                         reform_kernel_stats* out) {
    if (!batch || !out || family < 0 || family > 1) return kErrorArgument;
    const BatchImpl& b = *batch->impl;
    static const char* kNames[2][2] = {{"wpraq_node", "wpraq_edge"}, {"wpria_x", "wpria_yz"}};
    const std::vector<LogicState>& states = family == 0 ? b.praq_states : b.pria_states;
    for (int k = 0; k < 2; ++k) {
        std::snprintf(out[k].name, sizeof(out[k].name), "%s", kNames[family][k]);
        out[k].launches = 0;
        out[k].ms = 0.0;
        out[k].bytes = 0.0;
    }
    const auto& engines = b.wide[family];
    for (size_t g = 0; g < engines.size(); ++g) {
        // the group's states, in its table order
        const size_t lo = std::min<size_t>(size_t(b.group_first[family][g]), states.size());
        const size_t hi = std::min<size_t>(lo + size_t(engines[g]->count), states.size());
        const std::vector<LogicState> sub(states.begin() + lo, states.begin() + hi);
        double bytes[WideControl::kKinds] = {0.0, 0.0};
        engines[g]->algorithmic_bytes(sub, bytes);
        for (int k = 0; k < 2; ++k) {
            out[k].launches += engines[g]->launches[k];
            out[k].ms += engines[g]->kernel_ms[k];
            out[k].bytes += bytes[k];
        }
    }
    return 0;
}

extern "C" double reform_batch_instance_seconds(const reform_batch* batch, int family,
                                               int64_t index) {
    if (!batch || family < 0 || family > 1) return -1.0;
    const std::vector<double>& v =
        family == 0 ? batch->impl->praq_seconds : batch->impl->pria_seconds;
    return index >= 0 && index < int64_t(v.size()) ? v[index] : -1.0;
}

extern "C" int64_t reform_batch_device_bytes(const reform_batch* batch) {
    if (!batch) return kErrorArgument;
    const BatchImpl& b = *batch->impl;
    size_t bytes = b.dev_bytes;
    for (const auto& family : b.wide)
        for (const auto& w : family) bytes += w->bytes;
    return int64_t(bytes);
}

extern "C" int reform_batch_praq_result(reform_batch* batch, int64_t index, double* x, double* y,
                                        PraqInfo* info) {
    return batch ? batch->impl->praq_result(index, x, y, info) : kErrorArgument;
}

extern "C" int reform_batch_pria_result(reform_batch* batch, int64_t index, double* x, double* y,
                                        double* z, pria_info* info) {
    return batch ? batch->impl->pria_result(index, x, y, z, info) : kErrorArgument;
}
