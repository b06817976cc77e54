// Ragged batches of PRAQ and PRIA instances: include/reform_batch.h.
//
// Device memory: one allocation holds every instance's image (inputs first,
// iteration state after, 256-byte aligned), the job records the kernels read
// and the final scalar states.  The batched mode launches one workgroup per
// instance -- PRAQ and PRIA on two streams so the families overlap -- with
// the instances ordered largest first, so the longest solves start first and
// the small ones fill the CUs around them.  The sequential mode is the plain
// solver: per instance the tier praq_solve_ex / pria_solve_ex would pick.
#include "reform_batch.h"
#include "praq_host.hpp"
#include "pria_host.hpp"

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <cstring>
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
    PraqInstance copy;  // pointers into owned storage below
    std::vector<double> diag, vals;
    std::vector<int64_t> rows, cols;
    LogicParams lp;
    PraqOptions o;
    PraqHostInstance host;
    size_t off = 0;      // image offset in the device allocation
    int tier = 0;
};

struct PriaEntry {
    int64_t n = 0, t = 0;
    std::vector<double> k, h;
    double k0 = 0.0;
    LogicParams lp;
    pria_options o;
    PriaHostInstance host;
    size_t off = 0;
    int tier = 0;
};

struct BatchBase {
    int use_float32 = 0;
    int praq_block = 256, pria_block = 1024;
    std::vector<PraqEntry> praq;
    std::vector<PriaEntry> pria;
    void* dev = nullptr;
    size_t dev_bytes = 0;
    hipStream_t s0 = nullptr, s1 = nullptr;
    std::vector<LogicState> praq_states, pria_states;
    bool uploaded = false;
    // events around the batched launches: [praq start, praq end, pria start, pria end]
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double praq_ms = 0.0, pria_ms = 0.0;
    virtual ~BatchBase() {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (dev) (void)hipFree(dev);
        if (s0) (void)hipStreamDestroy(s0);
        if (s1) (void)hipStreamDestroy(s1);
    }
    virtual int upload() = 0;
    virtual int solve(int mode) = 0;
    virtual int praq_result(int64_t q, double* x, double* y, PraqInfo* info) = 0;
    virtual int pria_result(int64_t q, double* x, double* y, double* z, pria_info* info) = 0;
};

template <class Real>
struct Batch : BatchBase {
    std::vector<PraqJob<Real>> praq_jobs;   // in launch order (largest first)
    std::vector<PriaJob<Real>> pria_jobs;
    std::vector<int64_t> praq_order, pria_order;  // launch slot -> index
    std::vector<int64_t> praq_slot, pria_slot;    // index -> launch slot
    size_t praq_jobs_off = 0, pria_jobs_off = 0, praq_states_off = 0, pria_states_off = 0;
    size_t work_off = 0;
    int pria_max_T = 2;

    int upload() override {
        if (uploaded) return kErrorArgument;
        // host builds, in parallel
        std::vector<int> fail(praq.size() + pria.size(), 0);
        parallel_for(praq.size(), [&](size_t q) {
            PraqEntry& e = praq[q];
            fail[q] = praq_build_host(e.copy, e.o.dense_degree, nullptr, nullptr, e.host);
        });
        parallel_for(pria.size(), [&](size_t q) {
            PriaEntry& e = pria[q];
            fail[praq.size() + q] = pria_build_host(e.n, e.t, e.k.data(), e.k0, e.h.data(),
                                                    nullptr, nullptr, nullptr, e.host);
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
        praq_slot.resize(praq.size());
        pria_slot.resize(pria.size());
        for (size_t s = 0; s < praq_order.size(); ++s) praq_slot[praq_order[s]] = int64_t(s);
        for (size_t s = 0; s < pria_order.size(); ++s) pria_slot[pria_order[s]] = int64_t(s);

        // device allocation: images, jobs, states, one device-tier work area
        size_t at = 0, work = praq_device_work_bytes();
        std::vector<PraqLayout> pl(praq.size());
        std::vector<PriaLayout> rl(pria.size());
        for (size_t q = 0; q < praq.size(); ++q) {
            pl[q] = praq_layout(praq[q].host.n, praq[q].host.m, sizeof(Real));
            praq[q].off = at;
            at += align256(pl[q].total_bytes);
        }
        for (size_t q = 0; q < pria.size(); ++q) {
            rl[q] = pria_layout(pria[q].host.N, pria[q].host.ND, sizeof(Real));
            pria[q].off = at;
            at += align256(rl[q].total_bytes);
            work = std::max(work, pria_device_work_bytes(pria[q].host.N, pria[q].host.T));
            pria_max_T = std::max(pria_max_T, pria[q].host.T);
        }
        praq_jobs_off = at;
        at = align256(at + sizeof(PraqJob<Real>) * (praq.size() + 1));
        pria_jobs_off = at;
        at = align256(at + sizeof(PriaJob<Real>) * (pria.size() + 1));
        praq_states_off = at;
        at = align256(at + sizeof(LogicState) * (praq.size() + 1));
        pria_states_off = at;
        at = align256(at + sizeof(LogicState) * (pria.size() + 1));
        work_off = at;
        at = align256(at + work);
        dev_bytes = at;
        if (hipMalloc(&dev, dev_bytes) != hipSuccess) {
            dev = nullptr;
            return kErrorDevice;
        }
        if (hipStreamCreateWithFlags(&s0, hipStreamNonBlocking) != hipSuccess ||
            hipStreamCreateWithFlags(&s1, hipStreamNonBlocking) != hipSuccess)
            return kErrorDevice;
        for (auto& e : ev)
            if (hipEventCreate(&e) != hipSuccess) return kErrorDevice;
        char* base = static_cast<char*>(dev);

        // inputs: build each image into one host staging area, one copy each
        size_t up_total = 0;
        for (size_t q = 0; q < praq.size(); ++q) up_total = std::max(up_total, pl[q].input_bytes);
        for (size_t q = 0; q < pria.size(); ++q)
            up_total = std::max(up_total, pria_upload_bytes(pria[q].host, rl[q]));
        std::vector<char> image(up_total);
        const long long ticks = (long long)(kWgLimitSeconds * clock_rate_hz());
        praq_jobs.assign(praq.size(), PraqJob<Real>());
        for (size_t s = 0; s < praq_order.size(); ++s) {
            PraqEntry& e = praq[praq_order[s]];
            std::fill(image.begin(), image.begin() + pl[praq_order[s]].input_bytes, 0);
            praq_fill_image<Real>(e.host, pl[praq_order[s]], image.data());
            if (hipMemcpy(base + e.off, image.data(), pl[praq_order[s]].input_bytes,
                          hipMemcpyHostToDevice) != hipSuccess)
                return kErrorDevice;
            PraqJob<Real>& j = praq_jobs[s];
            std::memset(&j, 0, sizeof(j));
            j.a = praq_bind<Real>(pl[praq_order[s]], base + e.off, e.host.n, e.host.m);
            j.p = e.lp;
            j.scale = e.host.scale;
            j.lanes = e.host.lanes;
            j.time_limit_ticks = ticks;
            j.state = reinterpret_cast<LogicState*>(base + praq_states_off) + s;
            int64_t size = int64_t(e.host.n) + int64_t(e.host.m);
            e.tier = e.o.tier != PRAQ_TIER_AUTO
                         ? e.o.tier
                         : (size <= env_int("REFORM_PRAQ_WG_MAX", int64_t(1) << 17)
                                ? PRAQ_TIER_WORKGROUP
                                : PRAQ_TIER_DEVICE);
        }
        pria_jobs.assign(pria.size(), PriaJob<Real>());
        for (size_t s = 0; s < pria_order.size(); ++s) {
            PriaEntry& e = pria[pria_order[s]];
            const PriaLayout& L = rl[pria_order[s]];
            size_t up = pria_upload_bytes(e.host, L);
            std::fill(image.begin(), image.begin() + up, 0);
            pria_fill_image<Real>(e.host, L, image.data());
            if (hipMemcpy(base + e.off, image.data(), up, hipMemcpyHostToDevice) != hipSuccess)
                return kErrorDevice;
            PriaJob<Real>& j = pria_jobs[s];
            std::memset(&j, 0, sizeof(j));
            j.a = pria_bind<Real>(L, base + e.off, e.host);
            if (e.host.cold && pria_cold_start<Real>(j.a, s1)) return kErrorDevice;
            j.p = e.lp;
            j.scale = e.host.scale;
            j.time_limit_ticks = ticks;
            j.state = reinterpret_cast<LogicState*>(base + pria_states_off) + s;
            e.tier = e.o.tier != PRIA_TIER_AUTO
                         ? e.o.tier
                         : (e.host.ND <= env_int("REFORM_PRIA_WG_MAX", int64_t(1) << 15)
                                ? PRIA_TIER_WORKGROUP
                                : PRIA_TIER_DEVICE);
        }
        if (!praq_jobs.empty() &&
            hipMemcpy(base + praq_jobs_off, praq_jobs.data(), sizeof(PraqJob<Real>) * praq.size(),
                      hipMemcpyHostToDevice) != hipSuccess)
            return kErrorDevice;
        if (!pria_jobs.empty() &&
            hipMemcpy(base + pria_jobs_off, pria_jobs.data(), sizeof(PriaJob<Real>) * pria.size(),
                      hipMemcpyHostToDevice) != hipSuccess)
            return kErrorDevice;
        if (hipStreamSynchronize(s1) != hipSuccess) return kErrorDevice;
        praq_states.assign(praq.size(), LogicState());
        pria_states.assign(pria.size(), LogicState());
        uploaded = true;
        return 0;
    }

    int solve(int mode) override {
        if (!uploaded) return kErrorArgument;
        char* base = static_cast<char*>(dev);
        auto* pj = reinterpret_cast<PraqJob<Real>*>(base + praq_jobs_off);
        auto* rj = reinterpret_cast<PriaJob<Real>*>(base + pria_jobs_off);
        double* work = reinterpret_cast<double*>(base + work_off);
        praq_ms = pria_ms = 0.0;
        if (mode == REFORM_BATCHED) {
            if (!praq.empty()) {
                if (hipEventRecord(ev[0], s0) != hipSuccess) return kErrorDevice;
                if (praq_launch_wg<Real>(pj, int(praq.size()), praq_block, s0)) return kErrorDevice;
                if (hipEventRecord(ev[1], s0) != hipSuccess) return kErrorDevice;
            }
            if (!pria.empty()) {
                if (hipEventRecord(ev[2], s1) != hipSuccess) return kErrorDevice;
                if (pria_launch_wg<Real>(rj, int(pria.size()), pria_max_T, s1)) return kErrorDevice;
                if (hipEventRecord(ev[3], s1) != hipSuccess) return kErrorDevice;
            }
        } else if (mode == REFORM_SEQUENTIAL) {
            // in the order the instances were added, one at a time
            for (size_t q = 0; q < praq.size(); ++q) {
                int64_t s = praq_slot[q];
                if (praq[q].tier == PRAQ_TIER_WORKGROUP) {
                    if (praq_launch_wg<Real>(pj + s, 1, 1024, s0)) return kErrorDevice;
                    if (hipStreamSynchronize(s0) != hipSuccess) return kErrorDevice;
                } else {
                    LogicState st;
                    if (praq_device_solve<Real>(praq_jobs[s], st, work, 0.0, s0)) return kErrorDevice;
                    if (hipMemcpyAsync(praq_jobs[s].state, &st, sizeof(st), hipMemcpyHostToDevice,
                                       s0) != hipSuccess ||
                        hipStreamSynchronize(s0) != hipSuccess)
                        return kErrorDevice;
                }
            }
            for (size_t q = 0; q < pria.size(); ++q) {
                int64_t s = pria_slot[q];
                if (pria[q].tier == PRIA_TIER_WORKGROUP) {
                    if (pria_launch_wg<Real>(rj + s, 1, pria[q].host.T, s1)) return kErrorDevice;
                    if (hipStreamSynchronize(s1) != hipSuccess) return kErrorDevice;
                } else {
                    LogicState st;
                    if (pria_device_solve<Real>(pria_jobs[s], st, work, 0.0, s1)) return kErrorDevice;
                    if (hipMemcpyAsync(pria_jobs[s].state, &st, sizeof(st), hipMemcpyHostToDevice,
                                       s1) != hipSuccess ||
                        hipStreamSynchronize(s1) != hipSuccess)
                        return kErrorDevice;
                }
            }
        } else {
            return kErrorArgument;
        }
        if (!praq.empty() &&
            hipMemcpyAsync(praq_states.data(), base + praq_states_off,
                           sizeof(LogicState) * praq.size(), hipMemcpyDeviceToHost, s0) != hipSuccess)
            return kErrorDevice;
        if (!pria.empty() &&
            hipMemcpyAsync(pria_states.data(), base + pria_states_off,
                           sizeof(LogicState) * pria.size(), hipMemcpyDeviceToHost, s1) != hipSuccess)
            return kErrorDevice;
        if (hipStreamSynchronize(s0) != hipSuccess || hipStreamSynchronize(s1) != hipSuccess)
            return kErrorDevice;
        if (mode == REFORM_BATCHED) {
            float ms = 0.0f;
            if (!praq.empty() && hipEventElapsedTime(&ms, ev[0], ev[1]) == hipSuccess) praq_ms = ms;
            if (!pria.empty() && hipEventElapsedTime(&ms, ev[2], ev[3]) == hipSuccess) pria_ms = ms;
        }
        return 0;
    }

    int praq_result(int64_t q, double* x, double* y, PraqInfo* info) override {
        if (q < 0 || q >= int64_t(praq.size()) || !uploaded) return kErrorArgument;
        const PraqEntry& e = praq[q];
        const int64_t s = praq_slot[q];
        const PraqJob<Real>& j = praq_jobs[s];
        const LogicState& st = praq_states[s];
        if (x) {
            std::vector<Real> hx(e.host.n);
            if (hipMemcpy(hx.data(), j.a.x, sizeof(Real) * e.host.n, hipMemcpyDeviceToHost) !=
                hipSuccess)
                return kErrorDevice;
            for (int i = 0; i < e.host.n; ++i) x[i] = double(hx[i]);
        }
        if (y && e.host.m > 0) {
            std::vector<Real> hy(2 * size_t(e.host.m));
            if (hipMemcpy(hy.data(), j.a.y, sizeof(Real) * hy.size(), hipMemcpyDeviceToHost) !=
                hipSuccess)
                return kErrorDevice;
            for (int k = 0; k < e.host.m; ++k) {
                y[2 * size_t(k)] = double(hy[k]) * e.host.scale;
                y[2 * size_t(k) + 1] = double(hy[e.host.m + k]) * e.host.scale;
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
        }
        return 0;
    }

    int pria_result(int64_t q, double* x, double* y, double* z, pria_info* info) override {
        if (q < 0 || q >= int64_t(pria.size()) || !uploaded) return kErrorArgument;
        const PriaEntry& e = pria[q];
        const int64_t s = pria_slot[q];
        const PriaJob<Real>& j = pria_jobs[s];
        const LogicState& st = pria_states[s];
        std::vector<Real> v(e.host.ND);
        std::vector<double> xd(size_t(e.n) * e.t * e.t);
        double* outs[3] = {xd.data(), y, z};
        Real* srcs[3] = {j.a.x, j.a.y, j.a.z};
        for (int w = 0; w < 3; ++w) {
            if (!outs[w]) continue;
            if (hipMemcpy(v.data(), srcs[w], sizeof(Real) * v.size(), hipMemcpyDeviceToHost) !=
                hipSuccess)
                return kErrorDevice;
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
        }
        return 0;
    }
};

}  // namespace

struct reform_batch {
    std::unique_ptr<BatchBase> impl;
};

extern "C" reform_batch* reform_batch_create(int use_float32) {
    auto* b = new reform_batch;
    if (use_float32)
        b->impl.reset(new Batch<float>());
    else
        b->impl.reset(new Batch<double>());
    b->impl->use_float32 = use_float32;
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
    if (pria_block) batch->impl->pria_block = pria_block;
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

extern "C" int64_t reform_batch_device_bytes(const reform_batch* batch) {
    return batch ? int64_t(batch->impl->dev_bytes) : kErrorArgument;
}

extern "C" int reform_batch_praq_result(reform_batch* batch, int64_t index, double* x, double* y,
                                        PraqInfo* info) {
    return batch ? batch->impl->praq_result(index, x, y, info) : kErrorArgument;
}

extern "C" int reform_batch_pria_result(reform_batch* batch, int64_t index, double* x, double* y,
                                        double* z, pria_info* info) {
    return batch ? batch->impl->pria_result(index, x, y, z, info) : kErrorArgument;
}
