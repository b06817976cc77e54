// The serving loop of include/reform_serve.h.
//
// Threads: the caller's threads submit (copy inputs, enqueue); `host_threads`
// builders turn inputs into device images in host memory (half-edge build or
// slice layout, both precisions); one serving thread owns the GPU.  Its loop:
//
//   admit     move built requests into the device pool while they fit (first
//             fit, 256-byte aligned, coalescing on free): upload the images,
//             cold-start PRIA images, write the job records and a fresh
//             LogicState (kActNew: the request starts at its next round)
//   rebuild   when a family's membership changed, its wide engine's block
//             tables and job arrays are rebuilt for the new members
//   round     one round of each family's engine (PRAQ and PRIA on their own
//             streams), a hipGraph captured once per membership and replayed
//   retire    a per-member done flag comes back; finished members' answers
//             and states are copied out, their memory and slots freed
//
// so requests join and leave at round boundaries while the others keep
// their Halpern state on the device.
#include "reform_serve.h"
#include "tune.hpp"
#include "wide.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <memory>
#include <set>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

using namespace reform;

namespace {

constexpr int kErrorTicket = -4;
constexpr int kErrorTimeout = -5;
constexpr int kPending = -1;
// cancelled after admission (the answer is the point it stopped at) and
// before it (no answer)
constexpr int kCancelled = 3;
constexpr int kCancelledQueued = -6;
// host threads copying finished requests' answers out; at 1000 requests/s
// three threads served no faster than one (the staged copies share the link)
constexpr int kRetireThreads = 1;

// flags[k] for member k of a family's tables: 1 finished, 2 waiting for its
// FP64 image (an FP32 request at its handover), else 0
template <class Job>
__global__ void serve_done_flags(const Job* jobs, int count, int* flags) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const LogicState* st = jobs[k].state;
    flags[k] = st->action == kActDoneX ? 1 : (!st->in64 && st->stalled && st->hold64 ? 2 : 0);
}

// First-fit allocator over one device allocation.
class Pool {
public:
    int init(size_t bytes) {
        capacity_ = bytes & ~size_t(255);
        hipError_t e = device_malloc(&base_, capacity_);
        if (e != hipSuccess) {
            base_ = nullptr;
            record_error("hipMalloc(serving pool)", e, __FILE__, __LINE__);
            return kErrorDevice;
        }
        free_[0] = capacity_;
        return 0;
    }
    ~Pool() {
        if (base_) (void)hipFree(base_);
    }
    // offset of a block of `bytes`, or SIZE_MAX: the first free block that
    // fits and leaves a free block of at least `headroom` bytes
    size_t alloc(size_t bytes, size_t headroom = 0) {
        bytes = align256(bytes ? bytes : 1);
        size_t big1 = 0, big2 = 0;   // the two largest free blocks
        for (const auto& f : free_) {
            if (f.second > big1) {
                big2 = big1;
                big1 = f.second;
            } else if (f.second > big2) {
                big2 = f.second;
            }
        }
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            const size_t off = it->first, left = it->second - bytes;
            if (headroom && std::max(left, it->second == big1 ? big2 : big1) < headroom) continue;
            free_.erase(it);
            if (left) free_[off + bytes] = left;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return off;
        }
        return SIZE_MAX;
    }
    void release(size_t off, size_t bytes) {
        bytes = align256(bytes ? bytes : 1);
        used_ -= bytes;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }
    char* base() const { return static_cast<char*>(base_); }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }
    size_t capacity() const { return capacity_; }

private:
    void* base_ = nullptr;
    size_t capacity_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;  // offset -> bytes
};

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

struct Request {
    int64_t ticket = 0;
    int family = 0;
    // inputs: the caller's arrays (borrowed) or the copies below
    const double *diag = nullptr, *vals = nullptr, *k = nullptr, *h = nullptr;
    const int64_t *rows = nullptr, *cols = nullptr;
    std::vector<double> own_diag, own_vals, own_k, own_h;
    std::vector<int64_t> own_rows, own_cols;
    int64_t n = 0, m = 0, t = 0;
    double k0 = 0.0;
    PraqOptions po{};
    LogicParams lp{};
    // host images
    PraqHostInstance ph;
    PriaHostInstance rh;
    PraqLayout pl64{}, pl32{};
    PriaLayout rl64{}, rl32{};
    std::vector<char> img64, img32;
    size_t up64 = 0, up32 = 0;
    int build_status = 0;
    // device: the first image (FP32 on a server with FP32 iterates, else
    // FP64) from admission; on an FP32 server the FP64 image from the
    // request's handover (bytes64 its size, off64 SIZE_MAX until then)
    size_t off = SIZE_MAX, bytes = 0;
    size_t off64 = SIZE_MAX, bytes64 = 0;
    int slot = -1;
    PraqJob<double> pj64{};
    PraqJob<float> pj32{};
    PriaJob<double> rj64{};
    PriaJob<float> rj32{};
    LogicState init{};  // the first state, copied from here at admission
    // outcome
    LogicState final{};
    bool in64 = true;
    std::vector<double> d1, d2, d3;
    std::vector<float> f1, f2, f3;
    // PRIA: the upper bound of the unpadded original at the answer, computed
    // in the retire thread while borrowed inputs are still the caller's to
    // keep (until the request has finished)
    double upper_original = 0.0;
    int status = kPending;
    double t_submit = 0.0, t_admit = -1.0, t_done = -1.0;
    double time_limit = 0.0;  // seconds from admission; <= 0: none
    // set (under the server's lock) by reform_server_cancel; read by the
    // serving thread when it rebuilds the tables
    std::atomic<bool> cancel{false};
};

// Every input value finite (the server turns away what the plain solvers
// would iterate on to their caps).
bool all_finite(const double* v, size_t count) {
    for (size_t i = 0; i < count; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

// The holding costs a PRIA instance reads, H[i][s][t] for s <= t, finite.
bool holding_finite(const double* h, int64_t n, int64_t t) {
    for (int64_t i = 0; i < n; ++i)
        for (int64_t s = 0; s < t; ++s)
            if (!all_finite(h + (size_t(i) * t + s) * t + s, size_t(t - s))) return false;
    return true;
}

}  // namespace

struct reform_server {
    reform_server_config cfg{};
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::mutex mu;
    std::condition_variable cv_build, cv_ready, cv_done, cv_retire;
    std::deque<Request*> to_build, ready;
    std::deque<Request*> retiring;   // finished, left the tables, answers not yet copied
    std::deque<int64_t> finished;
    std::unordered_map<int64_t, std::unique_ptr<Request>> requests;
    std::atomic<bool> stop{false};
    int64_t next_ticket = 0;
    int64_t n_submitted = 0, n_admitted = 0, n_completed = 0, n_rejected = 0;
    int64_t n_cancelled = 0;
    // tickets of cancelled requests already admitted, for the serving thread
    // (under mu)
    std::vector<int64_t> cancels;
    int n_building = 0;  // requests on the builders
    std::atomic<int64_t> n_rounds{0}, n_rebuilds{0}, n_in_flight{0}, n_in_flight_peak{0};
    std::atomic<int64_t> n_handovers{0}, n_handover_waits{0};
    // FP64 image sizes of the FP32 requests in flight that have not handed
    // over (guarded by mu): admission keeps room for the largest
    std::multiset<size_t> pending64;
    std::atomic<double> retire_s{0.0}, rebuild_s{0.0};   // serving-thread host time
    std::vector<std::thread> builders;
    std::thread server;
    std::vector<std::thread> retirers;              // kRetireThreads, a stream each
    int device = 0;
    int start_status = 0;

    // serving-thread state
    Pool pool;
    LogicState* states = nullptr;
    std::vector<int> free_slots;
    hipStream_t streams[2] = {nullptr, nullptr};  // the families' rounds
    hipStream_t upload = nullptr;                  // admissions' copies
    std::vector<Request*> members[2];              // in table order
    std::vector<Request*> joining[2];              // admitted, join at the next round
    bool changed[2] = {false, false};
    // per family: a round in flight, its completion event and its done flags
    bool in_flight[2] = {false, false};
    hipEvent_t round_done[2] = {nullptr, nullptr};
    int* flags_dev[2] = {nullptr, nullptr};
    int* flags_host[2] = {nullptr, nullptr};       // pinned
    int flags_cap[2] = {0, 0};
    int stable[2] = {0, 0};                        // rounds since the tables changed
    WidePraq praq_engine;
    WidePria pria_engine;
    void* jobs_dev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [family][precision]
    size_t jobs_cap[2] = {0, 0};
    std::atomic<size_t> bytes_in_use{0}, bytes_peak{0};

    // Stop every thread: `stop` is set under the lock, so a thread that has
    // checked its wait predicate under the lock either sees it or is already
    // waiting when the notifications come (no lost wakeup).
    void halt() {
        {
            std::lock_guard<std::mutex> lock(mu);
            stop = true;
        }
        notify_every_thread();
    }
    void notify_every_thread() {
        cv_build.notify_all();
        cv_ready.notify_all();
        cv_retire.notify_all();
        cv_done.notify_all();
    }

    ~reform_server() {
        halt();
        for (auto& b : builders)
            if (b.joinable()) b.join();
        if (server.joinable()) server.join();
        for (auto& r : retirers)
            if (r.joinable()) r.join();
        for (auto& f : jobs_dev)
            for (void* p : f)
                if (p) (void)hipFree(p);
        for (int f = 0; f < 2; ++f) {
            if (flags_dev[f]) (void)hipFree(flags_dev[f]);
            if (flags_host[f]) (void)hipHostFree(flags_host[f]);
            if (round_done[f]) (void)hipEventDestroy(round_done[f]);
        }
        if (states) (void)hipFree(states);
        for (auto& s : streams)
            if (s) (void)hipStreamDestroy(s);
        if (upload) (void)hipStreamDestroy(upload);
    }

    void build_loop();
    void serve_loop();
    void retire_loop();
    int build(Request& r);
    int admit(Request& r);
    int grant64(Request& r, hipStream_t s);
    int rebuild(int family);
    int launch(int family);
    int retire_finished(int family);
    void retire(Request& r, int status);
    void finish_unadmitted(Request& r, int status);
    void stop_cancelled();
};

// ------------------------------------------------------------ host builds --

int reform_server::build(Request& r) {
    const bool f32 = cfg.use_float32 != 0;
    const bool finite =
        r.family == 0
            ? all_finite(r.diag, size_t(r.n)) && all_finite(r.vals, size_t(r.m))
            : all_finite(r.k, size_t(r.n)) && std::isfinite(r.k0) &&
                  holding_finite(r.h, r.n, r.t);
    if (!finite) return kErrorInstance;
    if (r.family == 0) {
        PraqInstance in{r.n, r.m, r.diag, r.rows, r.cols, r.vals};
        int st = praq_build_host(in, r.po.dense_degree, nullptr, nullptr, r.ph);
        if (st) return st;
        r.pl64 = praq_layout(r.ph, sizeof(double));
        r.img64.assign(r.pl64.input_bytes, 0);
        praq_fill_image<double>(r.ph, r.pl64, r.img64.data());
        r.up64 = r.pl64.input_bytes;
        r.bytes = align256(r.pl64.total_bytes);
        if (f32) {
            r.pl32 = praq_layout(r.ph, sizeof(float));
            r.img32.assign(r.pl32.input_bytes, 0);
            praq_fill_image<float>(r.ph, r.pl32, r.img32.data());
            r.up32 = r.pl32.input_bytes;
            r.bytes64 = r.bytes;
            r.bytes = align256(r.pl32.total_bytes);
        }
    } else {
        int st = pria_build_host(r.n, r.t, r.k, r.k0, r.h, nullptr, nullptr, nullptr, r.rh);
        if (st) return kErrorInstance;
        r.rl64 = pria_layout(r.rh.N, r.rh.T, r.rh.ND, sizeof(double));
        r.up64 = pria_upload_bytes(r.rh, r.rl64);
        r.img64.assign(r.up64, 0);
        pria_fill_image<double>(r.rh, r.rl64, r.img64.data());
        r.bytes = align256(r.rl64.total_bytes);
        if (f32) {
            r.rl32 = pria_layout(r.rh.N, r.rh.T, r.rh.ND, sizeof(float));
            r.up32 = pria_upload_bytes(r.rh, r.rl32);
            r.img32.assign(r.up32, 0);
            pria_fill_image<float>(r.rh, r.rl32, r.img32.data());
            r.bytes64 = r.bytes;
            r.bytes = align256(r.rl32.total_bytes);
        }
    }
    return 0;
}

void reform_server::build_loop() {
    for (;;) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lock(mu);
            // build ahead of admission only so far: built images wait in host memory
            cv_build.wait(lock, [&] {
                return stop || (!to_build.empty() &&
                                int(ready.size()) + n_building < cfg.build_ahead);
            });
            if (stop) return;
            r = to_build.front();
            to_build.pop_front();
            ++n_building;
        }
        int st = build(*r);
        // never fits: the first image and, at a handover, the FP64 one
        if (!st && r->bytes + r->bytes64 > pool.capacity()) st = kErrorInstance;
        std::lock_guard<std::mutex> lock(mu);
        --n_building;
        if (r->cancel) {
            finish_unadmitted(*r, kCancelledQueued);
        } else if (st) {
            ++n_rejected;
            finish_unadmitted(*r, st < 0 ? st : kErrorInstance);
        } else {
            ready.push_back(r);
            cv_ready.notify_all();
        }
    }
}

// A request that never reached the device finishes (under mu): rejected,
// or cancelled while it waited.
void reform_server::finish_unadmitted(Request& r, int status) {
    r.status = status;
    r.t_done = seconds_since(t0);
    r.img64 = std::vector<char>();
    r.img32 = std::vector<char>();
    if (status == kCancelledQueued) ++n_cancelled;
    ++n_completed;
    finished.push_back(r.ticket);
    cv_done.notify_all();
    cv_build.notify_one();
}

// -------------------------------------------------------------- admission --

int reform_server::admit(Request& r) {
    const bool f32 = cfg.use_float32 != 0;
    char* base = pool.base() + r.off;   // the first image
    hipStream_t s = upload;
    LogicState* state = states + r.slot;
    // On an FP32 server the FP64 record holds the shapes only (its arrays
    // are bound by grant64 at the handover); no kernel reads an instance's
    // FP64 arrays before its state says in64 or the conversion, which waits
    // for hold64 to clear.
    char* base64 = f32 ? nullptr : base;
    if (r.family == 0) {
        if (f32) {
            REFORM_HIP_CHECK(
                hipMemcpyAsync(base, r.img32.data(), r.up32, hipMemcpyHostToDevice, s));
        } else {
            REFORM_HIP_CHECK(
                hipMemcpyAsync(base, r.img64.data(), r.up64, hipMemcpyHostToDevice, s));
        }
        r.pj64 = PraqJob<double>{};
        r.pj64.a = praq_bind<double>(r.pl64, base64, r.ph);
        r.pj64.p = r.lp;
        r.pj64.scale = r.ph.scale;
        r.pj64.lanes = batch_praq_lanes(r.ph.n, r.ph.m);
        r.pj64.state = state;
        if (f32) {
            r.pj32 = PraqJob<float>{};
            r.pj32.a = praq_bind<float>(r.pl32, base, r.ph);
            r.pj32.p = r.lp;
            apply_fp32_policy(r.pj32.p);
            r.pj32.scale = r.ph.scale;
            r.pj32.lanes = r.pj64.lanes;
            r.pj32.state = state;
        }
    } else {
        r.rj64 = PriaJob<double>{};
        r.rj64.a = pria_bind<double>(r.rl64, base64, r.rh);
        r.rj64.p = r.lp;
        r.rj64.scale = r.rh.scale;
        r.rj64.state = state;
        if (f32) {
            REFORM_HIP_CHECK(
                hipMemcpyAsync(base, r.img32.data(), r.up32, hipMemcpyHostToDevice, s));
            r.rj32 = PriaJob<float>{};
            r.rj32.a = pria_bind<float>(r.rl32, base, r.rh);
            if (r.rh.cold && pria_cold_start<float>(r.rj32.a, s)) return kErrorDevice;
            r.rj32.p = r.lp;
            apply_fp32_policy(r.rj32.p);
            r.rj32.scale = r.rh.scale;
            r.rj32.state = state;
        } else {
            REFORM_HIP_CHECK(
                hipMemcpyAsync(base, r.img64.data(), r.up64, hipMemcpyHostToDevice, s));
            if (r.rh.cold && pria_cold_start<double>(r.rj64.a, s)) return kErrorDevice;
        }
    }
    LogicState init{};
    logic_init(init, r.lp);
    init.action = kActNew;
    init.in64 = f32 ? 0 : 1;
    init.last_action = kActInit;
    init.steps32 = init.steps64 = 0;
    init.switch_iteration = -1;
    init.hold64 = f32 ? 1 : 0;
    r.init = init;
    REFORM_HIP_CHECK(hipMemcpyAsync(state, &r.init, sizeof(init), hipMemcpyHostToDevice, s));
    return 0;
}

// An FP32 request at its handover, its FP64 image just placed at r.off64:
// the image uploaded (and cold-started, as admission does on an FP64 server,
// so the conversion meets the same memory), the FP64 record bound, and the
// state's hold64 cleared -- all on the family's stream before its next
// round, whose conversion then takes the request to FP64.
int reform_server::grant64(Request& r, hipStream_t s) {
    static const int32_t kZero = 0;
    char* base = pool.base() + r.off64;
    REFORM_HIP_CHECK(hipMemcpyAsync(base, r.img64.data(), r.up64, hipMemcpyHostToDevice, s));
    if (r.family == 0) {
        r.pj64.a = praq_bind<double>(r.pl64, base, r.ph);
    } else {
        r.rj64.a = pria_bind<double>(r.rl64, base, r.rh);
        if (r.rh.cold && pria_cold_start<double>(r.rj64.a, s)) return kErrorDevice;
    }
    REFORM_HIP_CHECK(hipMemcpyAsync(&states[r.slot].hold64, &kZero, sizeof(kZero),
                                    hipMemcpyHostToDevice, s));
    return 0;
}

// The family's engine over its current members: job arrays uploaded in
// member order, block tables rebuilt.
int reform_server::rebuild(int family) {
    const bool f32 = cfg.use_float32 != 0;
    auto& mem = members[family];
    const size_t count = mem.size();
    n_rebuilds++;
    auto grow = [&](size_t rec32, size_t rec64) -> int {
        if (count <= jobs_cap[family]) return 0;
        size_t cap = std::max<size_t>(count, 2 * jobs_cap[family]);
        for (int p = 0; p < 2; ++p) {
            if (jobs_dev[family][p]) (void)hipFree(jobs_dev[family][p]);
            jobs_dev[family][p] = nullptr;
        }
        jobs_cap[family] = 0;
        REFORM_HIP_CHECK(device_malloc(&jobs_dev[family][0], rec32 * cap));
        REFORM_HIP_CHECK(device_malloc(&jobs_dev[family][1], rec64 * cap));
        jobs_cap[family] = cap;
        return 0;
    };
    if (family == 0) {
        if (grow(sizeof(PraqJob<float>), sizeof(PraqJob<double>))) return kErrorDevice;
        std::vector<PraqJob<double>> h64(count);
        std::vector<PraqJob<float>> h32(count);
        for (size_t q = 0; q < count; ++q) {
            h64[q] = mem[q]->pj64;
            h32[q] = mem[q]->pj32;
        }
        auto* d64 = static_cast<PraqJob<double>*>(jobs_dev[0][1]);
        auto* d32 = static_cast<PraqJob<float>*>(jobs_dev[0][0]);
        if (count) {
            REFORM_HIP_CHECK(hipMemcpy(d64, h64.data(), sizeof(h64[0]) * count, hipMemcpyHostToDevice));
            if (f32)
                REFORM_HIP_CHECK(
                    hipMemcpy(d32, h32.data(), sizeof(h32[0]) * count, hipMemcpyHostToDevice));
        }
        if (praq_engine.setup(std::move(h64), f32 ? d32 : nullptr, d64)) return kErrorDevice;
    } else {
        if (grow(sizeof(PriaJob<float>), sizeof(PriaJob<double>))) return kErrorDevice;
        std::vector<PriaJob<double>> h64(count);
        std::vector<PriaJob<float>> h32(count);
        for (size_t q = 0; q < count; ++q) {
            h64[q] = mem[q]->rj64;
            h32[q] = mem[q]->rj32;
        }
        auto* d64 = static_cast<PriaJob<double>*>(jobs_dev[1][1]);
        auto* d32 = static_cast<PriaJob<float>*>(jobs_dev[1][0]);
        if (count) {
            REFORM_HIP_CHECK(hipMemcpy(d64, h64.data(), sizeof(h64[0]) * count, hipMemcpyHostToDevice));
            if (f32)
                REFORM_HIP_CHECK(
                    hipMemcpy(d32, h32.data(), sizeof(h32[0]) * count, hipMemcpyHostToDevice));
        }
        if (pria_engine.setup(std::move(h64), f32 ? d32 : nullptr, d64)) return kErrorDevice;
    }
    // the members' own time limits, on the server's clock
    std::vector<double> due(count, std::numeric_limits<double>::infinity());
    bool any = false;
    for (size_t q = 0; q < count; ++q) {
        if (mem[q]->time_limit > 0.0) {
            due[q] = mem[q]->t_admit + mem[q]->time_limit;
            any = true;
        }
        if (mem[q]->cancel) {   // stops at its first round in these tables
            due[q] = -std::numeric_limits<double>::infinity();
            any = true;
        }
    }
    if (!any) due.clear();
    WideControl& w = family == 0 ? static_cast<WideControl&>(praq_engine) : pria_engine;
    w.set_deadlines(std::move(due));
    changed[family] = false;
    return 0;
}

// ------------------------------------------------------------- retirement --

template <class Real>
static int copy_out(std::vector<Real>& out, const Real* src, size_t count, hipStream_t s) {
    out.resize(count);
    if (count)
        REFORM_HIP_CHECK(hipMemcpyAsync(out.data(), src, sizeof(Real) * count,
                                        hipMemcpyDeviceToHost, s));
    return 0;
}

// With `mu` held by the caller's lock below: the pool and the slots are
// shared with admission, which allocates under `mu`.
void reform_server::retire(Request& r, int status) {
    std::lock_guard<std::mutex> lock(mu);
    pool.release(r.off, r.bytes);
    if (r.off64 != SIZE_MAX) pool.release(r.off64, r.bytes64);
    else if (r.bytes64) pending64.erase(pending64.find(r.bytes64));
    bytes_in_use = pool.used();
    free_slots.push_back(r.slot);
    n_in_flight--;
    // stopped by its cancellation (a request that converged or ran out of
    // iterations first keeps that status)
    if (r.cancel && status == 2) status = kCancelled;
    if (status == kCancelled) ++n_cancelled;
    r.status = status;
    r.t_done = seconds_since(t0);
    ++n_completed;
    finished.push_back(r.ticket);
    cv_done.notify_all();
}

// Cancelled requests in the tables stop at the family's next round: their
// deadlines moved to the past (a cancelled request not in the tables yet is
// stopped by the rebuild that brings it in, and one already finished keeps
// its outcome).
void reform_server::stop_cancelled() {
    // under the lock throughout: a request found here cannot be released
    std::lock_guard<std::mutex> lock(mu);
    std::vector<int64_t> todo;
    todo.swap(cancels);
    for (int64_t ticket : todo) {
        auto found = requests.find(ticket);
        if (found == requests.end()) continue;
        Request* r = found->second.get();
        const int f = r->family;
        if (changed[f]) continue;   // the coming rebuild reads the flag
        const auto& mem = members[f];
        auto it = std::find(mem.begin(), mem.end(), r);
        if (it == mem.end()) continue;
        WideControl& w = f == 0 ? static_cast<WideControl&>(praq_engine) : pria_engine;
        w.stop_position(int(it - mem.begin()));
    }
}

// After the family's round: members whose state reads finished (the done
// flags copied back with the round) leave the tables and go to the retire
// thread, which copies their answers out and frees their memory and slots
// while the family's next round (tables rebuilt without them) runs.  FP32
// members at their handover get their FP64 image when the pool has room
// (else they wait for a later round); the tables are then rebuilt with their
// FP64 records.
int reform_server::retire_finished(int f) {
    auto& mem = members[f];
    const int* flag = flags_host[f];
    std::vector<Request*> keep, gone;
    for (size_t q = 0; q < mem.size(); ++q) {
        Request* r = mem[q];
        if (flag[q] == 1) {
            gone.push_back(r);
            continue;
        }
        keep.push_back(r);
        if (flag[q] != 2 || r->off64 != SIZE_MAX) continue;
        size_t off;
        {
            std::lock_guard<std::mutex> lock(mu);
            off = pool.alloc(r->bytes64);
            if (off != SIZE_MAX) {
                pending64.erase(pending64.find(r->bytes64));
                bytes_in_use = pool.used();
                bytes_peak = pool.peak();
            }
        }
        if (off == SIZE_MAX) {
            n_handover_waits++;
            continue;
        }
        r->off64 = off;
        if (grant64(*r, streams[f])) return kErrorDevice;
        n_handovers++;
        changed[f] = true;
    }
    if (!gone.empty()) {
        {
            std::lock_guard<std::mutex> lock(mu);
            retiring.insert(retiring.end(), gone.begin(), gone.end());
        }
        cv_retire.notify_all();
        changed[f] = true;
    }
    mem.swap(keep);
    return 0;
}

// Each retire thread: final state and answer of a finished request, on a
// stream of its own (nothing else touches a request once it has left the
// tables), then its memory and slot back to admission.
void reform_server::retire_loop() {
    auto fail = [&](int code) {
        {
            std::lock_guard<std::mutex> lock(mu);
            start_status = code;
            stop = true;
        }
        notify_every_thread();
    };
    hipStream_t s = nullptr;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        return fail(kErrorDevice);
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } guard{s};
    for (;;) {
        std::vector<Request*> batch;
        {
            std::unique_lock<std::mutex> lock(mu);
            cv_retire.wait(lock, [&] { return stop || !retiring.empty(); });
            if (stop) return;
            batch.push_back(retiring.front());   // one at a time: the threads share the queue
            retiring.pop_front();
        }
        const double t_start = seconds_since(t0);
        for (Request* rp : batch) {
            Request& r = *rp;
            if (hipMemcpyAsync(&r.final, states + r.slot, sizeof(LogicState),
                               hipMemcpyDeviceToHost, s) != hipSuccess ||
                hipStreamSynchronize(s) != hipSuccess)
                return fail(kErrorDevice);
            r.in64 = r.final.in64 != 0;
            int st = 0;
            if (r.family == 0) {
                const size_t n = r.ph.n, m2 = 2 * size_t(r.ph.m);
                st = r.in64 ? (copy_out(r.d1, r.pj64.a.x, n, s) || copy_out(r.d2, r.pj64.a.y, m2, s))
                            : (copy_out(r.f1, r.pj32.a.x, n, s) || copy_out(r.f2, r.pj32.a.y, m2, s));
            } else {
                const size_t nd = r.rh.ND;
                st = r.in64 ? (copy_out(r.d1, r.rj64.a.x, nd, s) ||
                               copy_out(r.d2, r.rj64.a.y, nd, s) ||
                               copy_out(r.d3, r.rj64.a.z, nd, s))
                            : (copy_out(r.f1, r.rj32.a.x, nd, s) ||
                               copy_out(r.f2, r.rj32.a.y, nd, s) ||
                               copy_out(r.f3, r.rj32.a.z, nd, s));
            }
            if (st || hipStreamSynchronize(s) != hipSuccess) return fail(kErrorDevice);
            r.img64 = std::vector<char>();
            if (r.family == 1) {
                std::vector<double> x(size_t(r.n) * r.t * r.t);
                if (r.in64) pria_unpack(r.rh, r.d1, 0, x.data());
                else pria_unpack(r.rh, r.f1, 0, x.data());
                r.upper_original = pria_objective_host(r.n, r.t, r.k, r.k0, r.h, x.data());
            }
            retire(r, r.final.status);
        }
        retire_s = retire_s + (seconds_since(t0) - t_start);
    }
}

// One round of the family on its stream, then its done flags into pinned
// memory and the completion event.  The round is replayed from a graph once
// the tables have held for two rounds, else launched directly (a capture
// costs more than the launches of one round).
int reform_server::launch(int f) {
    WideControl& w = f == 0 ? static_cast<WideControl&>(praq_engine) : pria_engine;
    const int count = int(members[f].size());
    hipStream_t s = streams[f];
    if (count > flags_cap[f]) {
        if (flags_dev[f]) REFORM_HIP_CHECK(hipFree(flags_dev[f]));
        if (flags_host[f]) REFORM_HIP_CHECK(hipHostFree(flags_host[f]));
        flags_dev[f] = flags_host[f] = nullptr;
        flags_cap[f] = std::max(count, 2 * flags_cap[f]);
        REFORM_HIP_CHECK(device_malloc(&flags_dev[f], sizeof(int) * flags_cap[f]));
        REFORM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&flags_host[f]),
                                       sizeof(int) * flags_cap[f], hipHostMallocDefault));
    }
    // members past their own time limits stop at their current points; the
    // round then passes them by and the done flags retire them (status 2)
    if (w.expire_due(s, seconds_since(t0)) < 0) return kErrorDevice;
    w.use_graph = stable[f] >= 2;
    if (w.launch_round(s)) return kErrorDevice;
    ++stable[f];
    const dim3 grid((count + 255) / 256), blk(256);
    if (f == 0)
        hipLaunchKernelGGL(serve_done_flags<PraqJob<double>>, grid, blk, 0, s,
                           static_cast<const PraqJob<double>*>(jobs_dev[0][1]), count, flags_dev[0]);
    else
        hipLaunchKernelGGL(serve_done_flags<PriaJob<double>>, grid, blk, 0, s,
                           static_cast<const PriaJob<double>*>(jobs_dev[1][1]), count, flags_dev[1]);
    REFORM_HIP_CHECK(hipGetLastError());
    REFORM_HIP_CHECK(hipMemcpyAsync(flags_host[f], flags_dev[f], sizeof(int) * count,
                                    hipMemcpyDeviceToHost, s));
    REFORM_HIP_CHECK(hipEventRecord(round_done[f], s));
    in_flight[f] = true;
    n_rounds++;
    return 0;
}

// ---------------------------------------------------------- serving thread --

void reform_server::serve_loop() {
    auto fail = [&](int code) {
        // no round or copy of this thread's left running on the pool
        for (hipStream_t st : streams)
            if (st) (void)hipStreamSynchronize(st);
        if (upload) (void)hipStreamSynchronize(upload);
        {
            std::lock_guard<std::mutex> lock(mu);
            start_status = code;
            stop = true;
        }
        notify_every_thread();
    };
    if (hipSetDevice(device) != hipSuccess) return fail(kErrorDevice);
    for (int f = 0; f < 2; ++f) {
        if (hipStreamCreateWithFlags(&streams[f], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&round_done[f], hipEventDisableTiming) != hipSuccess)
            return fail(kErrorDevice);
    }
    if (hipStreamCreateWithFlags(&upload, hipStreamNonBlocking) != hipSuccess)
        return fail(kErrorDevice);
    while (!stop) {
        bool progress = false;
        // admit what fits, in submission order; the copies run on their own
        // stream beside the rounds in flight
        std::vector<Request*> admitted;
        {
            std::unique_lock<std::mutex> lock(mu);
            const bool idle = !in_flight[0] && !in_flight[1] && members[0].empty() &&
                              members[1].empty() && joining[0].empty() && joining[1].empty();
            if (idle && ready.empty())
                cv_ready.wait_for(lock, std::chrono::milliseconds(5),
                                  [&] { return stop || !ready.empty(); });
            if (stop) break;
            while (!ready.empty() && !free_slots.empty()) {
                Request* r = ready.front();
                // room kept for the largest FP64 image a request in flight
                // (this one included) may still ask for at its handover
                const size_t headroom =
                    r->bytes64 ? std::max(r->bytes64, pending64.empty() ? size_t(0)
                                                                        : *pending64.rbegin())
                               : size_t(0);
                const size_t off = pool.alloc(r->bytes, headroom);
                if (off == SIZE_MAX) break;  // wait for memory: the queue keeps its order
                if (r->bytes64) pending64.insert(r->bytes64);
                r->off = off;
                r->slot = free_slots.back();
                free_slots.pop_back();
                ready.pop_front();
                admitted.push_back(r);
            }
            bytes_in_use = pool.used();
            bytes_peak = pool.peak();
        }
        if (!admitted.empty()) {
            cv_build.notify_all();
            for (Request* r : admitted)
                if (admit(*r)) return fail(kErrorDevice);
            if (hipStreamSynchronize(upload) != hipSuccess) return fail(kErrorDevice);
            const double now = seconds_since(t0);
            {
                std::lock_guard<std::mutex> lock(mu);
                n_admitted += int64_t(admitted.size());
                for (Request* r : admitted) r->t_admit = now;
            }
            for (Request* r : admitted) {
                // the device holds the first image now (an FP32 request's
                // FP64 image waits in host memory for its handover)
                if (!r->bytes64) r->img64 = std::vector<char>();
                r->img32 = std::vector<char>();
                const int64_t now_in = ++n_in_flight;
                int64_t peak = n_in_flight_peak.load();
                while (now_in > peak && !n_in_flight_peak.compare_exchange_weak(peak, now_in)) {
                }
                joining[r->family].push_back(r);
            }
            progress = true;
        }
        stop_cancelled();
        // each family on its own: retire after its round, then the next round
        for (int f = 0; f < 2; ++f) {
            if (in_flight[f]) {
                const hipError_t q = hipEventQuery(round_done[f]);
                if (q == hipErrorNotReady) continue;
                if (q != hipSuccess) return fail(kErrorDevice);
                in_flight[f] = false;
                if (retire_finished(f)) return fail(kErrorDevice);
                progress = true;
            }
            if (!joining[f].empty()) {
                members[f].insert(members[f].end(), joining[f].begin(), joining[f].end());
                joining[f].clear();
                changed[f] = true;
            }
            if (members[f].empty()) continue;
            if (changed[f]) {
                const double t_rebuild = seconds_since(t0);
                if (rebuild(f)) return fail(kErrorDevice);
                rebuild_s = rebuild_s + (seconds_since(t0) - t_rebuild);
                stable[f] = 0;
            }
            if (launch(f)) return fail(kErrorDevice);
            progress = true;
        }
        if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    // leave no round running on the images the pool is about to free
    for (int f = 0; f < 2; ++f)
        if (streams[f]) (void)hipStreamSynchronize(streams[f]);
}

// ------------------------------------------------------------------ C ABI --

extern "C" reform_server_config reform_server_default_config(void) {
    reform_server_config c;
    c.use_float32 = 1;
    c.host_threads = 4;
    c.hbm_budget = int64_t(200) * 1000 * 1000 * 1000;
    c.max_in_flight = 16384;
    c.borrow_inputs = 0;
    c.build_ahead = 32;
    return c;
}

extern "C" reform_server* reform_server_create(const reform_server_config* config) {
    reform_server_config c = config ? *config : reform_server_default_config();
    if (c.host_threads <= 0) c.host_threads = 4;
    if (c.hbm_budget <= 0) c.hbm_budget = int64_t(200) * 1000 * 1000 * 1000;
    if (c.max_in_flight <= 0) c.max_in_flight = 16384;
    if (c.build_ahead <= 0) c.build_ahead = 32;
    auto srv = std::make_unique<reform_server>();
    srv->cfg = c;
    if (hipGetDevice(&srv->device) != hipSuccess) return nullptr;
    if (srv->pool.init(size_t(c.hbm_budget))) return nullptr;
    if (device_malloc(&srv->states, sizeof(LogicState) * size_t(c.max_in_flight)) != hipSuccess) {
        srv->states = nullptr;
        return nullptr;
    }
    for (int64_t q = c.max_in_flight - 1; q >= 0; --q) srv->free_slots.push_back(int(q));
    reform_server* raw = srv.get();
    for (int q = 0; q < c.host_threads; ++q) raw->builders.emplace_back([raw] { raw->build_loop(); });
    raw->server = std::thread([raw] { raw->serve_loop(); });
    for (int q = 0; q < kRetireThreads; ++q)
        raw->retirers.emplace_back([raw] { raw->retire_loop(); });
    return srv.release();
}

extern "C" void reform_server_destroy(reform_server* server) { delete server; }

static int64_t enqueue(reform_server* srv, std::unique_ptr<Request> r) {
    std::lock_guard<std::mutex> lock(srv->mu);
    if (srv->stop) return srv->start_status ? srv->start_status : kErrorDevice;
    r->ticket = srv->next_ticket++;
    r->t_submit = seconds_since(srv->t0);
    const int64_t ticket = r->ticket;
    srv->to_build.push_back(r.get());
    srv->requests[ticket] = std::move(r);
    ++srv->n_submitted;
    srv->cv_build.notify_one();
    return ticket;
}

extern "C" int64_t reform_server_submit_praq(reform_server* server, const PraqInstance* instance,
                                             const PraqParameters* parameters,
                                             const PraqOptions* options) {
    if (!server || !instance || !parameters || !praq_params_valid(*parameters) ||
        (options && !(options->eta > 0.0)))
        return kErrorArgument;
    if (instance->n <= 0 || instance->m < 0 || !instance->diag ||
        (instance->m > 0 && (!instance->rows || !instance->cols || !instance->vals)))
        return kErrorInstance;
    auto r = std::make_unique<Request>();
    r->family = 0;
    r->n = instance->n;
    r->m = instance->m;
    r->diag = instance->diag;
    r->rows = instance->rows;
    r->cols = instance->cols;
    r->vals = instance->vals;
    if (!server->cfg.borrow_inputs) {
        r->own_diag.assign(instance->diag, instance->diag + instance->n);
        r->diag = r->own_diag.data();
        if (instance->m > 0) {
            r->own_rows.assign(instance->rows, instance->rows + instance->m);
            r->own_cols.assign(instance->cols, instance->cols + instance->m);
            r->own_vals.assign(instance->vals, instance->vals + instance->m);
            r->rows = r->own_rows.data();
            r->cols = r->own_cols.data();
            r->vals = r->own_vals.data();
        }
    }
    PraqParameters p = *parameters;
    if (p.max_iter <= 0) p.max_iter = -1;
    if (p.op_steps <= 0) p.op_steps = praq_check_interval(instance->n, &p);
    if (p.op_steps < 2) p.op_steps = 2;
    r->po = options ? *options : praq_default_options();
    r->lp = praq_logic_params(p, r->po);
    r->time_limit = r->po.time_limit;
    return enqueue(server, std::move(r));
}

extern "C" int64_t reform_server_submit_pria(reform_server* server, int64_t n, int64_t t,
                                             const double* order_costs, double joint_cost,
                                             const double* holding_costs, const pria_params* params,
                                             const pria_options* options) {
    if (!server || !pria_params_valid(params) || n <= 0 || t <= 0 || !order_costs ||
        !holding_costs)
        return kErrorArgument;
    auto r = std::make_unique<Request>();
    r->family = 1;
    if (options) r->time_limit = options->time_limit;
    r->n = n;
    r->t = t;
    r->k = order_costs;
    r->h = holding_costs;
    if (!server->cfg.borrow_inputs) {
        r->own_k.assign(order_costs, order_costs + n);
        r->own_h.assign(holding_costs, holding_costs + size_t(n) * t * t);
        r->k = r->own_k.data();
        r->h = r->own_h.data();
    }
    r->k0 = joint_cost;
    r->lp = pria_logic_params(*params);
    return enqueue(server, std::move(r));
}

extern "C" int64_t reform_server_poll(reform_server* server, int64_t* tickets, int64_t max) {
    if (!server || (!tickets && max > 0)) return kErrorArgument;
    std::lock_guard<std::mutex> lock(server->mu);
    int64_t k = 0;
    while (k < max && !server->finished.empty()) {
        tickets[k++] = server->finished.front();
        server->finished.pop_front();
    }
    return k;
}

extern "C" int reform_server_wait_idle(reform_server* server, double timeout_seconds) {
    if (!server) return kErrorArgument;
    std::unique_lock<std::mutex> lock(server->mu);
    auto idle = [&] { return server->stop || server->n_completed == server->n_submitted; };
    if (timeout_seconds > 0.0) {
        if (!server->cv_done.wait_for(lock, std::chrono::duration<double>(timeout_seconds), idle))
            return kErrorTimeout;
    } else {
        server->cv_done.wait(lock, idle);
    }
    return server->start_status;
}

static Request* find(reform_server* server, int64_t ticket) {
    auto it = server->requests.find(ticket);
    return it == server->requests.end() ? nullptr : it->second.get();
}

extern "C" int reform_server_info(reform_server* server, int64_t ticket, reform_request_info* info) {
    if (!server || !info) return kErrorArgument;
    std::lock_guard<std::mutex> lock(server->mu);
    Request* r = find(server, ticket);
    if (!r) return kErrorTicket;
    std::memset(info, 0, sizeof(*info));
    info->ticket = ticket;
    info->family = r->family;
    info->status = r->status;
    info->submitted = r->t_submit;
    info->admitted = r->t_admit;
    info->completed = r->t_done;
    info->device_bytes = int64_t(r->bytes + (r->off64 != SIZE_MAX ? r->bytes64 : 0));
    info->switch_iteration = -1;
    info->bracket = std::numeric_limits<double>::quiet_NaN();
    if (r->status >= 0) {
        info->iterations = r->final.iteration;
        info->restarts = r->final.restarts;
        info->switch_iteration = r->final.switch_iteration;
        info->lower = r->final.lower;
        info->upper = r->final.upper;
        info->relative_gap = r->final.gap;
        if (r->family == 1) info->upper = r->upper_original;
        info->bracket = info->upper - info->lower;
    }
    return 0;
}

extern "C" int reform_server_praq_result(reform_server* server, int64_t ticket, double* x,
                                         double* y) {
    if (!server) return kErrorArgument;
    std::lock_guard<std::mutex> lock(server->mu);
    Request* r = find(server, ticket);
    if (!r || r->family != 0 || r->status < 0) return kErrorTicket;
    const size_t n = r->ph.n, m = r->ph.m;
    for (size_t i = 0; x && i < n; ++i) x[i] = r->in64 ? r->d1[i] : double(r->f1[i]);
    for (size_t e = 0; y && e < m; ++e) {
        const double a = r->in64 ? r->d2[e] : double(r->f2[e]);
        const double b = r->in64 ? r->d2[m + e] : double(r->f2[m + e]);
        y[2 * e] = a * r->ph.scale;
        y[2 * e + 1] = b * r->ph.scale;
    }
    return 0;
}

extern "C" int reform_server_pria_result(reform_server* server, int64_t ticket, double* x,
                                         double* y, double* z) {
    if (!server) return kErrorArgument;
    std::lock_guard<std::mutex> lock(server->mu);
    Request* r = find(server, ticket);
    if (!r || r->family != 1 || r->status < 0) return kErrorTicket;
    double* outs[3] = {x, y, z};
    for (int w = 0; w < 3; ++w) {
        if (!outs[w]) continue;
        if (r->in64) pria_unpack(r->rh, w == 0 ? r->d1 : (w == 1 ? r->d2 : r->d3), w, outs[w]);
        else pria_unpack(r->rh, w == 0 ? r->f1 : (w == 1 ? r->f2 : r->f3), w, outs[w]);
    }
    return 0;
}

extern "C" int reform_server_cancel(reform_server* server, int64_t ticket) {
    if (!server) return kErrorArgument;
    std::lock_guard<std::mutex> lock(server->mu);
    Request* r = find(server, ticket);
    if (!r || r->status != kPending || r->cancel) return kErrorTicket;
    r->cancel = true;
    auto drop = [&](std::deque<Request*>& q) {
        auto it = std::find(q.begin(), q.end(), r);
        if (it == q.end()) return false;
        q.erase(it);
        return true;
    };
    if (drop(server->to_build) || drop(server->ready))
        server->finish_unadmitted(*r, kCancelledQueued);
    else if (r->slot >= 0)
        server->cancels.push_back(ticket);
    // else on a builder, which finishes it when the build ends
    return 0;
}

extern "C" int reform_server_release(reform_server* server, int64_t ticket) {
    if (!server) return kErrorArgument;
    std::lock_guard<std::mutex> lock(server->mu);
    Request* r = find(server, ticket);
    if (!r || (r->status == kPending)) return kErrorTicket;
    server->requests.erase(ticket);
    return 0;
}

extern "C" int reform_server_stats(reform_server* server, reform_server_counters* stats) {
    if (!server || !stats) return kErrorArgument;
    std::lock_guard<std::mutex> lock(server->mu);
    stats->submitted = server->n_submitted;
    stats->admitted = server->n_admitted;
    stats->completed = server->n_completed;
    stats->rejected = server->n_rejected;
    stats->rounds = server->n_rounds;
    stats->rebuilds = server->n_rebuilds;
    stats->in_flight = server->n_in_flight;
    stats->bytes_in_use = int64_t(server->bytes_in_use);
    stats->bytes_peak = int64_t(server->bytes_peak);
    stats->bytes_budget = int64_t(server->pool.capacity());
    stats->in_flight_peak = server->n_in_flight_peak;
    stats->handovers = server->n_handovers;
    stats->handover_waits = server->n_handover_waits;
    stats->cancelled = server->n_cancelled;
    stats->retire_seconds = server->retire_s;
    stats->rebuild_seconds = server->rebuild_s;
    return server->start_status;
}
