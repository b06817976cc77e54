// The C ABI of include/pria_hip.h.  Entry points follow
// jrp/pria-cuda/src/pria.cu:15-134; the batch runs every instance of a
// same-shape batch as one workgroup of one launch instead of a thread fan-out.
#include "pria_hip.h"
#include "pria_host.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

using namespace reform;

namespace {

constexpr double kDefaultWgLimitSeconds = 600.0;

// Below this many entries (ND) the one-workgroup tier runs a plain solve.
int64_t pria_wg_max() {
    static int64_t v = [] {
        const char* env = std::getenv("REFORM_PRIA_WG_MAX");
        return env ? std::atoll(env) : int64_t(1) << 15;
    }();
    return v;
}

double wall_clock_rate_hz() {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0.0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess) return 0.0;
    return double(khz) * 1000.0;
}

struct Buffer {
    void* ptr = nullptr;
    ~Buffer() {
        if (ptr) (void)hipFree(ptr);
    }
};

// Device memory in a stream's order (hipMallocAsync / hipFreeAsync): a solve
// neither synchronizes the device nor waits for other threads' solves when it
// allocates or frees, so concurrent solves on their own streams compose.  A
// borrowed buffer is memory the caller provides; it is not freed.
struct StreamBuffer {
    void* ptr = nullptr;
    hipStream_t stream = nullptr;
    bool owned = false;
    int alloc(size_t bytes, hipStream_t s) {
        stream = s;
        if (device_malloc_async(&ptr, bytes ? bytes : 1, s) != hipSuccess) {
            ptr = nullptr;
            return kPriaErrorDevice;
        }
        owned = true;
        return 0;
    }
    void borrow(void* p) {
        ptr = p;
        owned = false;
    }
    ~StreamBuffer() {
        if (ptr && owned) (void)hipFreeAsync(ptr, stream);
    }
};

struct Stream {
    hipStream_t s = nullptr;
    ~Stream() {
        if (s) (void)hipStreamDestroy(s);
    }
};



// An instance whose arrays are in device memory: order_costs [n] and the
// dense holding costs [n][t][t] (upper triangle read), both FP64.
struct PriaDeviceInput {
    const double* k = nullptr;
    const double* h = nullptr;
    int t = 0;
};

constexpr int kDevBlock = 256;
constexpr int kDevScaleBlocks = 512;

// Sum of squares of the upper triangles of h's rows (i, a); rows are dealt
// to blocks as r = blockIdx.x + q gridDim.x, so the order of additions is
// fixed by the grid.
__global__ void __launch_bounds__(kDevBlock) pria_sumsq_rows(const double* h, int64_t rows, int t,
                                                             double* partial) {
    __shared__ double red[16];
    double acc = 0.0;
    for (int64_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const int a = int(r % t);
        const double* row = h + r * t;
        for (int b = a + threadIdx.x; b < t; b += blockDim.x) acc += row[b] * row[b];
    }
    acc = block_sum1(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc;
}

// ||(H, K, K0)||^2: the row partials, then K, then K0 (one block).
__global__ void __launch_bounds__(kDevBlock) pria_sumsq_total(const double* k, int n, double k0,
                                                              const double* partial, int parts,
                                                              double* out) {
    __shared__ double red[16];
    double acc = 0.0;
    for (int p = threadIdx.x; p < parts; p += blockDim.x) acc += partial[p];
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += k[i] * k[i];
    acc = block_sum1(acc, red);
    if (threadIdx.x == 0) out[0] = acc + k0 * k0;
}

// The scaled image from device inputs: one block per row (i, a) of h, its
// upper triangle to slice order (pria_build_host + pria_fill_image); block
// (i, 0) also writes K_i / scale.
template <class Real>
__global__ void __launch_bounds__(kDevBlock) pria_gather_rows(const double* k, const double* h,
                                                              int N, int t, int T, double scale,
                                                              Real* ak, Real* ah) {
    const int64_t r = blockIdx.x;
    const int i = int(r / t), a = int(r % t);
    const double* row = h + r * t;
    Real* dst = ah + pria_off(N, T, a) + int64_t(i) * (T - a) - a;
    for (int b = a + threadIdx.x; b < t; b += blockDim.x) dst[b] = Real(row[b] / scale);
    if (a == 0 && threadIdx.x == 0) ak[i] = Real(k[i] / scale);
}

// Slice order -> the caller's dense [n][t][t] (pria_unpack): one block per
// row (i, a); which 1: y / K_i, 2: z / K0 (0 where that budget is 0).
template <class Real>
__global__ void __launch_bounds__(kDevBlock) pria_scatter_rows(const Real* v, const double* k,
                                                               double k0s, int which, int N, int t,
                                                               int T, double scale, double* out) {
    const int64_t r = blockIdx.x;
    const int i = int(r / t), a = int(r % t);
    double norm = 1.0;
    if (which == 1) norm = k[i] / scale;
    if (which == 2) norm = k0s;
    const Real* src = v + pria_off(N, T, a) + int64_t(i) * (T - a) - a;
    double* dst = out + r * t;
    for (int b = threadIdx.x; b < t; b += blockDim.x) {
        double val = 0.0;
        if (b >= a) {
            val = double(src[b]);
            if (which) val = norm > 0.0 ? val / norm : 0.0;
        }
        dst[b] = val;
    }
}

// primal_objective of dense x (pria_objective_host), per row (i, a): the
// row's max over all t and its <h, x> over the upper triangle.
__global__ void __launch_bounds__(kDevBlock) pria_objective_rows(const double* h, const double* x,
                                                                 int t, double* row_max,
                                                                 double* row_dot) {
    __shared__ double red[16];
    const int64_t r = blockIdx.x;
    const int a = int(r % t);
    const double* hr = h + r * t;
    const double* xr = x + r * t;
    double mx = -INFINITY, dot = 0.0;
    for (int b = threadIdx.x; b < t; b += blockDim.x) {
        mx = fmax(mx, xr[b]);
        if (b >= a) dot += hr[b] * xr[b];
    }
    mx = block_max1(mx, red);
    dot = block_sum1(dot, red);
    if (threadIdx.x == 0) {
        row_max[r] = mx;
        row_dot[r] = dot;
    }
}

// sum_{i,a} K_i max_b x + K0 sum_a max_{i,b} x + <H, x> (one block).
__global__ void __launch_bounds__(kDevBlock) pria_objective_total(const double* k, double k0,
                                                                  int N, int t,
                                                                  const double* row_max,
                                                                  const double* row_dot,
                                                                  double* out) {
    __shared__ double red[16];
    double acc = 0.0;
    for (int64_t r = threadIdx.x; r < int64_t(N) * t; r += blockDim.x)
        acc += k[r / t] * row_max[r] + row_dot[r];
    for (int a = threadIdx.x; a < t; a += blockDim.x) {
        double mx = -INFINITY;
        for (int i = 0; i < N; ++i) mx = fmax(mx, row_max[int64_t(i) * t + a]);
        acc += k0 * mx;
    }
    acc = block_sum1(acc, red);
    if (threadIdx.x == 0) out[0] = acc;
}

int pick_tier(const PriaHostInstance& s, int tier) {
    if (tier != PRIA_TIER_AUTO) return tier;
    return s.ND <= pria_wg_max() && s.T <= kPriaWgMaxT ? PRIA_TIER_WORKGROUP
                                                        : PRIA_TIER_DEVICE;
}

// One instance's device image in one precision, and the tier that runs it.
template <class Real>
struct PriaImage {
    StreamBuffer buf;
    PriaLayout L;
    PriaJob<Real> job;
    size_t job_off = 0, state_off = 0, work_off = 0, bytes = 0;

    void place(const PriaHostInstance& s) {
        L = pria_layout(s.N, s.T, s.ND, sizeof(Real));
        job_off = pria_align256(L.total_bytes);
        state_off = pria_align256(job_off + sizeof(PriaJob<Real>));
        work_off = pria_align256(state_off + sizeof(LogicState));
        bytes = work_off + pria_device_work_bytes(s.N, s.T);
    }

    // device bytes of the image of an instance of s's shape
    static size_t size_for(const PriaHostInstance& s) {
        PriaImage probe;
        probe.place(s);
        return probe.bytes;
    }

    // `din` set: K and h gathered on the device from the caller's arrays
    // (cold start only); else the host image of `s` uploaded.  `at` set: the
    // image lives there (size_for(s) bytes, 256-byte aligned), else in a
    // stream-ordered allocation.
    int init(const PriaHostInstance& s, const LogicParams& lp, hipStream_t stream,
             const PriaDeviceInput* din = nullptr, char* at = nullptr) {
        place(s);
        if (at) buf.borrow(at);
        else if (buf.alloc(bytes, stream)) return kPriaErrorDevice;
        char* base = static_cast<char*>(buf.ptr);
        std::memset(&job, 0, sizeof(job));
        job.a = pria_bind<Real>(L, base, s);
        std::vector<char> image;
        if (din) {
            Real* k = reinterpret_cast<Real*>(base + L.off[PriaLayout::kK]);
            Real* h = reinterpret_cast<Real*>(base + L.off[PriaLayout::kH]);
            if (s.T != din->t &&  // padded entries of h are zero
                hipMemsetAsync(h, 0, sizeof(Real) * size_t(s.ND), stream) != hipSuccess)
                return kPriaErrorDevice;
            hipLaunchKernelGGL(pria_gather_rows<Real>, dim3(unsigned(int64_t(s.N) * din->t)),
                               dim3(kDevBlock), 0, stream, din->k, din->h, s.N, din->t, s.T,
                               s.scale, k, h);
            if (hipGetLastError() != hipSuccess) return kPriaErrorDevice;
        } else {
            const size_t up = pria_upload_bytes(s, L);
            image.assign(up, 0);
            pria_fill_image<Real>(s, L, image.data());
            if (hipMemcpyAsync(base, image.data(), up, hipMemcpyHostToDevice, stream) !=
                hipSuccess)
                return kPriaErrorDevice;
        }
        if (s.cold && pria_cold_start<Real>(job.a, stream)) return kPriaErrorDevice;
        job.p = lp;
        job.scale = s.scale;
        job.state = reinterpret_cast<LogicState*>(base + state_off);
        return hipStreamSynchronize(stream) == hipSuccess ? 0 : kPriaErrorDevice;
    }

    int run(int tier, LogicState& st, double time_limit, bool resume, int T, hipStream_t stream) {
        char* base = static_cast<char*>(buf.ptr);
        if (tier == PRIA_TIER_WORKGROUP) {
            double limit = time_limit > 0.0 ? std::min(time_limit, kDefaultWgLimitSeconds)
                                            : kDefaultWgLimitSeconds;
            job.time_limit_ticks = (long long)(limit * wall_clock_rate_hz());
            job.resume = resume ? 1 : 0;
            if (resume && hipMemcpyAsync(job.state, &st, sizeof(st), hipMemcpyHostToDevice,
                                         stream) != hipSuccess)
                return kPriaErrorDevice;
            if (hipMemcpyAsync(base + job_off, &job, sizeof(job), hipMemcpyHostToDevice, stream) !=
                hipSuccess)
                return kPriaErrorDevice;
            if (pria_launch_wg<Real>(reinterpret_cast<PriaJob<Real>*>(base + job_off), 1, T,
                                     stream))
                return kPriaErrorDevice;
            if (hipMemcpyAsync(&st, job.state, sizeof(st), hipMemcpyDeviceToHost, stream) !=
                hipSuccess)
                return kPriaErrorDevice;
            return hipStreamSynchronize(stream) == hipSuccess ? 0 : kPriaErrorDevice;
        }
        return pria_device_solve<Real>(job, st, reinterpret_cast<double*>(base + work_off),
                                       time_limit, stream, resume);
    }

    // slice-order x, y, z of the final point, widened to FP64
    int outputs(const PriaHostInstance& s, std::vector<double>* out[3], hipStream_t stream) {
        Real* srcs[3] = {job.a.x, job.a.y, job.a.z};
        std::vector<Real> v(s.ND);
        for (int q = 0; q < 3; ++q) {
            if (!out[q]) continue;
            if (hipMemcpyAsync(v.data(), srcs[q], sizeof(Real) * s.ND, hipMemcpyDeviceToHost,
                               stream) != hipSuccess ||
                hipStreamSynchronize(stream) != hipSuccess)
                return kPriaErrorDevice;
            out[q]->assign(v.begin(), v.end());
        }
        return 0;
    }
};

// Device bytes of solve_mixed's images when they are carved from an arena:
// the FP64 image, or the FP32 image and the FP64 one of a handover.
size_t mixed_image_bytes(const PriaHostInstance& s, bool fp32) {
    const size_t d = PriaImage<double>::size_for(s);
    return fp32 ? pria_align256(PriaImage<float>::size_for(s)) + d : d;
}

// FP64 iterates, or FP32 iterates handing over to FP64 when they stall; all
// work on `stream`.  `emit(image)` takes the answer of the image the solve
// finished in.  `arena` set: the images are carved from it (at least
// mixed_image_bytes(s, fp32) bytes, 256-byte aligned) instead of allocated.
template <class Emit>
int solve_mixed(const PriaHostInstance& s, const LogicParams& lp, const pria_options& o,
                bool fp32, const PriaDeviceInput* din, hipStream_t stream, Emit&& emit,
                pria_info* info, char* arena = nullptr) {
    const int tier = pick_tier(s, o.tier);
    if (tier == PRIA_TIER_WORKGROUP && s.T > kPriaWgMaxT) return kPriaErrorArgument;
    LogicState st;
    std::memset(&st, 0, sizeof(st));
    size_t held = 0;
    if (!fp32) {
        PriaImage<double> d;
        if (d.init(s, lp, stream, din, arena) ||
            d.run(tier, st, o.time_limit, false, s.T, stream) || emit(d))
            return kPriaErrorDevice;
        held = d.bytes;
    } else {
        LogicParams lp32 = lp;
        apply_fp32_policy(lp32);
        PriaImage<float> f;
        if (f.init(s, lp32, stream, din, arena) ||
            f.run(tier, st, o.time_limit, false, s.T, stream))
            return kPriaErrorDevice;
        held = f.bytes;
        if (st.stalled) {
            PriaImage<double> d;
            st.switch_iteration = st.iteration;
            char* at = arena ? arena + pria_align256(f.bytes) : nullptr;
            if (d.init(s, lp, stream, din, at) || pria_convert_state(f.job.a, d.job.a, stream) ||
                d.run(tier, st, o.time_limit, true, s.T, stream) || emit(d))
                return kPriaErrorDevice;
            held += d.bytes;
        } else if (emit(f)) {
            return kPriaErrorDevice;
        }
    }
    if (info) {
        info->iterations = st.iteration;
        info->restarts = st.restarts;
        info->checks = st.checks;
        info->lower = st.lower;
        info->upper_iterate = st.upper;
        info->relative_gap = st.gap;
        info->omega = st.omega;
        info->scale = s.scale;
        info->status = st.status;
        info->tier = tier;
        info->switch_iteration = st.switch_iteration;
        info->device_bytes = int64_t(held);
    }
    return 0;
}

// Shared body of every single-instance entry: FP64 in and out.
int solve_f64io(int64_t n, int64_t t, const double* k, double k0, const double* h,
                const pria_params* params, const pria_options* options, int use_float32,
                const double* wx, const double* wy, const double* wz, double* x, double* y,
                double* z, pria_info* info) {
    if (info) {
        std::memset(info, 0, sizeof(*info));
        info->lower = info->upper = info->relative_gap = info->bracket =
            std::numeric_limits<double>::quiet_NaN();
        info->status = kPriaErrorArgument;
    }
    if (!pria_params_valid(params)) return kPriaErrorArgument;
    pria_options o = options ? *options : pria_default_options();
    PriaHostInstance s;
    if (pria_build_host(n, t, k, k0, h, wx, wy, wz, s)) return kPriaErrorArgument;
    LogicParams lp = pria_logic_params(*params);
    pria_info local;
    pria_info* out = info ? info : &local;
    std::vector<double> xd(size_t(n) * t * t);
    std::vector<double> vx, vy, vz;
    std::vector<double>* outs[3] = {&vx, y ? &vy : nullptr, z ? &vz : nullptr};
    Stream stream;
    if (hipStreamCreateWithFlags(&stream.s, hipStreamNonBlocking) != hipSuccess)
        return kPriaErrorDevice;
    int status = solve_mixed(s, lp, o, use_float32 != 0, nullptr, stream.s,
                             [&](auto& image) { return image.outputs(s, outs, stream.s); }, out);
    if (status == 0) {
        pria_unpack(s, vx, 0, xd.data());
        if (y) pria_unpack(s, vy, 1, y);
        if (z) pria_unpack(s, vz, 2, z);
    }
    if (status) {
        out->status = status;
        return status;
    }
    out->upper = pria_objective_host(n, t, k, k0, h, xd.data());
    out->bracket = out->upper - out->lower;
    if (x) std::memcpy(x, xd.data(), sizeof(double) * xd.size());
    return out->status == 0 ? 0 : 1;
}

template <class Real>
std::vector<double> widen(const Real* p, size_t count) {
    std::vector<double> v(count);
    for (size_t q = 0; q < count; ++q) v[q] = double(p[q]);
    return v;
}

// pria_solve_f64/f32: the reference's return convention, 0 or a device code.
template <class Real>
int solve_reference_abi(int64_t n, int64_t t, const Real* k, Real k0, const Real* h,
                        const pria_params* params, const Real* wx, const Real* wy,
                        const Real* wz, Real* x, Real* y, Real* z, pria_result_info* info) {
    if (n <= 0 || t <= 0 || !k || !h || !pria_params_valid(params)) return kPriaErrorArgument;
    const size_t sz = size_t(n) * t * t;
    std::vector<double> kd = widen(k, n), hd = widen(h, sz), wxd, wyd, wzd;
    if (wx) wxd = widen(wx, sz);
    if (wy) wyd = widen(wy, sz);
    if (wz) wzd = widen(wz, sz);
    std::vector<double> xd(sz), yd(sz), zd(sz);
    pria_info pi;
    int st = solve_f64io(n, t, kd.data(), double(k0), hd.data(), params, nullptr,
                         sizeof(Real) == 4, wx ? wxd.data() : nullptr, wy ? wyd.data() : nullptr,
                         wz ? wzd.data() : nullptr, xd.data(), y ? yd.data() : nullptr,
                         z ? zd.data() : nullptr, &pi);
    if (st < 0) return st;
    for (size_t q = 0; q < sz; ++q) {
        if (x) x[q] = Real(xd[q]);
        if (y) y[q] = Real(yd[q]);
        if (z) z[q] = Real(zd[q]);
    }
    if (info) {
        info->iterations = pi.iterations;
        info->primal_objective = pi.upper_iterate;
        info->dual_objective = pi.lower;
        info->relative_gap = pi.relative_gap;
    }
    return 0;
}

// pria_solve_batch_f64/f32: independent solo solves fanned out over at most
// 16 host threads, each solve on a non-blocking stream of its own, so small
// instances overlap on the device while each computes exactly what a solo
// pria_solve_f64/f32 computes -- the same tier choice, arithmetic and
// summation order, hence bitwise the solo answer (the promise of
// jrp/pria-cuda/include/pria/pria.h:85-99; the same fan-out as
// jrp/pria-cuda/src/pria.cu:15-44).  Returns 0 when every instance succeeded,
// else the error code of a failing one.
template <class Real>
int solve_batch(int64_t count, int64_t n, int64_t t, const Real* k, const Real* k0,
                const Real* h, const pria_params* params, Real* x, Real* y, Real* z,
                pria_result_info* infos) {
    if (count <= 0 || n <= 0 || t <= 0 || !k || !k0 || !h || !pria_params_valid(params))
        return kPriaErrorArgument;
    const size_t sz = size_t(n) * t * t;
    const int workers = int(std::min<int64_t>(count, 16));
    std::atomic<int64_t> next(0);
    std::atomic<int> err(0);
    auto work = [&]() {
        for (int64_t b; (b = next.fetch_add(1)) < count;) {
            const int st = solve_reference_abi<Real>(
                n, t, k + b * n, k0[b], h + b * sz, params, nullptr, nullptr, nullptr,
                x ? x + b * sz : nullptr, y ? y + b * sz : nullptr, z ? z + b * sz : nullptr,
                infos ? infos + b : nullptr);
            if (st != 0) err.store(st);
        }
    };
    std::vector<std::thread> threads;
    for (int w = 1; w < workers; ++w) threads.emplace_back(work);
    work();
    for (auto& th : threads) th.join();
    return err.load();
}

}  // namespace

extern "C" pria_params pria_params_auto(int64_t max_iter, double tolerance) {
    pria_params p;
    p.omega = 0.10835;
    p.eta = 0.999 / std::sqrt(2.0);
    p.beta_sufficient = std::exp(-1.0);
    p.beta_artificial = 0.2468;
    p.beta_necessary = 0.5706;
    p.theta_initial = 0.81930;
    p.theta_target = 0.42818;
    p.theta_decay = 0.90474;
    p.tolerance = tolerance;
    p.max_iter = max_iter;
    p.use_halpern = 1;
    p.warm_steps = 0;
    p.op_steps = 16;
    p.verbose = 0;
    return p;
}

extern "C" pria_options pria_default_options(void) {
    pria_options o;
    o.tier = PRIA_TIER_AUTO;
    o.time_limit = 0.0;
    return o;
}

extern "C" int pria_solve_ex(int64_t n, int64_t t, const double* order_costs, double joint_cost,
                             const double* holding_costs, const pria_params* params,
                             const pria_options* options, int use_float32, const double* warm_x,
                             const double* warm_y, const double* warm_z, double* x, double* y,
                             double* z, pria_info* info) {
    return solve_f64io(n, t, order_costs, joint_cost, holding_costs, params, options, use_float32,
                       warm_x, warm_y, warm_z, x, y, z, info);
}

namespace {

// The shape of a device-IO solve (scale unknown until the device has summed
// the squares); kPriaErrorArgument for inputs the entry points reject.
int device_io_shape(int64_t n, int64_t t, const double* order_costs, const double* holding_costs,
                    const pria_params* params, PriaHostInstance& s) {
    if (!pria_params_valid(params) || n <= 0 || t <= 0 || !order_costs || !holding_costs)
        return kPriaErrorArgument;
    const int64_t T = t + (t & 1);
    const int64_t nd = n * T * (T + 1) / 2;
    if (nd >= (int64_t(1) << 31) || T > kPriaMaxT) return kPriaErrorArgument;
    s.N = int(n);
    s.T = int(T);
    s.Torig = int(t);
    s.ND = int(nd);
    s.cold = true;
    return 0;
}

// Scratch of a device-IO solve besides the images, in doubles: the scale
// partials and total, the objective's row maxima and dots, and x when the
// caller takes none.
struct DeviceIoScratch {
    size_t f_partial = 0, f_total = 0, f_max = 0, f_dot = 0, f_x = 0, doubles = 0;
    DeviceIoScratch(int64_t rows, int64_t t, bool x_given) {
        f_total = kDevScaleBlocks;
        f_max = f_total + 1;
        f_dot = f_max + size_t(rows);
        f_x = f_dot + size_t(rows);
        doubles = f_x + (x_given ? 0 : size_t(rows) * size_t(t));
    }
    size_t bytes() const { return pria_align256(sizeof(double) * doubles); }
};

// Shared body of pria_solve_device and pria_solve_device_noalloc.  `arena`
// null: scratch and images are stream-ordered allocations; else everything
// is carved from `arena` (device memory of at least DeviceIoScratch::bytes()
// + mixed_image_bytes() bytes, 256-byte aligned), and nothing is allocated.
int solve_device_io(PriaHostInstance& s, const double* order_costs, double joint_cost,
                    const double* holding_costs, const pria_params* params,
                    const pria_options* options, int use_float32, double* x, double* y,
                    double* z, pria_info* info, hipStream_t stream, char* arena) {
    const pria_options o = options ? *options : pria_default_options();
    const int64_t n = s.N, t = s.Torig, rows = n * t;
    const DeviceIoScratch sc(rows, t, x != nullptr);
    StreamBuffer scratch_buf;
    if (arena) scratch_buf.borrow(arena);
    else if (scratch_buf.alloc(sizeof(double) * sc.doubles, stream)) return kPriaErrorDevice;
    double* w = static_cast<double*>(scratch_buf.ptr);
    hipLaunchKernelGGL(pria_sumsq_rows, dim3(kDevScaleBlocks), dim3(kDevBlock), 0, stream,
                       holding_costs, rows, int(t), w + sc.f_partial);
    hipLaunchKernelGGL(pria_sumsq_total, dim3(1), dim3(kDevBlock), 0, stream, order_costs,
                       int(n), joint_cost, w + sc.f_partial, kDevScaleBlocks, w + sc.f_total);
    double sq = 0.0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&sq, w + sc.f_total, sizeof(double), hipMemcpyDeviceToHost, stream) !=
            hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return kPriaErrorDevice;
    // shape and scale only: the arrays stay on the device
    s.scale = sq > 0.0 ? std::sqrt(sq) : 1.0;
    s.k0 = joint_cost / s.scale;
    const PriaDeviceInput din{order_costs, holding_costs, int(t)};
    const LogicParams lp = pria_logic_params(*params);
    pria_info local;
    pria_info* out = info ? info : &local;
    double* xd = x ? x : w + sc.f_x;
    const dim3 grid{static_cast<unsigned>(rows)}, blk{unsigned(kDevBlock)};
    auto emit = [&](auto& image) -> int {
        using Real = std::remove_reference_t<decltype(*image.job.a.x)>;
        hipLaunchKernelGGL(pria_scatter_rows<Real>, grid, blk, 0, stream, image.job.a.x,
                           order_costs, s.k0, 0, s.N, int(t), s.T, s.scale, xd);
        if (y)
            hipLaunchKernelGGL(pria_scatter_rows<Real>, grid, blk, 0, stream, image.job.a.y,
                               order_costs, s.k0, 1, s.N, int(t), s.T, s.scale, y);
        if (z)
            hipLaunchKernelGGL(pria_scatter_rows<Real>, grid, blk, 0, stream, image.job.a.z,
                               order_costs, s.k0, 2, s.N, int(t), s.T, s.scale, z);
        return hipGetLastError() == hipSuccess ? 0 : kPriaErrorDevice;
    };
    int status = solve_mixed(s, lp, o, use_float32 != 0, &din, stream, emit, out,
                             arena ? arena + sc.bytes() : nullptr);
    if (status) {
        out->status = status;
        return status;
    }
    out->device_bytes += int64_t(sizeof(double) * sc.doubles);
    // the upper bound of the returned point on the caller's instance
    hipLaunchKernelGGL(pria_objective_rows, grid, blk, 0, stream, holding_costs, xd, int(t),
                       w + sc.f_max, w + sc.f_dot);
    hipLaunchKernelGGL(pria_objective_total, dim3(1), blk, 0, stream, order_costs, joint_cost,
                       int(n), int(t), w + sc.f_max, w + sc.f_dot, w + sc.f_total);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&out->upper, w + sc.f_total, sizeof(double), hipMemcpyDeviceToHost,
                       stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return kPriaErrorDevice;
    out->bracket = out->upper - out->lower;
    return out->status == 0 ? 0 : 1;
}

void clear_info(pria_info* info) {
    if (!info) return;
    std::memset(info, 0, sizeof(*info));
    info->lower = info->upper = info->relative_gap = info->bracket =
        std::numeric_limits<double>::quiet_NaN();
    info->status = kPriaErrorArgument;
}

// ---- pria_solve_device_noalloc's memory ----
//
// One static array per GPU serves every call on that GPU: scratch, the
// FP32 image and the FP64 image of a handover.  Host code reaches the array
// only through the address hipGetSymbolAddress gives for the current device
// (the array's name in host code is a stand-in, not a device address).
constexpr size_t kPriaNoAllocBytes = size_t(PRIA_NOALLOC_BYTES);
constexpr int kNoAllocDevices = 64;

static __device__ __align__(256) char scratch[kPriaNoAllocBytes];

struct NoAllocSlot {
    std::mutex mu;          // held for the whole of a call: calls on one GPU take turns
    char* ptr = nullptr;    // the array's device address on this GPU
};

NoAllocSlot no_alloc_slots[kNoAllocDevices];

// Bytes of the array a pria_solve_device_noalloc call uses (-1: shape rejected).
int64_t noalloc_bytes(int64_t n, int64_t t, int use_float32, int with_x) {
    if (n <= 0 || t <= 0) return -1;
    const int64_t T = t + (t & 1);
    const int64_t nd = n * T * (T + 1) / 2;
    if (nd >= (int64_t(1) << 31) || T > kPriaMaxT) return -1;
    PriaHostInstance s;
    s.N = int(n);
    s.T = int(T);
    s.Torig = int(t);
    s.ND = int(nd);
    const DeviceIoScratch sc(n * t, t, with_x != 0);
    return int64_t(sc.bytes() + mixed_image_bytes(s, use_float32 != 0));
}

}  // namespace

extern "C" int pria_solve_device(int64_t n, int64_t t, const double* order_costs,
                                 double joint_cost, const double* holding_costs,
                                 const pria_params* params, const pria_options* options,
                                 int use_float32, double* x, double* y, double* z,
                                 pria_info* info, void* stream_handle) {
    clear_info(info);
    PriaHostInstance s;
    if (device_io_shape(n, t, order_costs, holding_costs, params, s)) return kPriaErrorArgument;
    return solve_device_io(s, order_costs, joint_cost, holding_costs, params, options,
                           use_float32, x, y, z, info, static_cast<hipStream_t>(stream_handle),
                           nullptr);
}

extern "C" int pria_solve_device_noalloc(int64_t n, int64_t t, const double* order_costs,
                                         double joint_cost, const double* holding_costs,
                                         const pria_params* params, const pria_options* options,
                                         int use_float32, double* x, double* y, double* z,
                                         pria_info* info, void* stream_handle) {
    clear_info(info);
    PriaHostInstance s;
    if (device_io_shape(n, t, order_costs, holding_costs, params, s)) return kPriaErrorArgument;
    const int64_t need = noalloc_bytes(n, t, use_float32, x != nullptr);
    if (need > int64_t(kPriaNoAllocBytes)) {
        record_message("pria_solve_device_noalloc: an instance of n = " + std::to_string(n) +
                       ", t = " + std::to_string(t) + " needs " + std::to_string(need) +
                       " bytes of scratch, more than the " + std::to_string(kPriaNoAllocBytes) +
                       " the entry point holds per GPU");
        return kPriaErrorArgument;
    }
    int dev = 0;
    const hipError_t de = hipGetDevice(&dev);
    if (de != hipSuccess) {
        record_error("hipGetDevice", de, __FILE__, __LINE__);
        return kPriaErrorDevice;
    }
    if (dev < 0 || dev >= kNoAllocDevices) {
        record_message("pria_solve_device_noalloc: device ordinal " + std::to_string(dev) +
                       " beyond the " + std::to_string(kNoAllocDevices) + " it serves");
        return kPriaErrorArgument;
    }
    NoAllocSlot& slot = no_alloc_slots[dev];
    std::lock_guard<std::mutex> lock(slot.mu);
    if (!slot.ptr) {  // first call on this GPU, with it current
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            record_error("hipGetSymbolAddress(scratch)", e, __FILE__, __LINE__);
            return kPriaErrorDevice;
        }
        slot.ptr = static_cast<char*>(p);
    }
    return solve_device_io(s, order_costs, joint_cost, holding_costs, params, options,
                           use_float32, x, y, z, info, static_cast<hipStream_t>(stream_handle),
                           slot.ptr);
}

extern "C" int pria_solve_f64(int64_t n, int64_t t, const double* order_costs, double joint_cost,
                              const double* holding_costs, const pria_params* params,
                              const double* warm_x, const double* warm_y, const double* warm_z,
                              double* x, double* y, double* z, pria_result_info* info) {
    return solve_reference_abi<double>(n, t, order_costs, joint_cost, holding_costs, params, warm_x,
                                       warm_y, warm_z, x, y, z, info);
}

extern "C" int pria_solve_f32(int64_t n, int64_t t, const float* order_costs, float joint_cost,
                              const float* holding_costs, const pria_params* params,
                              const float* warm_x, const float* warm_y, const float* warm_z,
                              float* x, float* y, float* z, pria_result_info* info) {
    return solve_reference_abi<float>(n, t, order_costs, joint_cost, holding_costs, params, warm_x,
                                      warm_y, warm_z, x, y, z, info);
}

extern "C" int pria_solve_batch_f64(int64_t count, int64_t n, int64_t t, const double* order_costs,
                                    const double* joint_costs, const double* holding_costs,
                                    const pria_params* params, double* x, double* y, double* z,
                                    pria_result_info* infos) {
    return solve_batch<double>(count, n, t, order_costs, joint_costs, holding_costs, params, x, y,
                               z, infos);
}

extern "C" int pria_solve_batch_f32(int64_t count, int64_t n, int64_t t, const float* order_costs,
                                    const float* joint_costs, const float* holding_costs,
                                    const pria_params* params, float* x, float* y, float* z,
                                    pria_result_info* infos) {
    return solve_batch<float>(count, n, t, order_costs, joint_costs, holding_costs, params, x, y,
                              z, infos);
}

extern "C" double pria_objective_f64(int64_t n, int64_t t, const double* order_costs,
                                     double joint_cost, const double* holding_costs,
                                     const double* x) {
    return pria_objective_host(n, t, order_costs, joint_cost, holding_costs, x);
}

namespace {

template <class Real>
int debug_step(int64_t n, int64_t t, const double* k, double k0, const double* h,
               const double* x1, const double* y1, const double* z1, double omega, double eta,
               int tier, double* x2, double* y2, double* z2) {
    if (n <= 0 || t <= 0 || (t & 1)) return kPriaErrorArgument;
    PriaHostInstance s;
    if (pria_build_host(n, t, k, k0, h, nullptr, nullptr, nullptr, s)) return kPriaErrorArgument;
    // the inputs are already scaled: undo the build's scaling
    for (auto& v : s.h) v *= s.scale;
    for (auto& v : s.k) v *= s.scale;
    s.k0 *= s.scale;
    s.scale = 1.0;
    PriaLayout L = pria_layout(s.N, s.T, s.ND, sizeof(Real));
    Buffer buf;
    if (device_malloc(&buf.ptr, L.total_bytes) != hipSuccess) return kPriaErrorDevice;
    char* base = static_cast<char*>(buf.ptr);
    std::vector<char> image(L.input_bytes, 0);
    pria_fill_image<Real>(s, L, image.data());
    if (hipMemcpy(base, image.data(), L.off[PriaLayout::kXS], hipMemcpyHostToDevice) != hipSuccess)
        return kPriaErrorDevice;
    PriaArrays<Real> a = pria_bind<Real>(L, base, s);
    std::vector<Real> v(s.ND);
    const double* ins[3] = {x1, y1, z1};
    Real* dst[3] = {a.sx, a.sy, a.sz};
    for (int q = 0; q < 3; ++q) {
        for (int i = 0; i < s.N; ++i)
            for (int sa = 0; sa < s.T; ++sa)
                for (int tb = sa; tb < s.T; ++tb)
                    v[pria_idx(s.N, s.T, i, sa, tb)] = Real(ins[q][(int64_t(i) * t + sa) * t + tb]);
        if (hipMemcpy(dst[q], v.data(), sizeof(Real) * s.ND, hipMemcpyHostToDevice) != hipSuccess)
            return kPriaErrorDevice;
    }
    if (pria_one_step<Real>(a, omega, eta, tier, nullptr)) return kPriaErrorDevice;
    if (hipDeviceSynchronize() != hipSuccess) return kPriaErrorDevice;
    double* outs[3] = {x2, y2, z2};
    for (int q = 0; q < 3; ++q) {
        if (hipMemcpy(v.data(), dst[q], sizeof(Real) * s.ND, hipMemcpyDeviceToHost) != hipSuccess)
            return kPriaErrorDevice;
        std::memset(outs[q], 0, sizeof(double) * size_t(n) * t * t);
        for (int i = 0; i < s.N; ++i)
            for (int sa = 0; sa < s.T; ++sa)
                for (int tb = sa; tb < s.T; ++tb)
                    outs[q][(int64_t(i) * t + sa) * t + tb] = double(v[pria_idx(s.N, s.T, i, sa, tb)]);
    }
    return 0;
}

}  // namespace

extern "C" int pria_debug_step(int64_t n, int64_t t, const double* k, double k0, const double* h,
                               const double* x1, const double* y1, const double* z1, double omega,
                               double eta, int use_float32, int tier, double* x2, double* y2,
                               double* z2) {
    return use_float32 ? debug_step<float>(n, t, k, k0, h, x1, y1, z1, omega, eta, tier, x2, y2, z2)
                       : debug_step<double>(n, t, k, k0, h, x1, y1, z1, omega, eta, tier, x2, y2, z2);
}
