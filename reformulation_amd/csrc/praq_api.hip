// The C ABI of include/praq_hip.h.  Entry points and validation follow
// qubo/praq-cuda/src/praq.cu:8-79; the extensions add options, the dual
// output and the tier choice.
#include "praq_hip.h"
#include "praq_host.hpp"
#include "reform_tune.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

using namespace reform;

namespace {

constexpr double kDefaultWgLimitSeconds = 600.0;

// Below this many nodes + edges the one-workgroup tier runs the plain solve.
int64_t wg_tier_max() {
    static int64_t v = [] {
        const char* env = std::getenv("REFORM_PRAQ_WG_MAX");
        return env ? std::atoll(env) : int64_t(1) << 17;
    }();
    return v;
}

double wall_clock_rate_hz() {
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0.0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess) return 0.0;
    return double(khz) * 1000.0;
}

struct DeviceBuffer {
    void* ptr = nullptr;
    ~DeviceBuffer() {
        if (ptr) (void)hipFree(ptr);
    }
    int alloc(size_t bytes) { return device_malloc(&ptr, bytes ? bytes : 1) == hipSuccess ? 0 : -1; }
};


// The answer into caller-owned device memory: x widened to FP64, the dual
// from slot-major halves to [m][2] pairs in instance units (as outputs()).
template <class Real>
__global__ void __launch_bounds__(256) praq_put_outputs(const Real* x, const Real* y, int n, int m,
                                                        double scale, double* dx, double* dy) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) dx[g] = double(x[g]);
    if (dy && g < m) {
        dy[2 * size_t(g)] = double(y[g]) * scale;
        dy[2 * size_t(g) + 1] = double(y[m + g]) * scale;
    }
}

int pick_tier(const PraqHostInstance& s, int tier) {
    if (tier != PRAQ_TIER_AUTO) return tier;
    return int64_t(s.n) + int64_t(s.m) <= wg_tier_max() ? PRAQ_TIER_WORKGROUP : PRAQ_TIER_DEVICE;
}

// One instance's device image in one precision, and the tier that runs it.
template <class Real>
struct PraqImage {
    DeviceBuffer buf;
    PraqLayout L;
    PraqJob<Real> job;
    size_t job_off = 0, state_off = 0, work_off = 0, bytes = 0;

    // `dev` set: the input part written on the device from its build
    // (praq_solve_device); else the host image of `s` uploaded
    int init(const PraqHostInstance& s, const LogicParams& lp, hipStream_t stream,
             const PraqDeviceBuild* dev = nullptr) {
        L = praq_layout(s, sizeof(Real));
        job_off = align256(L.total_bytes);
        state_off = align256(job_off + sizeof(PraqJob<Real>));
        work_off = align256(state_off + sizeof(LogicState));
        bytes = work_off + praq_device_work_bytes(s.n, s.m);
        if (buf.alloc(bytes)) return kErrorDevice;
        char* base = static_cast<char*>(buf.ptr);
        if (dev) {
            if (praq_fill_image_device<Real>(*dev, L, base, stream)) return kErrorDevice;
        } else {
            std::vector<char> image(L.input_bytes, 0);
            praq_fill_image<Real>(s, L, image.data());
            REFORM_HIP_CHECK(hipMemcpyAsync(base, image.data(), L.input_bytes,
                                            hipMemcpyHostToDevice, stream));
        }
        REFORM_HIP_CHECK(hipStreamSynchronize(stream));
        std::memset(&job, 0, sizeof(job));
        job.a = praq_bind<Real>(L, base, s);
        job.p = lp;
        job.scale = s.scale;
        job.lanes = s.lanes;
        job.state = reinterpret_cast<LogicState*>(base + state_off);
        return 0;
    }

    int run(int tier, LogicState& st, double time_limit, bool resume, hipStream_t stream) {
        char* base = static_cast<char*>(buf.ptr);
        if (tier == PRAQ_TIER_WORKGROUP) {
            double limit = time_limit > 0.0 ? std::min(time_limit, kDefaultWgLimitSeconds)
                                            : kDefaultWgLimitSeconds;
            job.time_limit_ticks = (long long)(limit * wall_clock_rate_hz());
            job.resume = resume ? 1 : 0;
            if (resume)
                REFORM_HIP_CHECK(hipMemcpyAsync(job.state, &st, sizeof(st), hipMemcpyHostToDevice,
                                                stream));
            REFORM_HIP_CHECK(hipMemcpyAsync(base + job_off, &job, sizeof(job),
                                            hipMemcpyHostToDevice, stream));
            if (praq_launch_wg<Real>(reinterpret_cast<PraqJob<Real>*>(base + job_off), 1, 1024,
                                     stream))
                return kErrorDevice;
            REFORM_HIP_CHECK(hipMemcpyAsync(&st, job.state, sizeof(st), hipMemcpyDeviceToHost,
                                            stream));
            REFORM_HIP_CHECK(hipStreamSynchronize(stream));
            return 0;
        }
        return praq_device_solve<Real>(job, st, reinterpret_cast<double*>(base + work_off),
                                       time_limit, stream, resume);
    }

    int outputs(const PraqHostInstance& s, double* x, double* y, hipStream_t stream) {
        std::vector<Real> hx(s.n), hy(2 * size_t(s.m));
        REFORM_HIP_CHECK(hipMemcpyAsync(hx.data(), job.a.x, sizeof(Real) * s.n,
                                        hipMemcpyDeviceToHost, stream));
        if (y && s.m > 0)
            REFORM_HIP_CHECK(hipMemcpyAsync(hy.data(), job.a.y, sizeof(Real) * 2 * size_t(s.m),
                                            hipMemcpyDeviceToHost, stream));
        REFORM_HIP_CHECK(hipStreamSynchronize(stream));
        for (int i = 0; i < s.n; ++i) x[i] = double(hx[i]);
        if (y)
            for (int e = 0; e < s.m; ++e) {
                y[2 * size_t(e)] = double(hy[e]) * s.scale;
                y[2 * size_t(e) + 1] = double(hy[s.m + e]) * s.scale;
            }
        return 0;
    }

    // the same answer written to device memory, in stream order
    int outputs_device(const PraqHostInstance& s, double* x, double* y, hipStream_t stream) {
        const int count = std::max(s.n, s.m);
        if (count > 0)
            hipLaunchKernelGGL(praq_put_outputs<Real>, dim3((count + 255) / 256), dim3(256), 0,
                               stream, job.a.x, job.a.y, s.n, s.m, s.scale, x,
                               s.m > 0 ? y : nullptr);
        REFORM_HIP_CHECK(hipGetLastError());
        REFORM_HIP_CHECK(hipStreamSynchronize(stream));
        return 0;
    }
};

// FP64 iterates, or FP32 iterates handing over to FP64 when they stall; all
// work on `stream`.  `emit(image)` writes the answer of the image the solve
// finished in.
template <class Emit>
int solve_mixed(const PraqHostInstance& s, LogicParams lp, const PraqOptions& o, bool fp32,
                hipStream_t stream, Emit&& emit, PraqInfo* info,
                const PraqDeviceBuild* dev = nullptr) {
    const int tier = pick_tier(s, o.tier);
    LogicState st;
    std::memset(&st, 0, sizeof(st));
    size_t held = 0;
    if (!fp32) {
        PraqImage<double> d;
        if (d.init(s, lp, stream, dev) || d.run(tier, st, o.time_limit, false, stream) ||
            emit(d))
            return kErrorDevice;
        held = d.bytes;
    } else {
        LogicParams lp32 = lp;
        apply_fp32_policy(lp32);
        PraqImage<float> f;
        if (f.init(s, lp32, stream, dev) || f.run(tier, st, o.time_limit, false, stream))
            return kErrorDevice;
        held = f.bytes;
        if (st.stalled) {
            PraqImage<double> d;
            st.switch_iteration = st.iteration;
            if (d.init(s, lp, stream, dev) || praq_convert_state(f.job.a, d.job.a, stream) ||
                d.run(tier, st, o.time_limit, true, stream) || emit(d))
                return kErrorDevice;
            held += d.bytes;
        } else if (emit(f)) {
            return kErrorDevice;
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
        info->scale = s.scale;
        info->status = st.status;
        info->tier = tier;
        info->switch_iteration = st.switch_iteration;
        info->device_bytes = int64_t(held);
        info->bracket = st.upper - st.lower;
    }
    return 0;
}

}  // namespace

extern "C" PraqParameters praq_default_parameters(void) {
    PraqParameters p;
    p.tolerance = 1e-8;
    p.max_iter = -1;
    p.use_float32 = 0;
    p.omega = 1.0;
    p.beta_sufficient = 0.36787944117144233;  // exp(-1)
    p.beta_necessary = 0.5706;
    p.beta_artificial = 0.2468;
    p.theta_initial = 0.81930;
    p.theta_target = 0.42818;
    p.theta_decay = 0.90474;
    p.op_steps = 16;
    p.interval_offset = 0.0;
    p.interval_scale = 4.328085122666891;  // 3 / ln 2
    p.interval_power = 0.0;
    return p;
}

extern "C" PraqOptions praq_default_options(void) {
    PraqOptions o;
    o.use_halpern = 1;
    o.warm_steps = 0;
    o.dense_degree = 0;
    o.tier = PRAQ_TIER_AUTO;
    o.eta = 0.999;
    o.time_limit = 0.0;
    return o;
}

extern "C" int64_t praq_check_interval(int64_t n, const PraqParameters* parameters) {
    if (!parameters || n <= 0) return 2;
    double u = std::log(double(n) / 50000.0);
    double power = parameters->interval_power;
    double b = power == 0.0 ? u : std::expm1(power * u) / power;
    double v = std::floor(parameters->interval_offset + parameters->interval_scale * b + 0.5);
    if (!(v >= 2.0)) return 2;
    if (v > 1e9) return 1000000000;
    return int64_t(v);
}

namespace {

struct StreamOwner {
    hipStream_t s = nullptr;
    ~StreamOwner() {
        if (s) (void)hipStreamDestroy(s);
    }
};

// The body of praq_solve_ex and praq_solve_device: `instance` in host
// memory, every launch and copy on `stream` (own_stream: on a stream of its
// own, created once the instance has been validated), the answer through
// `emit(image, host instance, stream)`.
template <class Emit>
int solve_entry(const PraqInstance* instance, const PraqParameters* parameters,
                const PraqOptions* options, const double* warm_x, const double* warm_y,
                hipStream_t stream, bool own_stream, Emit&& emit, PraqInfo* info) {
    PraqParameters p = *parameters;
    if (!praq_params_valid(p)) return kErrorArgument;
    PraqOptions o = options ? *options : praq_default_options();
    if (o.tier < PRAQ_TIER_AUTO || o.tier > PRAQ_TIER_DEVICE) return kErrorArgument;

    PraqHostInstance s;
    int status = praq_build_host(*instance, o.dense_degree, warm_x, warm_y, s);
    if (status) {
        if (info) info->status = status;
        return status;
    }
    if (p.max_iter <= 0) p.max_iter = -1;
    if (p.op_steps <= 0) p.op_steps = praq_check_interval(instance->n, &p);
    if (p.op_steps < 2) p.op_steps = 2;
    LogicParams lp = praq_logic_params(p, o);
    if (pick_tier(s, o.tier) == PRAQ_TIER_DEVICE) praq_set_lanes(s, praq_device_lanes(s.n, s.m));

    PraqInfo local;
    PraqInfo* out = info ? info : &local;
    StreamOwner owned;
    if (own_stream) {
        const hipError_t e = hipStreamCreateWithFlags(&owned.s, hipStreamNonBlocking);
        if (e != hipSuccess) {
            owned.s = nullptr;
            record_error("hipStreamCreateWithFlags", e, __FILE__, __LINE__);
            out->status = kErrorDevice;
            return kErrorDevice;
        }
        stream = owned.s;
    }
    status = solve_mixed(s, lp, o, p.use_float32 != 0, stream,
                         [&](auto& image) { return emit(image, s, stream); }, out);
    if (status < 0) {
        out->status = status;
        return status;
    }
    return out->status == 0 ? 0 : 1;
}

void clear_info(PraqInfo* info) {
    if (!info) return;
    std::memset(info, 0, sizeof(*info));
    info->lower = info->upper = info->relative_gap = info->bracket =
        std::numeric_limits<double>::quiet_NaN();
    info->status = kErrorArgument;
}

}  // namespace

extern "C" int praq_solve_ex(const PraqInstance* instance, const PraqParameters* parameters,
                             const PraqOptions* options, const double* warm_x,
                             const double* warm_y, double* x, double* y, PraqInfo* info) {
    clear_info(info);
    if (!instance || !parameters || !x) return kErrorArgument;
    return solve_entry(instance, parameters, options, warm_x, warm_y, nullptr, true,
                       [&](auto& image, const PraqHostInstance& s, hipStream_t stream) {
                           return image.outputs(s, x, y, stream);
                       },
                       info);
}

extern "C" int praq_solve_device(const PraqInstance* instance, const PraqParameters* parameters,
                                 const PraqOptions* options, double* x, double* y,
                                 PraqInfo* info, void* stream_handle) {
    clear_info(info);
    if (!instance || !parameters || !x || !praq_params_valid(*parameters)) return kErrorArgument;
    if (options && (options->tier < PRAQ_TIER_AUTO || options->tier > PRAQ_TIER_DEVICE))
        return kErrorArgument;
    const int64_t n = instance->n, m = instance->m;
    if (n <= 0 || m < 0 || !instance->diag || (m > 0 && (!instance->rows || !instance->cols ||
                                                         !instance->vals)))
        return kErrorInstance;
    if (n > std::numeric_limits<int>::max() || m > std::numeric_limits<int>::max() / 2)
        return kErrorInstance;
    hipStream_t stream = static_cast<hipStream_t>(stream_handle);
    PraqParameters p = *parameters;
    const PraqOptions o = options ? *options : praq_default_options();
    // lanes per node for the tier the size picks, as praq_solve_ex's build
    PraqHostInstance shape;
    shape.n = int(n);
    shape.m = int(m);
    const int lanes = pick_tier(shape, o.tier) == PRAQ_TIER_DEVICE
                          ? praq_device_lanes(int(n), int(m))
                          : praq_lanes_for(int(n), int(m));
    // the instance built on the device: no copy of the COO arrays
    PraqDeviceBuild build;
    int status = praq_build_device(*instance, o.dense_degree, lanes, stream, build, shape);
    if (status) {
        if (info) info->status = status;
        return status;
    }
    if (p.max_iter <= 0) p.max_iter = -1;
    if (p.op_steps <= 0) p.op_steps = praq_check_interval(n, &p);
    if (p.op_steps < 2) p.op_steps = 2;
    const LogicParams lp = praq_logic_params(p, o);
    PraqInfo local;
    PraqInfo* out = info ? info : &local;
    status = solve_mixed(shape, lp, o, p.use_float32 != 0, stream,
                         [&](auto& image) { return image.outputs_device(shape, x, y, stream); },
                         out, &build);
    if (status < 0) {
        out->status = status;
        return status;
    }
    return out->status == 0 ? 0 : 1;
}

extern "C" int praq_solve(const PraqInstance* instance, const PraqParameters* parameters,
                          double* x, PraqResult* result) {
    if (result) {
        result->lower = std::numeric_limits<double>::quiet_NaN();
        result->upper = std::numeric_limits<double>::quiet_NaN();
        result->iterations = 0;
        result->status = kErrorArgument;
    }
    if (!instance || !parameters || !x || !result) return kErrorArgument;
    PraqInfo info;
    int status = praq_solve_ex(instance, parameters, nullptr, nullptr, nullptr, x, nullptr, &info);
    if (status < 0) {
        result->status = status;
        return status;
    }
    result->lower = info.lower;
    result->upper = info.upper;
    result->iterations = info.iterations;
    result->status = status;
    return status;
}

extern "C" int reform_time_praq_steps(const PraqInstance* instance, int use_float32, int lanes,
                                      int warmup, int steps, double* ms) {
    if (!instance || !ms || steps <= 0 || warmup < 0 || lanes < 0 || lanes > 64 ||
        (lanes & (lanes - 1)) != 0)
        return kErrorArgument;
    PraqHostInstance s;
    if (int status = praq_build_host(*instance, 0, nullptr, nullptr, s)) return status;
    praq_set_lanes(s, lanes > 0 ? lanes : praq_device_lanes(s.n, s.m));
    PraqParameters p = praq_default_parameters();
    p.op_steps = 16;
    const LogicParams lp = praq_logic_params(p, praq_default_options());
    hipStream_t stream;
    REFORM_HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    int st;
    if (use_float32) {
        PraqImage<float> f;
        st = f.init(s, lp, stream) ? kErrorDevice
                                   : praq_time_steps<float>(f.job, warmup, steps, stream, ms);
    } else {
        PraqImage<double> d;
        st = d.init(s, lp, stream) ? kErrorDevice
                                   : praq_time_steps<double>(d.job, warmup, steps, stream, ms);
    }
    (void)hipStreamDestroy(stream);
    return st;
}
