// PRAQ host side: instance validation and the COO -> half-edge CSR build,
// the packed per-instance device image, and the job record the kernels read.
//
// Reference: Structure::build qubo/praq-cuda/src/solver.cuh:517-569 (CSR over
// half-edges, degree floored at 1, objective scale), HalfEdge.centroid
// qubo/praq-jax/praq/half_edge.py:87-94 (cold start), the warm-start
// projection qubo/praq-jax/praq/solver.py:524-527.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "halpern.hpp"
#include "praq_device.hpp"
#include "praq_hip.h"

namespace reform {

// Parameters every PRAQ entry rejects with an argument error (-1): a
// non-positive omega or restart factor, a NaN or negative tolerance, a
// non-finite schedule.
inline bool praq_params_valid(const PraqParameters& p) {
    return p.tolerance >= 0.0 && p.omega > 0.0 && p.beta_sufficient > 0.0 &&
           p.beta_necessary > 0.0 && p.beta_artificial >= 0.0 && std::isfinite(p.theta_initial) &&
           std::isfinite(p.theta_target) && std::isfinite(p.theta_decay) &&
           std::isfinite(p.interval_offset) && std::isfinite(p.interval_scale) &&
           p.interval_power >= 0.0;
}


constexpr int kErrorArgument = -1;
constexpr int kErrorInstance = -2;
constexpr int kErrorDevice = -3;

template <class Real>
struct PraqJob {
    using real_type = Real;
    PraqArrays<Real> a;
    LogicParams p;
    double scale;
    int lanes;
    int resume;                  // continue from *state (after a precision switch)
    long long time_limit_ticks;  // wall-clock ticks; <= 0 none
    LogicState* state;           // device address the final state goes to
};

struct PraqHostInstance {
    int n = 0, m = 0, lanes = 1;
    double scale = 1.0;
    std::vector<int> e0, e1, ptr, half;
    std::vector<double> w, diag, deg, x_start, y_start;  // y_start slot-major, scaled
    // hubs (degree > heavy_degree) and their chunks, for the device tier
    int heavy_degree = 0;
    std::vector<int> heavy, hchunk, chunk_node;
};

// Hubs: nodes whose CSR list is longer than 64 half-edges per lane of the
// instance's lane groups; a lane group would walk them alone.
inline int praq_heavy_degree(int lanes) { return 64 * lanes; }

// The tuning table's lanes for a mean degree, 0 if unset (tune.hip).
int praq_tuned_lanes(double mean_degree);

// Lanes per node in the scatter: the smallest power of two leaving each
// lane at most 8 half-edges at the mean degree, capped at a wavefront.  The
// batched and one-workgroup tiers use this rule (measured best there).
inline int praq_lanes_for(int n, int m) {
    double average = 2.0 * double(m) / double(n > 0 ? n : 1);
    int lanes = 1;
    while (lanes < 64 && double(lanes) * 8.0 < average) lanes *= 2;
    return lanes;
}

// Lanes of a single-instance device-tier solve: the tuning table's choice
// for the mean degree (it was measured on that tier), else the rule.
inline int praq_device_lanes(int n, int m) {
    const int tuned = praq_tuned_lanes(2.0 * double(m) / double(n > 0 ? n : 1));
    return tuned > 0 ? tuned : praq_lanes_for(n, m);
}

// Lanes per node and the hubs and chunks they imply.
inline void praq_set_lanes(PraqHostInstance& s, int lanes) {
    s.lanes = lanes;
    s.heavy_degree = praq_heavy_degree(lanes);
    s.heavy.clear();
    s.chunk_node.clear();
    s.hchunk.assign(1, 0);
    for (int i = 0; i < s.n; ++i) {
        const int len = s.ptr[i + 1] - s.ptr[i];
        if (len <= s.heavy_degree) continue;
        const int h = int(s.heavy.size());
        s.heavy.push_back(i);
        for (int c = 0; c < (len + kPraqChunk - 1) / kPraqChunk; ++c) s.chunk_node.push_back(h);
        s.hchunk.push_back(int(s.chunk_node.size()));
    }
}

inline double clamp_host(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

inline void project_edge_host(double* v, double w, bool covering) {
    if (covering) {
        double mu = clamp_host(0.5 * (v[0] + v[1]), -w, 0.0);
        v[0] = v[1] = mu;
    } else {
        double line = v[0] + 0.5 * (w - v[0] - v[1]);
        v[0] = clamp_host(line, 0.0, w);
        v[1] = w - v[0];
    }
}

// The objective scale's squares, ||(diag; 2 W)||^2, in a fixed blocked
// order the host build, the device build (praq_build.hip) and the oracle all
// follow, so they give the same bits: chunks of kPraqScaleChunk terms summed
// in order, the chunk sums in order, the diagonal's d d and the weights'
// (4 v) v as two such sums, added.
constexpr int kPraqScaleChunk = 4096;

inline double praq_sumsq_blocked(const double* v, int64_t len, bool weights) {
    double total = 0.0;
    for (int64_t c0 = 0; c0 < len; c0 += kPraqScaleChunk) {
        const int64_t c1 = c0 + kPraqScaleChunk < len ? c0 + kPraqScaleChunk : len;
        double part = 0.0;
        for (int64_t k = c0; k < c1; ++k) part += weights ? 4.0 * v[k] * v[k] : v[k] * v[k];
        total += part;
    }
    return total;
}

// Validates and builds; returns 0 or kErrorInstance.
inline int praq_build_host(const PraqInstance& in, int dense_degree, const double* warm_x,
                           const double* warm_y, PraqHostInstance& s) {
    const int64_t limit = std::numeric_limits<int>::max();
    if (in.n <= 0 || in.n > limit || in.m < 0 || in.m > limit / 2) return kErrorInstance;
    if (!in.diag || (in.m > 0 && (!in.rows || !in.cols || !in.vals))) return kErrorInstance;
    const int n = int(in.n), m = int(in.m);
    s.n = n;
    s.m = m;
    s.e0.resize(m);
    s.e1.resize(m);
    s.w.resize(m);
    for (int e = 0; e < m; ++e) {
        int64_t i = in.rows[e], j = in.cols[e];
        if (i < 0 || i >= in.n || j < 0 || j >= in.n || i == j) return kErrorInstance;
        s.e0[e] = int(i);
        s.e1[e] = int(j);
    }
    const double squares =
        praq_sumsq_blocked(in.diag, n, false) + praq_sumsq_blocked(in.vals, m, true);
    s.scale = squares > 0.0 ? std::sqrt(squares) : 1.0;
    s.diag.resize(n);
    for (int i = 0; i < n; ++i) s.diag[i] = in.diag[i] / s.scale;
    for (int e = 0; e < m; ++e) s.w[e] = 2.0 * in.vals[e] / s.scale;  // signed

    s.ptr.assign(size_t(n) + 1, 0);
    for (int e = 0; e < m; ++e) {
        ++s.ptr[s.e0[e] + 1];
        ++s.ptr[s.e1[e] + 1];
    }
    for (int i = 0; i < n; ++i) s.ptr[i + 1] += s.ptr[i];
    s.half.resize(2 * size_t(m));
    std::vector<int> fill(s.ptr.begin(), s.ptr.end() - 1);
    for (int e = 0; e < m; ++e) {
        s.half[fill[s.e0[e]]++] = e;        // slot 0 of edge e
        s.half[fill[s.e1[e]]++] = m + e;    // slot 1
    }
    s.deg.resize(n);
    for (int i = 0; i < n; ++i) {
        double incident = dense_degree ? double(n - 1) : double(s.ptr[i + 1] - s.ptr[i]);
        s.deg[i] = incident > 1.0 ? incident : 1.0;
    }
    praq_set_lanes(s, praq_lanes_for(n, m));

    s.x_start.resize(n);
    for (int i = 0; i < n; ++i) s.x_start[i] = clamp_host(warm_x ? warm_x[i] : 0.5, 0.0, 1.0);
    s.y_start.resize(2 * size_t(m));
    for (int e = 0; e < m; ++e) {
        double w = std::fabs(s.w[e]);
        bool covering = s.w[e] > 0.0;
        double v[2];
        if (warm_y) {
            v[0] = warm_y[2 * size_t(e)] / s.scale;
            v[1] = warm_y[2 * size_t(e) + 1] / s.scale;
        } else {
            v[0] = v[1] = -0.5 * s.w[e];  // centroid: -W_ij on both slots
        }
        project_edge_host(v, w, covering);
        s.y_start[e] = v[0];
        s.y_start[m + e] = v[1];
    }
    return 0;
}

// Byte offsets of one instance's arrays in its device image.  The input part
// (uploaded) comes first, the iteration state after it.
struct PraqLayout {
    enum { kE0, kE1, kPtr, kHalf, kHeavy, kHChunk, kChunkNode, kW, kDiag, kDeg, kXS, kYS, kX0, kY0,
           kX, kSX, kRefl, kYC, kY, kSY, kChunkSum, kCount };
    size_t off[kCount];
    size_t input_bytes = 0, total_bytes = 0;
};

inline size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

inline PraqLayout praq_layout(const PraqHostInstance& s, size_t real_size) {
    PraqLayout L;
    const size_t ni = sizeof(int), nn = size_t(s.n), mm = size_t(s.m);
    const size_t nh = s.heavy.size(), nc = s.chunk_node.size();
    size_t sizes[PraqLayout::kCount];
    sizes[PraqLayout::kE0] = ni * mm;
    sizes[PraqLayout::kE1] = ni * mm;
    sizes[PraqLayout::kPtr] = ni * (nn + 1);
    sizes[PraqLayout::kHalf] = ni * 2 * mm;
    sizes[PraqLayout::kHeavy] = ni * nh;
    sizes[PraqLayout::kHChunk] = ni * (nh + 1);
    sizes[PraqLayout::kChunkNode] = ni * nc;
    sizes[PraqLayout::kChunkSum] = sizeof(double) * nc;
    sizes[PraqLayout::kW] = real_size * mm;
    sizes[PraqLayout::kDiag] = real_size * nn;
    sizes[PraqLayout::kDeg] = real_size * nn;
    sizes[PraqLayout::kXS] = real_size * nn;
    sizes[PraqLayout::kYS] = real_size * 2 * mm;
    sizes[PraqLayout::kX0] = real_size * nn;
    sizes[PraqLayout::kY0] = real_size * 2 * mm;
    sizes[PraqLayout::kX] = real_size * nn;
    sizes[PraqLayout::kSX] = real_size * nn;
    sizes[PraqLayout::kRefl] = real_size * nn;
    sizes[PraqLayout::kYC] = real_size * 2 * mm;
    sizes[PraqLayout::kY] = real_size * 2 * mm;
    sizes[PraqLayout::kSY] = real_size * 2 * mm;
    size_t at = 0;
    for (int k = 0; k < PraqLayout::kCount; ++k) {
        L.off[k] = at;
        at += align256(sizes[k] ? sizes[k] : 1);
        if (k == PraqLayout::kYS) L.input_bytes = at;
    }
    L.total_bytes = at;
    return L;
}

template <class Real>
inline void praq_fill_image(const PraqHostInstance& s, const PraqLayout& L, char* image) {
    auto put_int = [&](int k, const std::vector<int>& v) {
        if (!v.empty()) std::memcpy(image + L.off[k], v.data(), sizeof(int) * v.size());
    };
    auto put_real = [&](int k, const std::vector<double>& v) {
        Real* dst = reinterpret_cast<Real*>(image + L.off[k]);
        for (size_t q = 0; q < v.size(); ++q) dst[q] = Real(v[q]);
    };
    put_int(PraqLayout::kE0, s.e0);
    put_int(PraqLayout::kE1, s.e1);
    put_int(PraqLayout::kPtr, s.ptr);
    put_int(PraqLayout::kHalf, s.half);
    put_int(PraqLayout::kHeavy, s.heavy);
    put_int(PraqLayout::kHChunk, s.hchunk);
    put_int(PraqLayout::kChunkNode, s.chunk_node);
    put_real(PraqLayout::kW, s.w);
    put_real(PraqLayout::kDiag, s.diag);
    put_real(PraqLayout::kDeg, s.deg);
    put_real(PraqLayout::kXS, s.x_start);
    put_real(PraqLayout::kYS, s.y_start);
}

// An instance built on the device from a COO instance in device memory
// (praq_build.hip): the arrays of praq_fill_image's image, in FP64 where
// the image holds reals, and their counts; freed in stream order.
struct PraqDeviceBuild {
    hipStream_t stream = nullptr;
    void* arena = nullptr;
    void* hub_arena = nullptr;
    int n = 0, m = 0, nh = 0, nc = 0;
    int *e0 = nullptr, *e1 = nullptr, *ptr = nullptr, *half = nullptr;
    int *heavy = nullptr, *hchunk = nullptr, *chunk_node = nullptr;
    double *w = nullptr, *diag = nullptr, *deg = nullptr, *xs = nullptr, *ys = nullptr;
    PraqDeviceBuild() = default;
    PraqDeviceBuild(const PraqDeviceBuild&) = delete;
    PraqDeviceBuild& operator=(const PraqDeviceBuild&) = delete;
    ~PraqDeviceBuild();
};

// The cold-start build of `in` (device pointers) with `lanes` per node:
// validation (kErrorInstance as praq_build_host), then the arrays on the
// device; `shape` receives the counts, scale and lanes the layout and the
// binding read (its arrays stay empty).  Work on stream s.
int praq_build_device(const PraqInstance& in, int dense_degree, int lanes, hipStream_t s,
                      PraqDeviceBuild& b, PraqHostInstance& shape);

// praq_fill_image from a device build, on the device (stream s).
template <class Real>
int praq_fill_image_device(const PraqDeviceBuild& b, const PraqLayout& L, char* image,
                           hipStream_t s);

// Binds with the hub path off (heavy_degree = INT_MAX): every tier but the
// device tier walks every node with its lane group.
template <class Real>
inline PraqArrays<Real> praq_bind(const PraqLayout& L, char* base, const PraqHostInstance& s) {
    PraqArrays<Real> a;
    a.n = s.n;
    a.m = s.m;
    a.nh = int(s.heavy.size());
    a.nc = int(s.chunk_node.size());
    a.heavy_degree = std::numeric_limits<int>::max();
    a.heavy = reinterpret_cast<const int*>(base + L.off[PraqLayout::kHeavy]);
    a.hchunk = reinterpret_cast<const int*>(base + L.off[PraqLayout::kHChunk]);
    a.chunk_node = reinterpret_cast<const int*>(base + L.off[PraqLayout::kChunkNode]);
    a.chunk_sum = reinterpret_cast<double*>(base + L.off[PraqLayout::kChunkSum]);
    a.e0 = reinterpret_cast<const int*>(base + L.off[PraqLayout::kE0]);
    a.e1 = reinterpret_cast<const int*>(base + L.off[PraqLayout::kE1]);
    a.ptr = reinterpret_cast<const int*>(base + L.off[PraqLayout::kPtr]);
    a.half = reinterpret_cast<const int*>(base + L.off[PraqLayout::kHalf]);
    a.w = reinterpret_cast<const Real*>(base + L.off[PraqLayout::kW]);
    a.diag = reinterpret_cast<const Real*>(base + L.off[PraqLayout::kDiag]);
    a.deg = reinterpret_cast<const Real*>(base + L.off[PraqLayout::kDeg]);
    a.xs = reinterpret_cast<const Real*>(base + L.off[PraqLayout::kXS]);
    a.ys = reinterpret_cast<const Real*>(base + L.off[PraqLayout::kYS]);
    a.x0 = reinterpret_cast<Real*>(base + L.off[PraqLayout::kX0]);
    a.y0 = reinterpret_cast<Real*>(base + L.off[PraqLayout::kY0]);
    a.x = reinterpret_cast<Real*>(base + L.off[PraqLayout::kX]);
    a.sx = reinterpret_cast<Real*>(base + L.off[PraqLayout::kSX]);
    a.refl = reinterpret_cast<Real*>(base + L.off[PraqLayout::kRefl]);
    a.yc = reinterpret_cast<Real*>(base + L.off[PraqLayout::kYC]);
    a.y = reinterpret_cast<Real*>(base + L.off[PraqLayout::kY]);
    a.sy = reinterpret_cast<Real*>(base + L.off[PraqLayout::kSY]);
    return a;
}

inline LogicParams praq_logic_params(const PraqParameters& p, const PraqOptions& o) {
    LogicParams q;
    q.omega = p.omega;
    q.eta = o.eta > 0.0 ? o.eta : 0.999;
    q.beta_sufficient = p.beta_sufficient;
    q.beta_necessary = p.beta_necessary;
    q.beta_artificial = p.beta_artificial;
    q.theta_initial = p.theta_initial;
    q.theta_target = p.theta_target;
    q.theta_decay = p.theta_decay;
    q.tolerance = p.tolerance;
    q.cross_sign = 2.0;
    q.max_iter = p.max_iter <= 0 ? -1 : p.max_iter;
    q.op_steps = p.op_steps;
    q.use_halpern = o.use_halpern ? 1 : 0;
    q.warm_max = q.use_halpern ? q.op_steps * (o.warm_steps > 0 ? o.warm_steps : 0) : q.max_iter;
    q.fallback_nonpositive = 1;
    q.verbose = 0;
    q.stall_checks = 0;
    q.switch_gap = 0.0;
    return q;
}

// Implemented in praq.hip.
template <class Real>
int praq_device_solve(const PraqJob<Real>& job, LogicState& st, double* work,
                      double time_limit_seconds, hipStream_t stream, bool resume = false);
// FP32 state -> FP64 state of the same instance (precision switch)
int praq_convert_state(const PraqArrays<float>& from, const PraqArrays<double>& to,
                       hipStream_t stream);
int praq_convert_jobs(const PraqJob<float>* from, const PraqJob<double>* to, int count,
                      hipStream_t stream);
template <class Real>
int praq_launch_wg(const PraqJob<Real>* jobs, int count, int block, hipStream_t stream);
// Device-tier work area of an instance with n nodes and m edges.
size_t praq_device_work_bytes(int64_t n, int64_t m);
// Mean milliseconds of one reflected device-tier step after `warmup` steps
// (HIP events over `steps` steps), from the job's start point.
template <class Real>
int praq_time_steps(const PraqJob<Real>& job, int warmup, int steps, hipStream_t stream,
                    double* ms);

}  // namespace reform
