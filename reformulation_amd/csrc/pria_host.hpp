// PRIA host side: padding, scaling, the slice-order image of one instance,
// the projected start, the job record, and the dense <-> slice conversions.
//
// Reference: solve_jrp_problem_warm_start jrp/pria-jax/pria/solver.py:608-724
// (padding odd T, scale ||(H, K, K0)||, warm duals scaled by K_i / K0 and
// projected, outputs normalised and unpadded); the scale falls back to 1 for
// an all-zero instance as jrp/pria-cuda does.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <vector>

#include "halpern.hpp"
#include "pria_device.hpp"
#include "pria_hip.h"

namespace reform {

// Parameters every PRIA entry rejects with an argument error: a missing
// record, a non-positive omega or eta, a NaN or negative tolerance.
inline bool pria_params_valid(const pria_params* p) {
    return p && p->omega > 0.0 && p->eta > 0.0 && p->tolerance >= 0.0;
}


constexpr int kPriaErrorArgument = -1;
constexpr int kPriaErrorDevice = -3;

// Largest padded period count T': the x tile of 16 FP64 columns (T' x 17
// doubles) fits the 160 KB LDS.  The one-workgroup tier, whose x tile shares
// the LDS with the workgroup's own arrays, takes T' up to 640.
constexpr int kPriaMaxT = 1200;
constexpr int kPriaWgMaxT = 640;

template <class Real>
struct PriaJob {
    using real_type = Real;
    PriaArrays<Real> a;
    LogicParams p;
    double scale;
    int resume;                  // continue from *state (after a precision switch)
    int pad;
    long long time_limit_ticks;
    LogicState* state;
};

struct PriaHostInstance {
    int N = 0, T = 0, Torig = 0, ND = 0;
    double scale = 1.0, k0 = 0.0;
    std::vector<double> k;                 // scaled
    std::vector<double> h, x0, y0, z0;     // slice order, scaled / projected
    bool cold = true;                      // start computed on the device
};

inline int64_t pria_idx(int N, int T, int i, int s, int t) {
    return int64_t(pria_off(N, T, s)) + int64_t(i) * (T - s) + (t - s);
}

// Euclidean projection of v onto {u >= 0, sum u = target} (sort, prefix sums,
// count-based rho); target <= 0 gives 0.
inline void simplex_project_host(std::vector<double>& v, double target) {
    if (v.empty()) return;
    if (!(target > 0.0)) {
        std::fill(v.begin(), v.end(), 0.0);
        return;
    }
    std::vector<double> u(v);
    std::sort(u.begin(), u.end(), std::greater<double>());
    double running = 0.0, at = 0.0;
    int64_t rho = 0;
    for (size_t q = 0; q < u.size(); ++q) {
        running += u[q];
        if (u[q] * double(q + 1) > running - target) {
            ++rho;
            at = running;
        }
    }
    double theta = (at - target) / double(rho);
    for (double& e : v) e = std::max(e - theta, 0.0);
}

// Builds the slice-order image.  warm_* are dense [n][t][t] in the caller's
// (unpadded) shape, duals normalised; NULL means zeros (the cold start).
inline int pria_build_host(int64_t n, int64_t t, const double* k, double k0, const double* h,
                           const double* warm_x, const double* warm_y, const double* warm_z,
                           PriaHostInstance& s) {
    if (n <= 0 || t <= 0 || !k || !h) return kPriaErrorArgument;
    const int64_t T = t + (t & 1);
    const int64_t nd = n * T * (T + 1) / 2;
    if (nd >= (int64_t(1) << 31) || T > kPriaMaxT) return kPriaErrorArgument;
    s.N = int(n);
    s.T = int(T);
    s.Torig = int(t);
    s.ND = int(nd);
    double sq = k0 * k0;
    for (int64_t i = 0; i < n; ++i) sq += k[i] * k[i];
    for (int64_t i = 0; i < n; ++i)
        for (int64_t a = 0; a < t; ++a)
            for (int64_t b = a; b < t; ++b) {
                double v = h[(i * t + a) * t + b];
                sq += v * v;
            }
    s.scale = sq > 0.0 ? std::sqrt(sq) : 1.0;
    s.k.resize(n);
    for (int64_t i = 0; i < n; ++i) s.k[i] = k[i] / s.scale;
    s.k0 = k0 / s.scale;
    s.h.assign(nd, 0.0);
    const int N = s.N, TT = s.T;
    for (int i = 0; i < N; ++i)
        for (int a = 0; a < t; ++a)
            for (int b = a; b < t; ++b)
                s.h[pria_idx(N, TT, i, a, b)] = h[(int64_t(i) * t + a) * t + b] / s.scale;
    s.cold = !warm_x && !warm_y && !warm_z;
    if (s.cold) return 0;  // the centre of each simplex, filled on the device
    s.x0.assign(nd, 0.0);
    s.y0.assign(nd, 0.0);
    s.z0.assign(nd, 0.0);
    for (int i = 0; i < N; ++i)
        for (int a = 0; a < t; ++a)
            for (int b = a; b < t; ++b) {
                int64_t g = pria_idx(N, TT, i, a, b), go = (int64_t(i) * t + a) * t + b;
                if (warm_x) s.x0[g] = warm_x[go];
                if (warm_y) s.y0[g] = warm_y[go] * s.k[i];
                if (warm_z) s.z0[g] = warm_z[go] * s.k0;
            }
    // project the start: x per column (i, t), y per row (i, s), z per slice s
    std::vector<double> seg;
    for (int i = 0; i < N; ++i)
        for (int b = 0; b < TT; ++b) {
            seg.resize(b + 1);
            for (int a = 0; a <= b; ++a) seg[a] = s.x0[pria_idx(N, TT, i, a, b)];
            simplex_project_host(seg, 1.0);
            for (int a = 0; a <= b; ++a) s.x0[pria_idx(N, TT, i, a, b)] = seg[a];
        }
    for (int a = 0; a < TT; ++a) {
        const int len = TT - a;
        for (int i = 0; i < N; ++i) {
            int64_t base = pria_idx(N, TT, i, a, a);
            seg.assign(s.y0.begin() + base, s.y0.begin() + base + len);
            simplex_project_host(seg, s.k[i]);
            std::copy(seg.begin(), seg.end(), s.y0.begin() + base);
        }
        int64_t base = pria_off(N, TT, a);
        seg.assign(s.z0.begin() + base, s.z0.begin() + base + int64_t(N) * len);
        simplex_project_host(seg, s.k0);
        std::copy(seg.begin(), seg.end(), s.z0.begin() + base);
    }
    return 0;
}

struct PriaLayout {
    enum {
        kK, kH, kXS, kYS, kZS, kX0, kY0, kZ0, kXC, kX, kSX, kYC, kY, kSY, kZC, kZ, kSZ, kR,
        kThX, kThY, kThZ,  // FP64 thresholds: N T', N T', T'
        kCount
    };
    size_t off[kCount];
    size_t input_bytes = 0, total_bytes = 0;
};

inline size_t pria_align256(size_t v) { return (v + 255) & ~size_t(255); }

inline PriaLayout pria_layout(int N, int T, int ND, size_t real_size) {
    PriaLayout L;
    size_t at = 0;
    for (int q = 0; q < PriaLayout::kCount; ++q) {
        L.off[q] = at;
        size_t bytes = real_size * size_t(ND);
        if (q == PriaLayout::kK) bytes = real_size * size_t(N);
        if (q == PriaLayout::kThX || q == PriaLayout::kThY) bytes = sizeof(double) * size_t(N) * T;
        if (q == PriaLayout::kThZ) bytes = sizeof(double) * size_t(T);
        at += pria_align256(bytes ? bytes : 1);
        if (q == PriaLayout::kZS) L.input_bytes = at;
    }
    L.total_bytes = at;
    return L;
}

template <class Real>
inline void pria_fill_image(const PriaHostInstance& s, const PriaLayout& L, char* image) {
    auto put = [&](int q, const std::vector<double>& v) {
        Real* dst = reinterpret_cast<Real*>(image + L.off[q]);
        for (size_t e = 0; e < v.size(); ++e) dst[e] = Real(v[e]);
    };
    put(PriaLayout::kK, s.k);
    put(PriaLayout::kH, s.h);
    if (!s.cold) {
        put(PriaLayout::kXS, s.x0);
        put(PriaLayout::kYS, s.y0);
        put(PriaLayout::kZS, s.z0);
    }
}

// Bytes of the image that must be uploaded: K and h only for a cold start.
inline size_t pria_upload_bytes(const PriaHostInstance& s, const PriaLayout& L) {
    return s.cold ? L.off[PriaLayout::kXS] : L.input_bytes;
}

template <class Real>
inline PriaArrays<Real> pria_bind(const PriaLayout& L, char* base, const PriaHostInstance& s) {
    PriaArrays<Real> a;
    a.N = s.N;
    a.T = s.T;
    a.ND = s.ND;
    a.k0 = Real(s.k0);
    auto at = [&](int q) { return reinterpret_cast<Real*>(base + L.off[q]); };
    a.k = at(PriaLayout::kK);
    a.h = at(PriaLayout::kH);
    a.xs = at(PriaLayout::kXS);
    a.ys = at(PriaLayout::kYS);
    a.zs = at(PriaLayout::kZS);
    a.x0 = at(PriaLayout::kX0);
    a.y0 = at(PriaLayout::kY0);
    a.z0 = at(PriaLayout::kZ0);
    a.xc = at(PriaLayout::kXC);
    a.x = at(PriaLayout::kX);
    a.sx = at(PriaLayout::kSX);
    a.yc = at(PriaLayout::kYC);
    a.y = at(PriaLayout::kY);
    a.sy = at(PriaLayout::kSY);
    a.zc = at(PriaLayout::kZC);
    a.z = at(PriaLayout::kZ);
    a.sz = at(PriaLayout::kSZ);
    a.r = at(PriaLayout::kR);
    a.thx = reinterpret_cast<double*>(base + L.off[PriaLayout::kThX]);
    a.thy = reinterpret_cast<double*>(base + L.off[PriaLayout::kThY]);
    a.thz = reinterpret_cast<double*>(base + L.off[PriaLayout::kThZ]);
    return a;
}

inline LogicParams pria_logic_params(const pria_params& p) {
    LogicParams q;
    q.omega = p.omega;
    q.eta = p.eta;
    q.beta_sufficient = p.beta_sufficient;
    q.beta_necessary = p.beta_necessary;
    q.beta_artificial = p.beta_artificial;
    q.theta_initial = p.theta_initial;
    q.theta_target = p.theta_target;
    q.theta_decay = p.theta_decay;
    q.tolerance = p.tolerance;
    q.cross_sign = -2.0;
    q.max_iter = p.max_iter < 0 ? -1 : p.max_iter;
    q.op_steps = p.op_steps < 2 ? 2 : p.op_steps;
    q.use_halpern = p.use_halpern ? 1 : 0;
    int64_t warm = p.warm_steps > 0 ? p.warm_steps : 0;
    q.warm_max = q.use_halpern ? q.op_steps * warm : q.max_iter;
    q.fallback_nonpositive = 0;
    q.verbose = p.verbose;
    q.stall_checks = 0;
    q.switch_gap = 0.0;
    return q;
}

// Slice-order device values -> dense [n][t][t] caller shape (unpadded);
// norm_k: divide by K_i (y) / K0 (z), 0 where that budget is 0.
template <class Real>
inline void pria_unpack(const PriaHostInstance& s, const std::vector<Real>& v, int which,
                        double* out) {
    const int N = s.N, T = s.T, t = s.Torig;
    std::memset(out, 0, sizeof(double) * size_t(N) * t * t);
    for (int i = 0; i < N; ++i)
        for (int a = 0; a < t; ++a)
            for (int b = a; b < t; ++b) {
                double val = double(v[pria_idx(N, T, i, a, b)]);
                if (which == 1) val = s.k[i] > 0.0 ? val / s.k[i] : 0.0;
                if (which == 2) val = s.k0 > 0.0 ? val / s.k0 : 0.0;
                out[(int64_t(i) * t + a) * t + b] = val;
            }
}

// primal_objective (jrp/pria-jax/pria/jrp_problem.py:143-161) of dense x.
inline double pria_objective_host(int64_t n, int64_t t, const double* k, double k0,
                                  const double* h, const double* x) {
    double obj = 0.0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t a = 0; a < t; ++a) {
            double m = -INFINITY;
            for (int64_t b = 0; b < t; ++b) m = std::max(m, x[(i * t + a) * t + b]);
            obj += k[i] * m;
        }
    for (int64_t a = 0; a < t; ++a) {
        double m = -INFINITY;
        for (int64_t i = 0; i < n; ++i)
            for (int64_t b = 0; b < t; ++b) m = std::max(m, x[(i * t + a) * t + b]);
        obj += k0 * m;
    }
    for (int64_t i = 0; i < n; ++i)
        for (int64_t a = 0; a < t; ++a)
            for (int64_t b = a; b < t; ++b) obj += h[(i * t + a) * t + b] * x[(i * t + a) * t + b];
    return obj;
}

// Implemented in pria.hip.
template <class Real>
int pria_device_solve(const PriaJob<Real>& job, LogicState& st, double* work,
                      double time_limit_seconds, hipStream_t stream, bool resume = false);
// FP32 state -> FP64 state of the same instance (precision switch)
int pria_convert_state(const PriaArrays<float>& from, const PriaArrays<double>& to,
                       hipStream_t stream);
int pria_convert_jobs(const PriaJob<float>* from, const PriaJob<double>* to, int count,
                      hipStream_t stream);
template <class Real>
int pria_launch_wg(const PriaJob<Real>* jobs, int count, int max_T, hipStream_t stream);
template <class Real>
int pria_cold_start(const PriaArrays<Real>& a, hipStream_t stream);
template <class Real>
int pria_one_step(const PriaArrays<Real>& a, double omega, double eta, int tier, hipStream_t stream);
size_t pria_device_work_bytes(int N, int T);

}  // namespace reform
