// The scalar side of restarted reflected Halpern PDHG, shared by PRAQ and
// PRIA, by the host loop of the device tier and by thread 0 of the one-
// workgroup kernels.  All FP64.
//
// The control flow is solve_*_warm_start + _halpern_step of the reference
// (qubo/praq-jax/praq/solver.py:397-561, jrp/pria-jax/pria/solver.py:514-724),
// written as a state machine whose array work (steps, anchor copies, check
// sums) is done by the caller between transitions:
//
//   start:       one uncounted vanilla step from the projected start, then a
//                check with the bounds at (x, y)           -> after_warm_check
//   WARM block:  op_steps counted vanilla steps, check at (x, y) -> after_warm_check
//   HALPERN block: [anchor copy if pending] op_steps reflected steps (the last
//                one storing y, sy), check at (sx, sy)     -> after_halpern_check
//   RESTART:     one uncounted vanilla step with the new omega, anchor copy,
//                norm sums                                  -> after_restart
//   DONE_X / DONE_SX: the answer is (x, y) / (sx, sy).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace reform {

enum Action : int32_t {
    kActWarmBlock = 0,
    kActHalpernBlock = 1,
    kActRestart = 2,
    kActDoneX = 3,
    kActDoneSX = 4,
};

// What one check pass reduces, in instance-independent form.
struct CheckSums {
    double pnorm;    // primal part of the canonical norm of (x - sx)
    double dnorm;    // dual part, |y - sy|^2 (+ |z - sz|^2)
    double cross;    // <A (x - sx), (y - sy)>-type cross term
    double pdrift;   // primal drift from the anchor, squared
    double ddrift;   // dual drift from the anchor, squared
    double upper;    // upper bound at the bound point, original scale
    double lower;    // lower bound at the bound point, original scale
};

struct LogicParams {
    double omega, eta, beta_sufficient, beta_necessary, beta_artificial;
    double theta_initial, theta_target, theta_decay, tolerance;
    double cross_sign;           // +2 PRAQ, -2 PRIA
    int64_t max_iter;            // < 0: none
    int64_t op_steps;
    int64_t warm_max;            // op_steps * warm_steps (Halpern) or max_iter
    int32_t use_halpern;
    int32_t fallback_nonpositive;  // PRAQ: omega fallback also on <= 0
    int32_t verbose;               // device tier: one stderr line per check
    int32_t stall_checks;          // > 0: stop as stalled when the best gap has
                                   // not halved within this many checks (used
                                   // by FP32 iterates to hand over to FP64)
    double switch_gap;             // > 0: stop as stalled once the gap is at or
                                   // below it (FP32 iterates: hand the last
                                   // stretch to FP64)
};

// The handover policy of FP32 iterates: the stall test (8 checks without a
// halving of the best gap) and the gap, 3e-5, below which the rest of the
// solve runs in FP64 -- FP32 iterates certify gaps down to about 1e-5 and
// then crawl, so a 1e-6 solve finishes its last stretch in FP64, while looser
// tolerances finish in FP32.  On the 64-instance mixture at 1e-6 this gives
// the FP64 solve's speed with objectives within 5e-7 of it (FP32 throughout:
// 1.6x slower, 2e-6).  Overridable by REFORM_FP32_STALL_CHECKS and
// REFORM_FP32_SWITCH_GAP (0 disables the gap test).
inline void apply_fp32_policy(LogicParams& p) {
    static const int stall = [] {
        const char* v = std::getenv("REFORM_FP32_STALL_CHECKS");
        return v ? std::atoi(v) : 8;
    }();
    static const double gap = [] {
        const char* v = std::getenv("REFORM_FP32_SWITCH_GAP");
        return v ? std::atof(v) : 3e-5;
    }();
    p.stall_checks = stall;
    p.switch_gap = gap;
}

struct LogicState {
    double omega, theta, fp_error, check_error;
    double upper, lower, gap, new_err;
    int64_t restart_iter, iteration, restarts, checks;
    int32_t action;
    int32_t anchor_pending;      // copy (x, y) to the anchor before the next block
    int32_t status;              // 0 converged, 1 limit, 2 time limit, 3 stalled
    int32_t stalled;             // set by the stall test; the next action is kept
    double best_gap;             // smallest gap by halvings, for the stall test
    int64_t best_check;          // check count when best_gap was last halved
    int64_t switch_iteration;    // iteration at which FP32 handed over (-1: never)
    // kept by the batched tier (no other tier reads them)
    int32_t in64;                // iterates currently in the FP64 image
    int32_t last_action;         // the last action executed
    int64_t steps32, steps64;    // steps taken in each precision
    // the FP64 image is not there yet (a serving request admitted with its
    // FP32 image only): a handover waits, stalled, until the host has
    // allocated the image and cleared this
    int32_t hold64;
    int32_t reserved;
};

// One line per check of the device tier's host loop, when verbose.
inline void logic_trace(const LogicState& st, const char* what) {
    fprintf(stderr,
            "[%s] it=%lld r=%lld gap=%.6e up=%.12e lo=%.12e new_err=%.12e fp=%.12e check=%.12e "
            "omega=%.12e action=%d\n",
            what, (long long)st.iteration, (long long)st.restart_iter, st.gap, st.upper, st.lower,
            st.new_err, st.fp_error, st.check_error, st.omega, st.action);
}

// Progress of a long host-loop solve: one stderr line per REFORM_PROGRESS
// seconds (unset or 0: none).
struct LogicProgress {
    double every = 0.0, next = 0.0;
    LogicProgress() {
        const char* v = std::getenv("REFORM_PROGRESS");
        every = v ? std::atof(v) : 0.0;
        next = every;
    }
    void maybe(const LogicState& st, const char* family, double elapsed) {
        if (every <= 0.0 || elapsed < next) return;
        next = elapsed + every;
        fprintf(stderr, "[%s] %.1f s: iteration %lld, restarts %lld, gap %.3e, upper %.9e\n",
                family, elapsed, (long long)st.iteration, (long long)st.restarts, st.gap,
                st.upper);
    }
};

__host__ __device__ inline double canonical_norm(const CheckSums& s, double eta, double omega,
                                                 double cross_sign) {
    double v = (omega / eta) * s.pnorm + (1.0 / (eta * omega)) * s.dnorm + cross_sign * s.cross;
    return sqrt(v > 0.0 ? v : 0.0);
}

__host__ __device__ inline double relative_gap(double upper, double lower) {
    return (upper - lower) / (1.0 + fabs(upper) + fabs(lower));
}

__host__ __device__ inline bool within_max_iter(int64_t iteration, int64_t max_iter) {
    return iteration < max_iter || max_iter < 0;
}

__host__ __device__ inline void logic_init(LogicState& st, const LogicParams& p) {
    st.omega = p.omega;
    st.theta = p.theta_initial;
    st.fp_error = 0.0;
    st.check_error = INFINITY;
    st.upper = st.lower = 0.0;
    st.gap = INFINITY;
    st.new_err = 0.0;
    st.restart_iter = 0;
    st.iteration = 0;
    st.restarts = 0;
    st.checks = 0;
    st.action = kActWarmBlock;
    st.anchor_pending = 0;
    st.status = 1;
    st.stalled = 0;
    st.best_gap = INFINITY;
    st.best_check = 0;
    st.switch_iteration = -1;
    st.hold64 = 0;
    st.reserved = 0;
}

__host__ __device__ inline void finish(LogicState& st, const LogicParams& p, int32_t action) {
    st.action = action;
    st.status = st.gap <= p.tolerance ? 0 : 1;
}

// FP32 iterates meeting a tolerance below the switch gap do not finish: they
// hand over, so the check that ends the solve is taken on FP64 iterates (whose
// primal is feasible to FP64 rounding, which makes upper a bound of the LP).
// Looser tolerances finish in FP32.
__host__ __device__ inline bool certify_in_fp64(const LogicState& st, const LogicParams& p) {
    return p.switch_gap > 0.0 && p.tolerance < p.switch_gap && st.gap <= p.tolerance &&
           within_max_iter(st.iteration, p.max_iter);
}

// After the start step or a WARM block: bounds were taken at (x, y), drift
// is from the projected start (the anchor arrays still hold it).
__host__ __device__ inline void logic_after_warm_check(LogicState& st, const LogicParams& p,
                                                       const CheckSums& s) {
    st.upper = s.upper;
    st.lower = s.lower;
    st.gap = relative_gap(s.upper, s.lower);
    if (within_max_iter(st.iteration, p.warm_max) && !(st.gap <= p.tolerance)) {
        st.action = kActWarmBlock;
        return;
    }
    if (!p.use_halpern) {
        finish(st, p, kActDoneX);
        return;
    }
    // omega from the warm phase's drift; NaN (no warm steps: 0/0) falls back
    double pn = sqrt(s.pdrift), dn = sqrt(s.ddrift);
    double om = dn / pn;
    bool bad = (om != om) || (p.fallback_nonpositive && om <= 0.0);  // om != om: NaN
    st.omega = bad ? p.omega : om;
    st.theta = p.theta_initial;
    st.fp_error = canonical_norm(s, p.eta, st.omega, p.cross_sign);
    st.check_error = INFINITY;
    st.restart_iter = 0;
    st.anchor_pending = 1;
    bool terminated = !(within_max_iter(st.iteration, p.max_iter) && !(st.gap <= p.tolerance));
    if (terminated && certify_in_fp64(st, p)) {
        st.stalled = 1;
        st.status = 3;
    } else if (terminated) {
        finish(st, p, kActDoneX);
        return;
    }
    st.action = kActHalpernBlock;
}

// After a HALPERN block: norms at (x - sx), drift (x - x0), bounds at (sx, sy).
__host__ __device__ inline void logic_after_halpern_check(LogicState& st, const LogicParams& p,
                                                          const CheckSums& s) {
    st.checks += 1;
    double new_err = canonical_norm(s, p.eta, st.omega, p.cross_sign);
    st.new_err = new_err;
    st.upper = s.upper;
    st.lower = s.lower;
    st.gap = relative_gap(s.upper, s.lower);
    bool terminated = !(within_max_iter(st.iteration, p.max_iter) && !(st.gap <= p.tolerance));

    bool sufficient = new_err <= p.beta_sufficient * st.fp_error;
    bool necessary = (new_err <= p.beta_necessary * st.fp_error) && (new_err > st.check_error);
    bool artificial = double(st.restart_iter) >= p.beta_artificial * double(st.iteration);
    bool restart = sufficient || necessary || artificial;

    if (terminated && !certify_in_fp64(st, p)) {
        finish(st, p, kActDoneSX);
        return;
    }
    // (a certify_in_fp64 case has gap <= tolerance < switch_gap: it stalls here)
    if (p.switch_gap > 0.0 && st.gap <= p.switch_gap) {
        st.stalled = 1;
        st.status = 3;
    }
    if (p.stall_checks > 0) {
        if (st.gap < 0.5 * st.best_gap) {
            st.best_gap = st.gap;
            st.best_check = st.checks;
        } else if (st.checks - st.best_check >= p.stall_checks) {
            st.stalled = 1;
            st.status = 3;
        }
    }
    if (restart) {
        double pn = sqrt(s.pdrift), dn = sqrt(s.ddrift);
        if (pn > 1e-16 && dn > 1e-16) {
            double om = exp(st.theta * (log(dn) - log(pn)) + (1.0 - st.theta) * log(st.omega));
            st.theta = p.theta_target + p.theta_decay * (st.theta - p.theta_target);
            st.omega = om;
        }
        st.action = kActRestart;
        return;
    }
    st.check_error = new_err;
    st.action = kActHalpernBlock;
}

// After a RESTART: one uncounted vanilla step from (sx, sy), anchors copied,
// norm sums of the new (x - sx).
__host__ __device__ inline void logic_after_restart(LogicState& st, const LogicParams& p,
                                                    const CheckSums& s) {
    st.fp_error = canonical_norm(s, p.eta, st.omega, p.cross_sign);
    st.check_error = INFINITY;
    st.restart_iter = 0;
    st.restarts += 1;
    st.action = kActHalpernBlock;
}

}  // namespace reform
