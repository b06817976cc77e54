// The batched device tier (wide.hpp): block tables over all instances of a
// batch, per-instance round control and precision, one round captured in a
// hipGraph.
#include "wide.hpp"
#include "tune.hpp"

#include <hip/hip_ext.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>
#include <type_traits>
#include <vector>

namespace reform {

namespace {

constexpr int kWideBlock = 256;
constexpr int kRoundsPerPoll = 2;
// tables of fewer instances are not compacted further
constexpr int kShrinkMin = 8;

int grid_for(int threads) { return (threads + kWideBlock - 1) / kWideBlock; }

template <class Job>
using RealOf = typename std::decay<Job>::type::real_type;

// Run `f` on the instance's job record in its current precision.
template <class J32, class J64, class F>
__device__ __forceinline__ void dispatch(const J32* j32, const J64* j64, int inst, F&& f) {
    if (j64[inst].state->in64) f(j64[inst]);
    else f(j32[inst]);
}

__device__ __forceinline__ bool wide_runs_steps(int act) {
    return act == kActWarmBlock || act == kActHalpernBlock || act == kActInit;
}

__device__ __forceinline__ bool wide_restarting(const LogicState* st) {
    return st->action == kActRestart && !st->stalled;
}

// The step this block's instance takes in launch j of this round, or false.
// mode 0: the round's steps -- init and warm rounds vanilla, Halpern rounds
// reflected, each instance for its own check interval; mode 1: the one
// vanilla step of an instance restarting.
template <class Step, class Job, class MakeStep>
__device__ __forceinline__ bool wide_step_for(const Job& job, const int* round_action, int inst,
                                              int mode, int j, MakeStep make, Step* out) {
    const LogicState* st = job.state;
    if (mode == 0) {
        const int act = round_action[inst];
        if (!wide_runs_steps(act) || j >= (act == kActInit ? 1 : job.p.op_steps)) return false;
        const bool reflect = act == kActHalpernBlock;
        *out = make(st->restart_iter + j, st->omega, job.p.eta, reflect,
                    !reflect || j + 1 == job.p.op_steps);
        return true;
    }
    if (!wide_restarting(st)) return false;
    *out = make(0, st->omega, job.p.eta, false, true);
    return true;
}

// The copy a copy launch makes for this instance: mode 0 (a round's start)
// 3 = the start copies of an instance starting, 0 = anchors where pending;
// mode 1: 1 = restart anchors; mode 2: 2 = finished step points; -1: none.
__device__ __forceinline__ int wide_copy_kind(const LogicState* st, const int* round_action,
                                              const int* pending, int inst, int mode) {
    if (mode == 0) {
        if (round_action[inst] == kActInit) return 3;
        return pending[inst] ? 0 : -1;
    }
    if (mode == 1) return wide_restarting(st) ? 1 : -1;
    return st->action == kActDoneSX ? 2 : -1;
}

// Whether a copy launch of this mode can have work for the instance, from
// per-instance arrays alone, before its job record and state are read:
// mode 0 a start or a pending anchor copy; mode 1 the instance is in the
// restart step's work lists (list_at of the latest list build); mode 2
// (finished step points) is decided by the state.
__device__ __forceinline__ bool copy_may_apply(int mode, const int* round_action,
                                               const int* pending, const int* list_at, int inst) {
    if (mode == 0) return round_action[inst] == kActInit || pending[inst];
    if (mode == 1) return list_at[3 * inst] >= 0;
    return true;
}

// The action a logic launch applies to this instance, or -1.
template <class Job>
__device__ __forceinline__ int wide_logic_action(const Job& job, const int* round_action, int inst,
                                                 int mode) {
    if (mode == 0) {
        const int act = round_action[inst];
        return wide_runs_steps(act) ? act : -1;
    }
    return wide_restarting(job.state) ? int(kActRestart) : -1;
}

// The scalar logic after a check, on one thread; the same sequence as the
// device tier's host loop.  `steps` counts the steps this instance took in
// its current precision.
template <class Job>
__device__ void wide_apply(const Job& job, int act, const CheckSums& cs) {
    LogicState st = *job.state;
    const int64_t taken = act == kActWarmBlock || act == kActHalpernBlock ? job.p.op_steps : 1;
    if (st.in64) st.steps64 += taken;
    else st.steps32 += taken;
    if (act == kActInit) {
        logic_after_warm_check(st, job.p, cs);
    } else if (act == kActWarmBlock) {
        st.restart_iter += job.p.op_steps;
        st.iteration += job.p.op_steps;
        logic_after_warm_check(st, job.p, cs);
    } else if (act == kActHalpernBlock) {
        st.restart_iter += job.p.op_steps;
        st.iteration += job.p.op_steps;
        logic_after_halpern_check(st, job.p, cs);
    } else {
        logic_after_restart(st, job.p, cs);
    }
    st.last_action = act;
    *job.state = st;
}

// One thread per instance: the round's action and pending anchor copies.
template <class Job>
__global__ void wide_pre(const Job* jobs, int count, int* round_action, int* pending) {
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= count) return;
    LogicState* st = jobs[inst].state;
    int act = st->action;
    if (act == kActNew) {
        st->action = kActWarmBlock;  // what logic_init leaves; the Init round's check decides
        act = kActInit;
    } else if (st->stalled || (act != kActWarmBlock && act != kActHalpernBlock)) {
        act = kActSkip;
    }
    round_action[inst] = act;
    const int pend = act == kActHalpernBlock && st->anchor_pending;
    pending[inst] = pend;
    if (pend) st->anchor_pending = 0;
}

// One thread per instance: a cold start at its next round, in the first
// precision.
template <class Job>
__global__ void wide_init(const Job* jobs, int count, int first_in64) {
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= count) return;
    LogicState s;
    logic_init(s, jobs[inst].p);
    s.action = kActNew;
    s.in64 = first_in64;
    s.last_action = kActInit;
    s.steps32 = s.steps64 = 0;
    s.hold64 = 0;
    *jobs[inst].state = s;
}

// An FP32 instance whose logic asked for FP64 and whose FP64 image is there:
// its state is converted, then it continues in FP64.
__device__ __forceinline__ bool wide_hands_over(const LogicState* st) {
    return !st->in64 && st->stalled && !st->hold64;
}

// One thread per instance: FP32 instances whose logic asked for FP64 (their
// state was converted by the launch before) continue in FP64 from here.
template <class Job>
__global__ void wide_promote(const Job* jobs, int count) {
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= count) return;
    LogicState* st = jobs[inst].state;
    if (wide_hands_over(st)) {
        st->in64 = 1;
        st->stalled = 0;
        st->switch_iteration = st->iteration;
    }
}

// One thread per instance: out of time (every instance, or those with a
// nonzero mask entry).  The answer is the last step's input after a warm
// block or a restart, else the step point; as the device tier.
template <class Job>
__global__ void wide_expire(const Job* jobs, int count, const int* mask) {
    const int inst = blockIdx.x * blockDim.x + threadIdx.x;
    if (inst >= count) return;
    if (mask && !mask[inst]) return;
    LogicState* st = jobs[inst].state;
    if (st->action == kActDoneX || st->action == kActDoneSX) return;
    if (st->action == kActNew) {  // never started: its start point is the answer
        st->action = kActDoneX;
        st->status = 2;
        return;
    }
    st->action = st->last_action == kActHalpernBlock ? kActDoneSX : kActDoneX;
    st->status = 2;
}

// One thread per table position: 1 while the instance has not finished.
template <class Job>
__global__ void wide_live(const Job* jobs, int count, int* live) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= count) return;
    const LogicState* st = jobs[q].state;
    live[q] = st->action == kActDoneX ? 0 : (st->in64 ? 2 : 1);
}

// One thread per subset position: the job record of full-table position keep[q].
template <class Job>
__global__ void wide_gather_jobs(const Job* src, const int* keep, int n, Job* dst) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) dst[q] = src[keep[q]];
}

// One block: finished step points have been copied (DONE_SX -> DONE_X);
// count the instances still active.
template <class Job>
__global__ void wide_count(const Job* jobs, int count, int* active) {
    __shared__ double red[16];
    double n = 0.0;
    for (int inst = threadIdx.x; inst < count; inst += blockDim.x) {
        LogicState* st = jobs[inst].state;
        if (st->action == kActDoneSX) st->action = kActDoneX;
        if (st->action != kActDoneX) n += 1.0;
    }
    n = block_sum1(n, red);
    if (threadIdx.x == 0) *active = int(n);
}

// Step j's scalars (wide_step_for's, from a RoundStep the list build wrote),
// or false when the instance takes no step j.
template <class Step, class MakeStep>
__device__ __forceinline__ bool round_step(const RoundStep& rs, int j, MakeStep make, Step* out) {
    if (j >= rs.steps) return false;
    if (rs.act == kActRestart) {
        *out = make(0, rs.omega, rs.eta, false, true);
        return true;
    }
    const bool reflect = rs.act == kActHalpernBlock;
    *out = make(rs.r0 + j, rs.omega, rs.eta, reflect, !reflect || j + 1 == rs.steps);
    return true;
}

// The RoundStep of an instance that takes steps in a round (mode 0: action
// `act`) or a restart step (mode 1).
template <class Job>
__device__ __forceinline__ RoundStep round_step_of(const Job& job, const LogicState* st, int act,
                                                  int mode) {
    RoundStep rs;
    rs.r0 = mode == 0 ? st->restart_iter : 0;
    rs.omega = st->omega;
    rs.eta = job.p.eta;
    rs.act = mode == 0 ? act : int(kActRestart);
    rs.steps = mode == 0 && act != kActInit ? int(job.p.op_steps) : 1;
    return rs;
}

// ---------------------------------------------------------------- PRAQ ----

using PJ32 = PraqJob<float>;
using PJ64 = PraqJob<double>;

// The step kernels run over the round's work lists (wpraq_lists): entries
// (instance, local block, precision) of the instances taking steps; block b
// takes entry b, blocks past the list's length exit.
template <class F>
__device__ __forceinline__ void praq_listed(const PJ32* j32, const PJ64* j64, const int4& t,
                                            F&& f) {
    if (t.z) f(j64[t.x]);
    else f(j32[t.x]);
}

__global__ void __launch_bounds__(256) wpraq_node(const PJ32* j32, const PJ64* j64,
                                                  const int4* list, const int* lcount,
                                                  const RoundStep* rstep, int j) {
    const int n = *lcount;
    const int4 t = list[blockIdx.x];   // the grid is the list's capacity
    if (int(blockIdx.x) >= n) return;
    praq_listed(j32, j64, t, [&](const auto& job) {
        using Real = RealOf<decltype(job)>;
        PraqStep<Real> sc;
        if (!round_step(rstep[t.x], j, praq_step_scalars<Real>, &sc)) return;
        const int lanes = job.lanes;
        const int gid = t.y * kWideBlock + threadIdx.x;
        praq_node(praq_view(job.a), sc, gid / lanes, gid % lanes, lanes);
    });
}

__global__ void __launch_bounds__(256) wpraq_edge(const PJ32* j32, const PJ64* j64,
                                                  const int4* list, const int* lcount,
                                                  const RoundStep* rstep, int j) {
    const int n = *lcount;
    const int4 t = list[blockIdx.x];
    if (int(blockIdx.x) >= n) return;
    praq_listed(j32, j64, t, [&](const auto& job) {
        using Real = RealOf<decltype(job)>;
        PraqStep<Real> sc;
        if (!round_step(rstep[t.x], j, praq_step_scalars<Real>, &sc)) return;
        const int e = t.y * kWideBlock + threadIdx.x;
        if (e < job.a.m) praq_edge(praq_view(job.a), sc, e);
    });
}

constexpr int kPraqListBlock = 1024;

// One block: the instances taking steps (mode 0) or a restart step (mode
// 1), in table order: per instance its RoundStep and its node and edge
// blocks' offsets in the two lists (at[3q] = -1: no steps), with its
// precision at[3q + 2]; the lengths go to lcount[0..1] (node, edge).
__global__ void __launch_bounds__(kPraqListBlock)
wpraq_lists(const PJ64* j64, int count, const int* first, const int* round_action, int mode,
            RoundStep* rstep, int* at_out, int* lcount) {
    __shared__ int scan[2][kPraqListBlock];
    const int tid = threadIdx.x;
    const int chunk = (count + kPraqListBlock - 1) / kPraqListBlock;
    const int q0 = min(count, tid * chunk), q1 = min(count, q0 + chunk);
    int mine[2] = {0, 0};
    for (int q = q0; q < q1; ++q) {
        const LogicState* st = j64[q].state;
        const int act = mode == 0 ? round_action[q] : int(kActRestart);
        const bool runs = mode == 0 ? wide_runs_steps(act) : wide_restarting(st);
        if (!runs) continue;
        mine[0] += first[4 * q + 1];
        mine[1] += first[4 * q + 3];
        rstep[q] = round_step_of(j64[q], st, act, mode);
    }
    scan[0][tid] = mine[0];
    scan[1][tid] = mine[1];
    __syncthreads();
    for (int off = 1; off < kPraqListBlock; off <<= 1) {
        const int v0 = tid >= off ? scan[0][tid - off] : 0;
        const int v1 = tid >= off ? scan[1][tid - off] : 0;
        __syncthreads();
        scan[0][tid] += v0;
        scan[1][tid] += v1;
        __syncthreads();
    }
    int at0 = scan[0][tid] - mine[0], at1 = scan[1][tid] - mine[1];
    for (int q = q0; q < q1; ++q) {
        const LogicState* st = j64[q].state;
        const bool runs = mode == 0 ? wide_runs_steps(round_action[q]) : wide_restarting(st);
        if (!runs) {
            at_out[3 * q] = -1;
            continue;
        }
        at_out[3 * q] = at0;
        at_out[3 * q + 1] = at1;
        at_out[3 * q + 2] = st->in64 ? 1 : 0;
        at0 += first[4 * q + 1];
        at1 += first[4 * q + 3];
    }
    if (tid == kPraqListBlock - 1) {
        lcount[0] = scan[0][tid];
        lcount[1] = scan[1][tid];
    }
}

// One block per instance: its node and edge blocks into the lists.
__global__ void __launch_bounds__(256)
wpraq_fill_lists(const int* first, const int* at, int4* node_list, int4* edge_list) {
    const int q = blockIdx.x;
    const int an = at[3 * q];
    if (an < 0) return;
    const int ae = at[3 * q + 1], p = at[3 * q + 2];
    const int nb = first[4 * q + 1], eb = first[4 * q + 3];
    for (int k = threadIdx.x; k < nb; k += blockDim.x) node_list[an + k] = make_int4(q, k, p, 0);
    for (int k = threadIdx.x; k < eb; k += blockDim.x) edge_list[ae + k] = make_int4(q, k, p, 0);
}

// The point a sums block of a listed instance bounds: the step point (sx,
// sy) after Halpern steps and a restart step, else (x, y).
__device__ __forceinline__ int listed_sums_point(const RoundStep& rs) {
    return rs.act == kActHalpernBlock || rs.act == kActRestart ? 2 : 1;
}

// Over the work lists of the mode (the instances that took its steps):
// blocks [0, node_blocks) the node list, the rest the edge list; one
// partial row per table block, at its table position.
__global__ void __launch_bounds__(256) wpraq_sums(const PJ32* j32, const PJ64* j64,
                                                  const int4* node_list, int node_blocks,
                                                  const int4* edge_list, const int* lcount,
                                                  const RoundStep* rstep, const int* first,
                                                  double* partial) {
    __shared__ double red[kPraqSums * 16];
    const bool node = int(blockIdx.x) < node_blocks;
    const int k = node ? int(blockIdx.x) : int(blockIdx.x) - node_blocks;
    const int n = lcount[node ? 0 : 1];
    const int4 t = node ? node_list[k] : edge_list[k];
    if (k >= n) return;
    const size_t row = node ? size_t(first[4 * t.x] + t.y)
                            : size_t(node_blocks + first[4 * t.x + 2] + t.y);
    praq_listed(j32, j64, t, [&](const auto& job) {
        using Real = RealOf<decltype(job)>;
        const int where = listed_sums_point(rstep[t.x]);
        const PraqView<Real> a = praq_view(job.a);
        const gptr<const Real> bx = where == 2 ? a.sx : a.x;
        const gptr<const Real> by = where == 2 ? a.sy : a.y;
        double acc[kPraqSums];
#pragma unroll
        for (int k = 0; k < kPraqSums; ++k) acc[k] = 0.0;
        const int gid = t.y * kWideBlock + threadIdx.x;
        if (node)
            praq_node_sums(a, gid / job.lanes, gid % job.lanes, job.lanes, bx, by, acc);
        else if (gid < a.m)
            praq_edge_sums(a, gid, bx, by, true, acc);
        block_sums<kPraqSums>(acc, red);
        if (threadIdx.x == 0)
            for (int q = 0; q < kPraqSums; ++q) partial[row * kPraqSums + q] = acc[q];
    });
}

// one block per instance: its rows in a fixed order, then the logic
__global__ void __launch_bounds__(256) wpraq_logic(const PJ32* j32, const PJ64* j64,
                                                   const int* first, int node_blocks,
                                                   const double* partial, const int* round_action,
                                                   int mode) {
    __shared__ double red[kPraqSums * 16];
    const int inst = blockIdx.x;
    dispatch(j32, j64, inst, [&](const auto& job) {
        const int act = wide_logic_action(job, round_action, inst, mode);
        if (act < 0) return;
        const int nf = first[4 * inst], nc = first[4 * inst + 1];
        const int ef = first[4 * inst + 2], ec = first[4 * inst + 3];
        double acc[kPraqSums];
#pragma unroll
        for (int k = 0; k < kPraqSums; ++k) acc[k] = 0.0;
        for (int r = threadIdx.x; r < nc; r += blockDim.x)
            for (int k = 0; k < kPraqSums; ++k) acc[k] += partial[size_t(nf + r) * kPraqSums + k];
        for (int r = threadIdx.x; r < ec; r += blockDim.x)
            for (int k = 0; k < kPraqSums; ++k)
                acc[k] += partial[size_t(node_blocks + ef + r) * kPraqSums + k];
        block_sums<kPraqSums>(acc, red);
        if (threadIdx.x == 0) wide_apply(job, act, praq_check_sums(acc, job.scale));
    });
}

// copies over node blocks then edge blocks, in the instance's precision
__global__ void __launch_bounds__(256) wpraq_copy(const PJ32* j32, const PJ64* j64,
                                                  const int2* node_tab, int node_blocks,
                                                  const int2* edge_tab, const int* round_action,
                                                  const int* pending, const int* list_at,
                                                  int mode) {
    const bool node = int(blockIdx.x) < node_blocks;
    const int2 t = node ? node_tab[blockIdx.x] : edge_tab[blockIdx.x - node_blocks];
    if (!copy_may_apply(mode, round_action, pending, list_at, t.x)) return;
    dispatch(j32, j64, t.x, [&](const auto& job) {
        const int kind = wide_copy_kind(job.state, round_action, pending, t.x, mode);
        if (kind < 0) return;
        const auto a = praq_view(job.a);
        const int gid = t.y * kWideBlock + threadIdx.x;
        if (node) {
            if (gid % job.lanes) return;
            const int i = gid / job.lanes;
            if (i >= a.n) return;
            if (kind <= 1) a.x0[i] = a.x[i];
            else if (kind == 2) a.x[i] = a.sx[i];
            else a.x0[i] = a.x[i] = a.sx[i] = a.xs[i];
        } else {
            if (gid >= a.m) return;
            for (int slot = 0; slot < 2; ++slot) {
                const int k = slot * a.m + gid;
                if (kind <= 1) a.y0[k] = a.y[k];
                else if (kind == 2) a.y[k] = a.sy[k];
                else a.y0[k] = a.y[k] = a.sy[k] = a.ys[k];
            }
        }
    });
}

// FP32 instances whose logic asked for FP64: their Halpern state, block by
// block, into the FP64 image (the precision flips in wide_promote)
__global__ void __launch_bounds__(256) wpraq_convert(const PJ32* j32, const PJ64* j64,
                                                     const int2* node_tab, int node_blocks,
                                                     const int2* edge_tab) {
    const bool node = int(blockIdx.x) < node_blocks;
    const int2 t = node ? node_tab[blockIdx.x] : edge_tab[blockIdx.x - node_blocks];
    if (!wide_hands_over(j64[t.x].state)) return;
    const PraqView<float> f = praq_view(j32[t.x].a);
    const PraqView<double> d = praq_view(j64[t.x].a);
    const int gid = t.y * kWideBlock + threadIdx.x;
    if (node) {
        const int lanes = j32[t.x].lanes;
        if (gid % lanes) return;
        const int i = gid / lanes;
        if (i >= f.n) return;
        d.x[i] = f.x[i];
        d.sx[i] = f.sx[i];
        d.x0[i] = f.x0[i];
    } else {
        if (gid >= f.m) return;
        for (int slot = 0; slot < 2; ++slot) {
            const int k = slot * f.m + gid;
            d.yc[k] = f.yc[k];
            d.y[k] = f.y[k];
            d.sy[k] = f.sy[k];
            d.y0[k] = f.y0[k];
        }
    }
}

// ---------------------------------------------------------------- PRIA ----

using RJ32 = PriaJob<float>;
using RJ64 = PriaJob<double>;

// ---- work lists ----
//
// Each round the step kernels run over lists of the table entries whose
// instances take steps, one list per (table, precision): x tiles FP32, x
// tiles FP64, slices FP32, slices FP64 (kListX32 ...), in table order.  A
// step launch has a block per table entry; block b takes item b of its list,
// so the blocks with work are dispatched first and the rest exit after one
// load: no block reads the job record of an instance of the other precision
// or of one that takes no step.  Which block takes an item does not change
// the item's arithmetic.  The lists hold the table entries themselves, and
// the step scalars of each listed instance sit in a RoundStep record, so a
// step block's first loads are its entry, then its job record and
// RoundStep side by side, then its data.
enum { kListX32 = 0, kListX64 = 1, kListS32 = 2, kListS64 = 3, kLists = 4 };

constexpr int kListBlock = 1024;

// One block: the instances taking steps in this round (mode 0) or taking a
// restart step (mode 1), in table order: per instance its RoundStep and the
// offsets at[3q..3q+2] = (x offset, slice offset, precision) of its ranges in
// the lists of its precision (-1: no steps); the lengths go to
// lcount[0..3].  wpria_fill_lists then copies the entries.  Every instance's
// x range is a whole number of groups of kXcds entries, so each range starts
// in its list at a multiple of kXcds.
__global__ void __launch_bounds__(kListBlock)
wpria_lists(const RJ64* j64, int count, const int* first, const int* xr,
            const int* round_action, int mode, RoundStep* rstep, int* at_out, int* lcount) {
    __shared__ int scan[kLists][kListBlock];
    const int tid = threadIdx.x;
    // a contiguous chunk of instances per thread keeps the lists in table order
    const int chunk = (count + kListBlock - 1) / kListBlock;
    const int q0 = min(count, tid * chunk), q1 = min(count, q0 + chunk);
    int mine[kLists] = {0, 0, 0, 0};
    for (int q = q0; q < q1; ++q) {
        const LogicState* st = j64[q].state;
        const int act = mode == 0 ? round_action[q] : int(kActRestart);
        const bool runs = mode == 0 ? wide_runs_steps(act) : wide_restarting(st);
        if (!runs) continue;
        const int p = st->in64 ? 1 : 0;
        mine[kListX32 + p] += xr[2 * q + 1];
        mine[kListS32 + p] += first[4 * q + 1];
        rstep[q] = round_step_of(j64[q], st, act, mode);
    }
    for (int l = 0; l < kLists; ++l) scan[l][tid] = mine[l];
    __syncthreads();
    // inclusive scan (Hillis-Steele) of each list's counts over the threads
    for (int off = 1; off < kListBlock; off <<= 1) {
        int v[kLists];
        for (int l = 0; l < kLists; ++l) v[l] = tid >= off ? scan[l][tid - off] : 0;
        __syncthreads();
        for (int l = 0; l < kLists; ++l) scan[l][tid] += v[l];
        __syncthreads();
    }
    int at[kLists];
    for (int l = 0; l < kLists; ++l) at[l] = scan[l][tid] - mine[l];
    for (int q = q0; q < q1; ++q) {
        const LogicState* st = j64[q].state;
        const bool runs = mode == 0 ? wide_runs_steps(round_action[q]) : wide_restarting(st);
        if (!runs) {
            at_out[3 * q] = -1;
            continue;
        }
        const int p = st->in64 ? 1 : 0;
        at_out[3 * q] = at[kListX32 + p];
        at_out[3 * q + 1] = at[kListS32 + p];
        at_out[3 * q + 2] = p;
        at[kListX32 + p] += xr[2 * q + 1];
        at[kListS32 + p] += first[4 * q + 1];
    }
    if (tid == kListBlock - 1)
        for (int l = 0; l < kLists; ++l) lcount[l] = scan[l][tid];
}

// One block per instance: its table entries into the lists at the offsets
// wpria_lists left.
__global__ void __launch_bounds__(256)
wpria_fill_lists(const int* first, const int* xr, const int* at, const int4* x_tab,
                 const int2* s_tab, int4* xlists, int2* slists, int x_cap, int s_cap) {
    const int q = blockIdx.x;
    const int ax = at[3 * q];
    if (ax < 0) return;
    const int as = at[3 * q + 1], p = at[3 * q + 2];
    int4* lx = xlists + size_t(p) * x_cap + ax;
    int2* ls = slists + size_t(p) * s_cap + as;
    const int x0 = xr[2 * q], nx = xr[2 * q + 1];
    const int s0 = first[4 * q], ns = first[4 * q + 1];
    for (int k = threadIdx.x; k < nx; k += blockDim.x) lx[k] = x_tab[x0 + k];
    for (int k = threadIdx.x; k < ns; k += blockDim.x) ls[k] = s_tab[s0 + k];
}

// Batches of x1 an FP32 x-tile thread of a `bs`-thread workgroup holds for
// its write phase when every T' <= 320 (bs / 32 rows per pass, kPriaXBatch
// passes per batch): 5 with 256 threads, 3 with 512
constexpr int x_hold_batches(int bs) {
    return (kPriaCompactT + (bs / kWideTW) * kPriaXBatch - 1) / ((bs / kWideTW) * kPriaXBatch);
}

template <class Real>
__device__ __forceinline__ const PriaJob<Real>& pria_job(const RJ32* j32, const RJ64* j64,
                                                         int inst) {
    if constexpr (sizeof(Real) == 8) return j64[inst];
    else return j32[inst];
}

// The x pass over the list of one precision, one x tile (32 columns of one
// item) per item: FP32 in 256-thread blocks (8 rows of the tile at a time;
// 512 threads and 16 rows with the tuning table's batch.pria.x_threads),
// FP64 in 512 threads (16 rows at a time; 256-byte tile rows, so LDS
// T' x 33 x 8 bytes: two blocks per CU).  TW = 16 is the FP64 fallback for
// T' whose 32-column tile does not fit the LDS: the tile's two halves in
// two 256-thread workgroups, still 16 lanes per column, so an instance's
// sums do not depend on which form its batch takes.
template <class Real, int TW, int HOLD = 0, int BS = (sizeof(Real) == 4 ? 256 : 512)>
__global__ void __launch_bounds__(BS)
wpria_x(const RJ32* j32, const RJ64* j64, const int4* list, const int* lcount,
        const RoundStep* rstep, int j) {
    extern __shared__ char dyn[];
    // TW < kWideTW: two workgroups per table entry, one per half tile
    constexpr int per = kWideTW / TW;
    const int n = *lcount, b = int(blockIdx.x) / per;
    const int4 t = list[b];   // the grid is per times the list's capacity
    if (b >= n || t.y < 0) return;   // past the list; item-group padding
    const int t0 = t.z + TW * (int(blockIdx.x) % per);
    const PriaJob<Real>& job = pria_job<Real>(j32, j64, t.x);
    if (t0 >= job.a.T) return;
    PriaStep<Real> sc;
    if (!round_step(rstep[t.x], j, pria_step_scalars<Real>, &sc)) return;
    const PriaView<Real> a = pria_view(job.a);
    pria_x_tile<Real, TW, HOLD>(a, sc, t.y, t0, reinterpret_cast<Real*>(dyn));
}

// Register shape of the y/z pass by T' (the compact one up to T' = 320).
template <class Real, bool compact>
struct PriaWideShape {
    static constexpr int VR = compact ? PriaCompactShape<Real>::VR : PriaShape<Real>::VR;
    static constexpr int VZ = compact ? PriaCompactShape<Real>::VZ : PriaShape<Real>::VZ;
};

// The y/z pass over the list of one precision, one slice per item, 512
// threads.  With `compact` (every instance's T' <= 320) the register shape is
// the compact one, held to 80 VGPRs so that three blocks share a CU (six
// waves per SIMD); the shapes give the same sums (row widths follow the row
// length only).
template <class Real, bool compact>
__global__ void __launch_bounds__(512, compact ? 6 : 3)
wpria_yz(const RJ32* j32, const RJ64* j64, const int2* list, const int* lcount,
         const RoundStep* rstep, int j) {
    __shared__ double red[kPriaYZRed];
    extern __shared__ double zcache[];   // kPriaZCacheBytes: z + r of a long slice
    const int n = *lcount;
    const int2 t = list[blockIdx.x];   // the grid is the list's capacity
    if (int(blockIdx.x) >= n) return;
    const PriaJob<Real>& job = pria_job<Real>(j32, j64, t.x);
    PriaStep<Real> sc;
    if (!round_step(rstep[t.x], j, pria_step_scalars<Real>, &sc)) return;
    using Shape = PriaWideShape<Real, compact>;
    pria_yz_slice<Real, Shape::VR, Shape::VZ>(pria_view(job.a), sc, t.y, red, 3, 0, 1 << 30,
                                             reinterpret_cast<Real*>(zcache),
                                             kPriaZCacheBytes / int(sizeof(Real)));
}

// Over the work lists of the mode (the instances that took its steps):
// blocks [0, s_blocks) the slice lists (FP32, then FP64), six sums each at
// the slice's table row; the rest the x-tile lists, a column-minimum sum
// each at the tile's (item, tile) row.
__global__ void __launch_bounds__(256) wpria_sums(const RJ32* j32, const RJ64* j64,
                                                  const int2* slists, int s_blocks,
                                                  const int4* xlists, int x_blocks,
                                                  const int* lcount, const RoundStep* rstep,
                                                  const int* first, double* partial) {
    __shared__ double red[kPriaSums * 16];
    const bool slice = int(blockIdx.x) < s_blocks;
    int k = slice ? int(blockIdx.x) : int(blockIdx.x) - s_blocks;
    const int n32 = lcount[slice ? kListS32 : kListX32], n64 = lcount[slice ? kListS64 : kListX64];
    const bool in64 = k >= n32;
    if (in64) k -= n32;
    if (k >= (in64 ? n64 : n32)) return;
    int4 t;
    if (slice) {
        const int2 e = slists[(in64 ? s_blocks : 0) + k];
        t = make_int4(e.x, e.y, 0, 0);
    } else {
        t = xlists[(in64 ? x_blocks : 0) + k];
    }
    if (!slice && t.y < 0) return;   // item-group padding
    auto body = [&](const auto& job) {
        using Real = RealOf<decltype(job)>;
        const int where = listed_sums_point(rstep[t.x]);
        const PriaView<Real> a = pria_view(job.a);
        double acc[kPriaSums];
        for (int q = 0; q < kPriaSums; ++q) acc[q] = 0.0;
        if (slice) {
            pria_slice_sums(a, t.y, where == 2 ? a.sx : a.x, red, acc);
            const size_t row = size_t(first[4 * t.x] + t.y);
            if (threadIdx.x == 0)
                for (int q = 0; q < kPLower; ++q) partial[row * kPriaSums + q] = acc[q];
        } else {
            pria_column_sums<Real, kWideTW>(a, t.y, t.z, where == 2 ? a.sy : a.y,
                                            where == 2 ? a.sz : a.z, red, acc);
            if (threadIdx.x == 0) partial[size_t(s_blocks + t.w) * kPriaSums] = acc[kPLower];
        }
    };
    if (in64) body(j64[t.x]);
    else body(j32[t.x]);
}

__global__ void __launch_bounds__(256) wpria_logic(const RJ32* j32, const RJ64* j64,
                                                   const int* first, int s_blocks,
                                                   const double* partial, const int* round_action,
                                                   int mode) {
    __shared__ double red[kPriaSums * 16];
    const int inst = blockIdx.x;
    dispatch(j32, j64, inst, [&](const auto& job) {
        const int act = wide_logic_action(job, round_action, inst, mode);
        if (act < 0) return;
        const int sf = first[4 * inst], sc = first[4 * inst + 1];
        const int tf = first[4 * inst + 2], tc = first[4 * inst + 3];
        double acc[kPriaSums];
        for (int q = 0; q < kPriaSums; ++q) acc[q] = 0.0;
        for (int r = threadIdx.x; r < sc; r += blockDim.x)
            for (int q = 0; q < kPLower; ++q) acc[q] += partial[size_t(sf + r) * kPriaSums + q];
        for (int r = threadIdx.x; r < tc; r += blockDim.x)
            acc[kPLower] += partial[size_t(s_blocks + tf + r) * kPriaSums];
        block_sums<kPriaSums>(acc, red);
        if (threadIdx.x == 0) wide_apply(job, act, pria_check_sums(acc, job.scale));
    });
}

// one block per slice, in the instance's precision; the start copies also
// clear the slice's thresholds
__global__ void __launch_bounds__(256) wpria_copy(const RJ32* j32, const RJ64* j64,
                                                  const int2* s_tab, const int* round_action,
                                                  const int* pending, const int* list_at,
                                                  int mode) {
    const int2 t = s_tab[blockIdx.x];
    if (!copy_may_apply(mode, round_action, pending, list_at, t.x)) return;
    dispatch(j32, j64, t.x, [&](const auto& job) {
        const int kind = wide_copy_kind(job.state, round_action, pending, t.x, mode);
        if (kind < 0) return;
        const auto a = pria_view(job.a);
        const int s = t.y, base = pria_off(a.N, a.T, s), L = a.N * (a.T - s);
        if (kind == 3) {  // no thresholds yet: column (i, s), row (i, s), slice s
            const double none = __builtin_nan("");
            for (int i = threadIdx.x; i < a.N; i += blockDim.x) {
                a.thx[i * a.T + s] = none;
                a.thy[i * a.T + s] = none;
            }
            if (threadIdx.x == 0) a.thz[s] = none;
        }
        for (int e = threadIdx.x; e < L; e += blockDim.x) {
            const int g = base + e;
            if (kind <= 1) {
                a.x0[g] = a.x[g];
                a.y0[g] = a.y[g];
                a.z0[g] = a.z[g];
            } else if (kind == 2) {
                a.x[g] = a.sx[g];
                a.y[g] = a.sy[g];
                a.z[g] = a.sz[g];
            } else {
                a.x0[g] = a.x[g] = a.sx[g] = a.xs[g];
                a.y0[g] = a.y[g] = a.sy[g] = a.ys[g];
                a.z0[g] = a.z[g] = a.sz[g] = a.zs[g];
            }
        }
    });
}

// FP32 instances whose logic asked for FP64: one slice's state (and the
// thresholds of the slice's column, row and slice) into the FP64 image
__global__ void __launch_bounds__(256) wpria_convert(const RJ32* j32, const RJ64* j64,
                                                     const int2* s_tab) {
    const int2 t = s_tab[blockIdx.x];
    if (!wide_hands_over(j64[t.x].state)) return;
    const PriaView<float> f = pria_view(j32[t.x].a);
    const PriaView<double> d = pria_view(j64[t.x].a);
    const int s = t.y, base = pria_off(f.N, f.T, s), L = f.N * (f.T - s);
    for (int i = threadIdx.x; i < f.N; i += blockDim.x) {
        d.thx[i * f.T + s] = f.thx[i * f.T + s];
        d.thy[i * f.T + s] = f.thy[i * f.T + s];
    }
    if (threadIdx.x == 0) d.thz[s] = f.thz[s];
    for (int e = threadIdx.x; e < L; e += blockDim.x) {
        const int g = base + e;
        d.x0[g] = f.x0[g];
        d.y0[g] = f.y0[g];
        d.z0[g] = f.z0[g];
        d.xc[g] = f.xc[g];
        d.yc[g] = f.yc[g];
        d.zc[g] = f.zc[g];
        d.x[g] = f.x[g];
        d.sx[g] = f.sx[g];
        d.y[g] = f.y[g];
        d.sy[g] = f.sy[g];
        d.z[g] = f.z[g];
        d.sz[g] = f.sz[g];
    }
}

}  // namespace

// ------------------------------------------------------------------ runs --

int WideRun::begin(WideControl* engine, hipStream_t s, double limit) {
    w = engine;
    stream = s;
    time_limit = limit;
    queued = retired = 0;
    stats = WideStats();
    t0 = std::chrono::steady_clock::now();
    finished = true;
    if (!w) return 0;
    if (w->shrunk() && w->restore()) return kErrorDevice;
    if (w->count == 0) return 0;
    w->expired.assign(w->deadlines.size(), 0);
    if (w->start(stream)) return kErrorDevice;
    // the round graph (timed rounds are launched directly instead: events
    // recorded by graph nodes carry no timestamps for hipEventElapsedTime)
    if (w->use_graph && !w->timing && !w->exec) {
        REFORM_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        const int st = w->enqueue_round(stream);
        REFORM_HIP_CHECK(hipStreamEndCapture(stream, &w->graph));
        if (st) return kErrorDevice;
        REFORM_HIP_CHECK(hipGraphInstantiate(&w->exec, w->graph, nullptr, nullptr, 0));
    }
    REFORM_HIP_CHECK(
        hipMemcpyAsync(w->host_active, w->active, sizeof(int), hipMemcpyDeviceToHost, stream));
    REFORM_HIP_CHECK(hipStreamSynchronize(stream));
    finished = w->host_active[0] == 0;
    return 0;
}

int WideRun::advance() {
    if (finished) return 0;
    int changed = 0;
    bool drained = false;
    while (retired < queued) {
        const int slot = retired & 1;
        const hipError_t q = hipEventQuery(w->polled[slot]);
        if (q == hipErrorNotReady) break;
        REFORM_HIP_CHECK(q);
        if (w->timing && w->collect(slot)) return kErrorDevice;
        ++retired;
        changed = 1;
        if (w->host_active[slot] == 0) {
            drained = true;
            break;
        }
    }
    const double elapsed =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // out of time: only once a round has run, so every instance has started
    if (drained || (time_limit > 0.0 && elapsed > time_limit && retired > 0)) {
        REFORM_HIP_CHECK(hipStreamSynchronize(stream));
        if (!drained) {
            if (w->expire(stream)) return kErrorDevice;
            REFORM_HIP_CHECK(hipStreamSynchronize(stream));
        }
        // the full tables again, for the results' accounting and the next solve
        if (w->shrunk() && w->restore()) return kErrorDevice;
        finished = true;
        stats.seconds = elapsed;
        return 1;
    }
    // compaction: once at most half the table is active, the tables are
    // rebuilt over the active instances, after the queued rounds have run
    if (shrink && retired > 0 && w->count >= kShrinkMin) {
        const int active_now = w->host_active[(retired - 1) & 1];
        if (active_now > 0 && 2 * active_now <= w->count) {
            if (queued > retired) return changed;
            if (w->shrink_to_live(stream)) return kErrorDevice;
            ++stats.shrinks;
            changed = 1;
        }
    }
    // instances past their own limits (again only once a round has run)
    if (!w->deadlines.empty() && retired > 0) {
        const int n = w->expire_due(stream, elapsed);
        if (n < 0) return n;
    }
    while (queued - retired < 2) {
        const int slot = queued & 1;
        const bool graph = w->use_graph && !w->timing;
        const int rounds = graph ? kRoundsPerPoll : 1;
        for (int r = 0; r < rounds; ++r) {
            if (graph) {
                if (w->launch_round(stream)) return kErrorDevice;  // captured after a shrink
            } else {  // the round's marks (if timing) in the group's event set
                w->mark_set = slot;
                w->marks_used[slot] = 0;
                w->mark_kinds[slot].clear();
                if (w->enqueue_round(stream)) return kErrorDevice;
            }
        }
        REFORM_HIP_CHECK(hipMemcpyAsync(w->host_active + slot, w->active, sizeof(int),
                                        hipMemcpyDeviceToHost, stream));
        REFORM_HIP_CHECK(hipEventRecord(w->polled[slot], stream));
        ++queued;
        stats.rounds += rounds;
        changed = 1;
    }
    return changed;
}

int WideControl::launch_round(hipStream_t stream) {
    if (!use_graph || timing) return enqueue_round(stream);
    if (!exec) {
        REFORM_HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        const int st = enqueue_round(stream);
        REFORM_HIP_CHECK(hipStreamEndCapture(stream, &graph));
        if (st) return kErrorDevice;
        REFORM_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    REFORM_HIP_CHECK(hipGraphLaunch(exec, stream));
    return 0;
}

int WideControl::solve(hipStream_t stream, double time_limit, WideStats* stats) {
    WideRun run;
    if (run.begin(this, stream, time_limit)) return kErrorDevice;
    while (!run.finished) {
        const int st = run.advance();
        if (st < 0) return st;
        if (!st) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    if (stats) *stats = run.stats;
    return 0;
}

// ---------------------------------------------------------------- control --

void WideControl::set_deadlines(std::vector<double> d) {
    deadlines = std::move(d);
    expired.assign(deadlines.size(), 0);
}

void WideControl::stop_position(int f) {
    if (f < 0 || f >= full_count) return;
    if (deadlines.empty())
        deadlines.assign(size_t(full_count), std::numeric_limits<double>::infinity());
    expired.resize(deadlines.size(), 0);
    deadlines[size_t(f)] = -std::numeric_limits<double>::infinity();
}

// The mask is copied from the host latches on the stream (pageable memory:
// the copy may read them after this call returns).  The latches only ever
// go from 0 to 1, and only once their deadline has passed, so a copy that
// sees a later write stops an instance that is already due, nothing else.
int WideControl::expire_due(hipStream_t stream, double now) {
    if (deadlines.empty() || !count) return 0;
    expired.resize(deadlines.size(), 0);
    const int nd = int(deadlines.size());
    auto full = [&](int q) { return table_index.empty() ? q : table_index[q]; };
    int due = 0;
    for (int q = 0; q < count; ++q) {
        const int f = full(q);
        if (f < nd && !expired[f] && now > deadlines[f]) {
            expired[f] = 1;
            ++due;
        }
    }
    if (!due) return 0;
    mask_host.resize(count);
    for (int q = 0; q < count; ++q) {
        const int f = full(q);
        mask_host[q] = f < nd ? expired[f] : 0;
    }
    REFORM_HIP_CHECK(hipMemcpyAsync(expire_mask, mask_host.data(), sizeof(int) * size_t(count),
                                    hipMemcpyHostToDevice, stream));
    if (expire(stream, expire_mask)) return kErrorDevice;
    return due;
}

int WideControl::shrink_to_live(hipStream_t stream) {
    if (live_flags(stream)) return kErrorDevice;
    std::vector<int> flags(count);
    REFORM_HIP_CHECK(hipMemcpyAsync(flags.data(), live, sizeof(int) * size_t(count),
                                    hipMemcpyDeviceToHost, stream));
    REFORM_HIP_CHECK(hipStreamSynchronize(stream));
    std::vector<int> keep;
    keep_has32 = false;
    for (int q = 0; q < count; ++q) {
        if (flags[q]) keep.push_back(table_index.empty() ? q : table_index[q]);
        keep_has32 = keep_has32 || flags[q] == 1;
    }
    if (keep.empty() || int(keep.size()) == count) return 0;
    return shrink(stream, keep);
}

template <class J32, class J64>
int WideSubset<J32, J64>::reserve(int n) {
    if (n <= cap) return 0;
    if (mem) (void)hipFree(mem);
    mem = nullptr;
    cap = 0;
    const size_t o64 = align256(sizeof(J32) * size_t(n));
    const size_t ok = o64 + align256(sizeof(J64) * size_t(n));
    REFORM_HIP_CHECK(device_malloc(&mem, ok + sizeof(int) * size_t(n)));
    char* b = static_cast<char*>(mem);
    jobs32 = reinterpret_cast<J32*>(b);
    jobs64 = reinterpret_cast<J64*>(b + o64);
    keep = reinterpret_cast<int*>(b + ok);
    cap = n;
    return 0;
}

template struct WideSubset<PraqJob<float>, PraqJob<double>>;
template struct WideSubset<PriaJob<float>, PriaJob<double>>;

namespace {

// The subset's records gathered on the device; the host copies for the
// table build.
template <class J32, class J64>
int gather_subset(WideSubset<J32, J64>& sub, const std::vector<J64>& full64, const J32* dev32,
                  const J64* dev64, const std::vector<int>& keep, hipStream_t stream,
                  std::vector<J64>& host64) {
    const int n = int(keep.size());
    if (sub.reserve(n)) return kErrorDevice;
    REFORM_HIP_CHECK(hipMemcpyAsync(sub.keep, keep.data(), sizeof(int) * size_t(n),
                                    hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(wide_gather_jobs<J64>, dim3(grid_for(n)), dim3(kWideBlock), 0, stream,
                       dev64, sub.keep, n, sub.jobs64);
    if (dev32)
        hipLaunchKernelGGL(wide_gather_jobs<J32>, dim3(grid_for(n)), dim3(kWideBlock), 0, stream,
                           dev32, sub.keep, n, sub.jobs32);
    REFORM_HIP_CHECK(hipGetLastError());
    REFORM_HIP_CHECK(hipStreamSynchronize(stream));
    host64.resize(n);
    for (int q = 0; q < n; ++q) host64[q] = full64[keep[q]];
    return 0;
}

}  // namespace

WideControl::~WideControl() {
    release();
    for (auto& set : marks)
        for (hipEvent_t e : set) (void)hipEventDestroy(e);
    for (auto& e : polled)
        if (e) (void)hipEventDestroy(e);
    if (host_active) (void)hipHostFree(host_active);
}

void WideControl::release() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (arena) (void)hipFree(arena);
    exec = nullptr;
    graph = nullptr;
    arena = nullptr;
    bytes = 0;
}

int WideControl::mark_events(int kind, hipEvent_t* begin, hipEvent_t* end) {
    *begin = *end = nullptr;
    if (!timing) return 0;
    std::vector<hipEvent_t>& set = marks[mark_set];
    for (hipEvent_t* out : {begin, end}) {
        const int k = marks_used[mark_set]++;
        if (k == int(set.size())) {
            // timing only: no system-scope fence (its cache writeback would
            // be timed with the kernels); the poll event that follows them
            // on the stream keeps its fence
            hipEvent_t e;
            REFORM_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
            set.push_back(e);
        }
        *out = set[k];
    }
    mark_kinds[mark_set].push_back(kind);
    return 0;
}

int WideControl::collect(int set) {
    const std::vector<int>& kinds = mark_kinds[set];
    for (size_t p = 0; p < kinds.size(); ++p) {
        float ms = 0.0f;
        REFORM_HIP_CHECK(hipEventElapsedTime(&ms, marks[set][2 * p], marks[set][2 * p + 1]));
        kernel_ms[kinds[p]] += ms;
        launches[kinds[p]] += 1;
    }
    return 0;
}

void WideControl::clear_timing() {
    for (int k = 0; k < kKinds; ++k) {
        kernel_ms[k] = 0.0;
        launches[k] = 0;
    }
}

void WideControl::set_timing(bool on) { timing = on; }

void WideControl::algorithmic_bytes(const std::vector<LogicState>& states,
                                    double (&out)[kKinds]) const {
    for (int k = 0; k < kKinds; ++k) {
        out[k] = 0.0;
        for (int q = 0; q < count && q < int(states.size()); ++q)
            out[k] += double(states[q].steps32) * step_bytes[k][0][q] +
                      double(states[q].steps64) * step_bytes[k][1][q];
    }
}

int WideControl::allocate(size_t table_bytes, size_t partial_rows, int partial_width) {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    size_t at = align256(table_bytes);
    const size_t o_first = at;
    at = align256(at + sizeof(int) * 4 * size_t(count));
    const size_t o_partial = at;
    at = align256(at + sizeof(double) * partial_width * (partial_rows + 1));
    const size_t o_ctrl = at;
    at = align256(at + sizeof(int) * (4 * size_t(count) + 1));
    // the previous arena serves while it is large enough (tables rebuilt
    // between rounds); a larger one grows by half again
    if (!arena || at > bytes) {
        const size_t grow = arena ? std::max(at, bytes + bytes / 2) : at;
        release();
        hipError_t e = device_malloc(&arena, grow);
        if (e != hipSuccess) {
            arena = nullptr;
            record_error("hipMalloc(wide arena)", e, __FILE__, __LINE__);
            return kErrorDevice;
        }
        bytes = grow;
    }
    char* b = static_cast<char*>(arena);
    first = reinterpret_cast<int*>(b + o_first);
    partial = reinterpret_cast<double*>(b + o_partial);
    round_action = reinterpret_cast<int*>(b + o_ctrl);
    pending = round_action + count;
    active = pending + count;
    expire_mask = active + 1;
    live = expire_mask + count;
    if (!host_active) {
        hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&host_active), 2 * sizeof(int),
                                     hipHostMallocDefault);
        if (e != hipSuccess) {
            host_active = nullptr;
            record_error("hipHostMalloc(active count)", e, __FILE__, __LINE__);
            return kErrorDevice;
        }
    }
    for (auto& e : polled)
        if (!e) REFORM_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return 0;
}

// ------------------------------------------------------------- PRAQ host --

int WidePraq::setup(std::vector<PraqJob<double>> host64, const PraqJob<float>* dev32,
                    const PraqJob<double>* dev64) {
    full64 = std::move(host64);
    full_dev32 = dev32;
    full_dev64 = dev64;
    full_count = int(full64.size());
    table_index.clear();
    return build_tables(full64, dev32, dev64);
}

int WidePraq::shrink(hipStream_t stream, const std::vector<int>& keep) {
    std::vector<PraqJob<double>> host64;
    if (gather_subset(subset, full64, full_dev32, full_dev64, keep, stream, host64))
        return kErrorDevice;
    table_index = keep;
    return build_tables(host64, full_dev32 ? subset.jobs32 : nullptr, subset.jobs64);
}

int WidePraq::restore() {
    table_index.clear();
    return build_tables(full64, full_dev32, full_dev64);
}

int WidePraq::live_flags(hipStream_t stream) {
    hipLaunchKernelGGL(wide_live<PraqJob<double>>, dim3(grid_for(count)), dim3(kWideBlock), 0,
                       stream, jobs64, count, live);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

int WidePraq::build_tables(const std::vector<PraqJob<double>>& host64,
                           const PraqJob<float>* dev32, const PraqJob<double>* dev64) {
    count = int(host64.size());
    jobs32 = dev32;
    jobs64 = dev64;
    has32 = dev32 != nullptr;
    if (!count) {
        release();
        return 0;
    }
    std::vector<int2> nt, et;
    std::vector<int> fi(4 * size_t(count));
    op_steps = 2;
    // algorithmic bytes of one step, b = bytes per value: node kernel: ptr 4,
    // six per-node values, the reflected gather, and (index 4 + value b) per
    // half-edge of the CSR scatter; edge kernel: ends 8, weight, the y pair,
    // the anchor pair and the written pair
    for (int k = 0; k < kKinds; ++k)
        for (int p = 0; p < 2; ++p) step_bytes[k][p].assign(count, 0.0);
    for (int q = 0; q < count; ++q) {
        const auto& a = host64[q].a;
        const int lanes = host64[q].lanes;
        op_steps = std::max(op_steps, int(host64[q].p.op_steps));
        const int nb = int((int64_t(a.n) * lanes + kWideBlock - 1) / kWideBlock);
        const int eb = (a.m + kWideBlock - 1) / kWideBlock;
        fi[4 * q] = int(nt.size());
        fi[4 * q + 1] = nb;
        fi[4 * q + 2] = int(et.size());
        fi[4 * q + 3] = eb;
        for (int b = 0; b < nb; ++b) nt.push_back(make_int2(q, b));
        for (int b = 0; b < eb; ++b) et.push_back(make_int2(q, b));
        for (int p = 0; p < 2; ++p) {
            const double b = p ? 8.0 : 4.0, n = a.n, m = a.m;
            step_bytes[0][p][q] = n * (4.0 + 7.0 * b) + 2.0 * m * (4.0 + b);
            step_bytes[1][p][q] = m * (8.0 + 7.0 * b);
        }
    }
    node_blocks = int(nt.size());
    edge_blocks = int(et.size());
    // arena: node and edge tables, the step lists (node, edge), their
    // lengths, the RoundStep records and list offsets
    const size_t t_bytes = align256(sizeof(int2) * (nt.size() + et.size()));
    const size_t l_bytes = align256(sizeof(int4) * (nt.size() + et.size()));
    const size_t c_bytes = align256(sizeof(int) * 2);
    const size_t r_bytes = align256(sizeof(RoundStep) * size_t(count));
    const size_t tab_bytes = t_bytes + l_bytes + c_bytes + r_bytes + sizeof(int) * 3 * size_t(count);
    if (allocate(tab_bytes, nt.size() + et.size(), kPraqSums)) return kErrorDevice;
    char* ab = static_cast<char*>(arena);
    node_tab = reinterpret_cast<int2*>(ab);
    edge_tab = node_tab + nt.size();
    node_list = reinterpret_cast<int4*>(ab + t_bytes);
    edge_list = node_list + nt.size();
    lcount = reinterpret_cast<int*>(ab + t_bytes + l_bytes);
    rstep = reinterpret_cast<RoundStep*>(ab + t_bytes + l_bytes + c_bytes);
    list_at = reinterpret_cast<int*>(ab + t_bytes + l_bytes + c_bytes + r_bytes);
    std::vector<int2> tabs(nt);
    tabs.insert(tabs.end(), et.begin(), et.end());
    REFORM_HIP_CHECK(
        hipMemcpy(node_tab, tabs.data(), sizeof(int2) * tabs.size(), hipMemcpyHostToDevice));
    REFORM_HIP_CHECK(hipMemcpy(first, fi.data(), sizeof(int) * fi.size(), hipMemcpyHostToDevice));
    return 0;
}

int WidePraq::build_lists(hipStream_t stream, int mode) {
    hipLaunchKernelGGL(wpraq_lists, dim3(1), dim3(kPraqListBlock), 0, stream, jobs64, count,
                       first, round_action, mode, rstep, list_at, lcount);
    hipLaunchKernelGGL(wpraq_fill_lists, dim3(count), dim3(256), 0, stream, first, list_at,
                       node_list, edge_list);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

// One step of every instance the mode selects (over the lists build_lists
// made for the mode): the node kernel, then the edge kernel, each bracketed
// by timing marks.
int WidePraq::step(hipStream_t stream, int mode, int j) {
    const dim3 blk(kWideBlock);
    hipEvent_t b, e;
    if (node_blocks) {
        if (mark_events(0, &b, &e)) return kErrorDevice;
        hipExtLaunchKernelGGL(wpraq_node, dim3(node_blocks), blk, 0, stream, b, e, 0, jobs32,
                              jobs64, node_list, lcount, rstep, j);
    }
    if (edge_blocks) {
        if (mark_events(1, &b, &e)) return kErrorDevice;
        hipExtLaunchKernelGGL(wpraq_edge, dim3(edge_blocks), blk, 0, stream, b, e, 0, jobs32,
                              jobs64, edge_list, lcount + 1, rstep, j);
    }
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

int WidePraq::start(hipStream_t stream) {
    const dim3 blk(kWideBlock);
    hipLaunchKernelGGL(wide_init<PraqJob<double>>, dim3(grid_for(count)), blk, 0, stream, jobs64,
                       count, has32 ? 0 : 1);
    hipLaunchKernelGGL(wide_count<PraqJob<double>>, dim3(1), blk, 0, stream, jobs64, count,
                       active);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

int WidePraq::enqueue_round(hipStream_t stream) {
    const int all = node_blocks + edge_blocks;
    const dim3 blk(kWideBlock);
    hipLaunchKernelGGL(wide_pre<PraqJob<double>>, dim3(grid_for(count)), blk, 0, stream, jobs64,
                       count, round_action, pending);
    if (build_lists(stream, 0)) return kErrorDevice;
    hipLaunchKernelGGL(wpraq_copy, dim3(all), blk, 0, stream, jobs32, jobs64, node_tab,
                       node_blocks, edge_tab, round_action, pending, list_at, 0);
    for (int j = 0; j < op_steps; ++j)
        if (step(stream, 0, j)) return kErrorDevice;
    hipLaunchKernelGGL(wpraq_sums, dim3(all), blk, 0, stream, jobs32, jobs64, node_list,
                       node_blocks, edge_list, lcount, rstep, first, partial);
    hipLaunchKernelGGL(wpraq_logic, dim3(count), blk, 0, stream, jobs32, jobs64, first,
                       node_blocks, partial, round_action, 0);
    if (has32) {
        hipLaunchKernelGGL(wpraq_convert, dim3(all), blk, 0, stream, jobs32, jobs64, node_tab,
                           node_blocks, edge_tab);
        hipLaunchKernelGGL(wide_promote<PraqJob<double>>, dim3(grid_for(count)), blk, 0, stream,
                           jobs64, count);
    }
    // restarts: one uncounted vanilla step, the new anchor, the norms
    if (build_lists(stream, 1) || step(stream, 1, 0)) return kErrorDevice;
    hipLaunchKernelGGL(wpraq_copy, dim3(all), blk, 0, stream, jobs32, jobs64, node_tab,
                       node_blocks, edge_tab, round_action, pending, list_at, 1);
    hipLaunchKernelGGL(wpraq_sums, dim3(all), blk, 0, stream, jobs32, jobs64, node_list,
                       node_blocks, edge_list, lcount, rstep, first, partial);
    hipLaunchKernelGGL(wpraq_logic, dim3(count), blk, 0, stream, jobs32, jobs64, first,
                       node_blocks, partial, round_action, 1);
    hipLaunchKernelGGL(wpraq_copy, dim3(all), blk, 0, stream, jobs32, jobs64, node_tab,
                       node_blocks, edge_tab, round_action, pending, list_at, 2);
    hipLaunchKernelGGL(wide_count<PraqJob<double>>, dim3(1), blk, 0, stream, jobs64, count,
                       active);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

int WidePraq::expire(hipStream_t stream, const int* mask) {
    const dim3 blk(kWideBlock);
    hipLaunchKernelGGL(wide_expire<PraqJob<double>>, dim3(grid_for(count)), blk, 0, stream,
                       jobs64, count, mask);
    hipLaunchKernelGGL(wpraq_copy, dim3(node_blocks + edge_blocks), blk, 0, stream, jobs32,
                       jobs64, node_tab, node_blocks, edge_tab, round_action, pending, list_at, 2);
    hipLaunchKernelGGL(wide_count<PraqJob<double>>, dim3(1), blk, 0, stream, jobs64, count,
                       active);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------- PRIA host --

// The FP32 x kernel of the table's shape.
const void* WidePria::x32_kernel() const {
    if (x32_threads == 512)
        return x32_hold ? reinterpret_cast<const void*>(
                              wpria_x<float, kWideTW, x_hold_batches(512), 512>)
                        : reinterpret_cast<const void*>(wpria_x<float, kWideTW, 0, 512>);
    return x32_hold
               ? reinterpret_cast<const void*>(wpria_x<float, kWideTW, x_hold_batches(256), 256>)
               : reinterpret_cast<const void*>(wpria_x<float, kWideTW, 0, 256>);
}

int WidePria::setup(std::vector<PriaJob<double>> host64, const PriaJob<float>* dev32,
                    const PriaJob<double>* dev64) {
    // the tuning table's shapes, fixed for the engine's life
    x32_threads = pria_knob(kPriaXThreads) == 512 ? 512 : 256;
    x32_hold_knob = pria_knob(kPriaXHold) != 2;
    yz_compact_knob = pria_knob(kPriaYZShape) != 2;
    full64 = std::move(host64);
    full_dev32 = dev32;
    full_dev64 = dev64;
    full_count = int(full64.size());
    table_index.clear();
    return build_tables(full64, dev32, dev64);
}

int WidePria::shrink(hipStream_t stream, const std::vector<int>& keep) {
    std::vector<PriaJob<double>> host64;
    if (gather_subset(subset, full64, full_dev32, full_dev64, keep, stream, host64))
        return kErrorDevice;
    table_index = keep;
    if (build_tables(host64, full_dev32 ? subset.jobs32 : nullptr, subset.jobs64)) return kErrorDevice;
    // instances only ever go from FP32 to FP64: a subset with none left in
    // FP32 needs no FP32 step launches
    step32 = has32 && keep_has32;
    return 0;
}

int WidePria::restore() {
    table_index.clear();
    return build_tables(full64, full_dev32, full_dev64);
}

int WidePria::live_flags(hipStream_t stream) {
    hipLaunchKernelGGL(wide_live<PriaJob<double>>, dim3(grid_for(count)), dim3(kWideBlock), 0,
                       stream, jobs64, count, live);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

int WidePria::build_tables(const std::vector<PriaJob<double>>& host64,
                           const PriaJob<float>* dev32, const PriaJob<double>* dev64) {
    constexpr int TW = kWideTW;
    count = int(host64.size());
    jobs32 = dev32;
    jobs64 = dev64;
    has32 = dev32 != nullptr;
    step32 = has32;
    if (!count) {
        release();
        return 0;
    }
    std::vector<int4> xt;
    std::vector<int2> st;
    std::vector<int> fi(4 * size_t(count)), xr(2 * size_t(count));
    op_steps = 2;
    max_T = 2;
    int tile_rows = 0;
    // algorithmic bytes of one step per entry, b = bytes per value: x kernel
    // reads x, x0, h, y, z and writes x, r (7 b); y/z kernel reads y, y0, r,
    // z, z0, r and writes y, z (8 b)
    for (int k = 0; k < kKinds; ++k)
        for (int p = 0; p < 2; ++p) step_bytes[k][p].assign(count, 0.0);
    for (int q = 0; q < count; ++q) {
        const auto& a = host64[q].a;
        op_steps = std::max(op_steps, int(host64[q].p.op_steps));
        max_T = std::max(max_T, a.T);
        const int tiles = (a.T + TW - 1) / TW;
        fi[4 * q] = int(st.size());
        fi[4 * q + 1] = a.T;
        fi[4 * q + 2] = tile_rows;
        fi[4 * q + 3] = a.N * tiles;
        for (int s = 0; s < a.T; ++s) st.push_back(make_int2(q, s));
        // Items in groups of kXcds; a group's entries by tile, heaviest
        // (largest t) first, then by item, so an item's tiles are kXcds
        // entries apart.  Launches deal workgroups to the XCDs round robin
        // and every range starts at a multiple of kXcds in its list, so all
        // tiles of an item run on one XCD at about the same time, and the
        // 128-byte lines that neighbouring tiles' rows share are fetched
        // into that XCD's L2 once.  A short last group is padded with
        // entries of item -1, which the kernels skip.  The check row is
        // (item, tile).
        xr[2 * q] = int(xt.size());
        for (int g = 0; g < a.N; g += kXcds)
            for (int k = tiles - 1; k >= 0; --k)
                for (int i = g; i < g + kXcds; ++i)
                    xt.push_back(i < a.N ? make_int4(q, i, k * TW, tile_rows + i * tiles + k)
                                         : make_int4(q, -1, 0, -1));
        xr[2 * q + 1] = int(xt.size()) - xr[2 * q];
        tile_rows += a.N * tiles;
        for (int p = 0; p < 2; ++p) {
            const double b = p ? 8.0 : 4.0, nd = a.ND;
            step_bytes[0][p][q] = 7.0 * b * nd;
            step_bytes[1][p][q] = 8.0 * b * nd;
        }
    }
    x_blocks = int(xt.size());
    s_blocks = int(st.size());
    // arena: x table, slice table, the x lists (two of x_blocks entries),
    // the slice lists (two of s_blocks), their lengths, the instances'
    // RoundStep records, list offsets and x ranges
    const size_t x_bytes = align256(sizeof(int4) * xt.size());
    const size_t s_bytes = align256(sizeof(int2) * st.size());
    const size_t lx_bytes = align256(sizeof(int4) * 2 * size_t(x_blocks));
    const size_t ls_bytes = align256(sizeof(int2) * 2 * size_t(s_blocks));
    const size_t c_bytes = align256(sizeof(int) * kLists);
    const size_t r_bytes = align256(sizeof(RoundStep) * size_t(count));
    const size_t a_bytes = align256(sizeof(int) * 3 * size_t(count));
    size_t o = 0;
    const size_t o_x = o, o_s = (o += x_bytes), o_lx = (o += s_bytes), o_ls = (o += lx_bytes),
                 o_c = (o += ls_bytes), o_r = (o += c_bytes), o_a = (o += r_bytes),
                 o_xr = (o += a_bytes);
    const size_t tab_bytes = o_xr + sizeof(int) * xr.size();
    if (allocate(tab_bytes, xt.size() + st.size(), kPriaSums)) return kErrorDevice;
    char* ab = static_cast<char*>(arena);
    x_tab = reinterpret_cast<int4*>(ab + o_x);
    s_tab = reinterpret_cast<int2*>(ab + o_s);
    xlists = reinterpret_cast<int4*>(ab + o_lx);
    slists = reinterpret_cast<int2*>(ab + o_ls);
    lcount = reinterpret_cast<int*>(ab + o_c);
    rstep = reinterpret_cast<RoundStep*>(ab + o_r);
    list_at = reinterpret_cast<int*>(ab + o_a);
    x_range = reinterpret_cast<int*>(ab + o_xr);
    REFORM_HIP_CHECK(hipMemcpy(x_range, xr.data(), sizeof(int) * xr.size(), hipMemcpyHostToDevice));
    REFORM_HIP_CHECK(hipMemcpy(x_tab, xt.data(), sizeof(int4) * xt.size(), hipMemcpyHostToDevice));
    REFORM_HIP_CHECK(hipMemcpy(s_tab, st.data(), sizeof(int2) * st.size(), hipMemcpyHostToDevice));
    REFORM_HIP_CHECK(hipMemcpy(first, fi.data(), sizeof(int) * fi.size(), hipMemcpyHostToDevice));
    // LDS of an x tile: FP32 32 columns; FP64 32 columns when T' x 33 x 8
    // bytes fit two tiles to a CU, else two halves of 16
    lds32 = sizeof(float) * (size_t(max_T) * (TW + 1) + TW);
    lds64 = sizeof(double) * (size_t(max_T) * (TW + 1) + TW);
    x64_full = lds64 <= kPriaX64MaxLds;
    if (!x64_full) lds64 = sizeof(double) * (size_t(max_T) * (TW / 2 + 1) + TW / 2);
    x32_hold = x32_hold_knob && max_T <= kPriaCompactT;
    compact = yz_compact_knob && max_T <= kPriaCompactT;
    const void* x32 = x32_kernel();
    const void* x64 = x64_full
                          ? reinterpret_cast<const void*>(wpria_x<double, kWideTW>)
                          : reinterpret_cast<const void*>(wpria_x<double, kWideTW / 2, 0, 256>);
    if (lds32 > 65536)
        REFORM_HIP_CHECK(hipFuncSetAttribute(x32, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             int(lds32)));
    if (lds64 > 65536)
        REFORM_HIP_CHECK(hipFuncSetAttribute(x64, hipFuncAttributeMaxDynamicSharedMemorySize,
                                             int(lds64)));
    return 0;
}

int WidePria::build_lists(hipStream_t stream, int mode) {
    hipLaunchKernelGGL(wpria_lists, dim3(1), dim3(kListBlock), 0, stream, jobs64, count, first,
                       x_range, round_action, mode, rstep, list_at, lcount);
    hipLaunchKernelGGL(wpria_fill_lists, dim3(count), dim3(256), 0, stream, first, x_range,
                       list_at, x_tab, s_tab, xlists, slists, x_blocks, s_blocks);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

// One step of every instance the mode selects: the x kernels, then the y/z
// kernels, over the round's work lists (built by build_lists for the mode),
// each pair bracketed by timing events (the first kernel's start, the last
// kernel's stop).
int WidePria::step(hipStream_t stream, int mode, int j) {
    const int4* lx32 = xlists;
    const int4* lx64 = xlists + x_blocks;
    const int2* ls32 = slists;
    const int2* ls64 = slists + s_blocks;
    // FP32 launches while the tables may hold an instance in FP32 (batches
    // with FP32 images, until a compaction finds every instance in FP64);
    // every batch can hold FP64 instances (FP64 batches, and FP32 ones
    // after the handover)
    hipEvent_t b, e;
    if (mark_events(0, &b, &e)) return kErrorDevice;
    if (step32) {
        const int* lcount_x32 = lcount + kListX32;
        const RoundStep* rs = rstep;
        void* args[] = {&jobs32, &jobs64, &lx32, &lcount_x32, &rs, &j};
        REFORM_HIP_CHECK(hipExtLaunchKernel(x32_kernel(), dim3(x_blocks), dim3(x32_threads), args,
                                            lds32, stream, b, nullptr, 0));
    }
    hipEvent_t b64 = step32 ? nullptr : b;
    if (x64_full)
        hipExtLaunchKernelGGL((wpria_x<double, kWideTW>), dim3(x_blocks), dim3(512), lds64,
                              stream, b64, e, 0, jobs32, jobs64, lx64, lcount + kListX64,
                              rstep, j);
    else   // half tiles in 256 threads: 16 lanes per column, as the full tile's
        hipExtLaunchKernelGGL((wpria_x<double, kWideTW / 2, 0, 256>), dim3(2 * x_blocks),
                              dim3(256), lds64, stream, b64, e, 0, jobs32, jobs64, lx64,
                              lcount + kListX64, rstep, j);
    if (mark_events(1, &b, &e)) return kErrorDevice;
    auto yz = [&](auto real, const int2* list, int l, hipEvent_t start, hipEvent_t stop) {
        using Real = decltype(real);
        if (compact)
            hipExtLaunchKernelGGL((wpria_yz<Real, true>), dim3(s_blocks), dim3(512),
                                  kPriaZCacheBytes, stream, start, stop, 0, jobs32, jobs64, list,
                                  lcount + l, rstep, j);
        else
            hipExtLaunchKernelGGL((wpria_yz<Real, false>), dim3(s_blocks), dim3(512),
                                  kPriaZCacheBytes, stream, start, stop, 0, jobs32, jobs64, list,
                                  lcount + l, rstep, j);
    };
    if (step32) yz(float(0), ls32, kListS32, b, nullptr);
    yz(double(0), ls64, kListS64, step32 ? nullptr : b, e);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

int WidePria::start(hipStream_t stream) {
    const dim3 blk(kWideBlock);
    hipLaunchKernelGGL(wide_init<PriaJob<double>>, dim3(grid_for(count)), blk, 0, stream, jobs64,
                       count, has32 ? 0 : 1);
    hipLaunchKernelGGL(wide_count<PriaJob<double>>, dim3(1), blk, 0, stream, jobs64, count,
                       active);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

int WidePria::enqueue_round(hipStream_t stream) {
    const dim3 blk(kWideBlock);
    hipLaunchKernelGGL(wide_pre<PriaJob<double>>, dim3(grid_for(count)), blk, 0, stream, jobs64,
                       count, round_action, pending);
    if (build_lists(stream, 0)) return kErrorDevice;
    hipLaunchKernelGGL(wpria_copy, dim3(s_blocks), blk, 0, stream, jobs32, jobs64, s_tab,
                       round_action, pending, list_at, 0);
    for (int j = 0; j < op_steps; ++j)
        if (step(stream, 0, j)) return kErrorDevice;
    hipLaunchKernelGGL(wpria_sums, dim3(s_blocks + x_blocks), blk, 0, stream, jobs32, jobs64,
                       slists, s_blocks, xlists, x_blocks, lcount, rstep, first, partial);
    hipLaunchKernelGGL(wpria_logic, dim3(count), blk, 0, stream, jobs32, jobs64, first, s_blocks,
                       partial, round_action, 0);
    if (has32) {
        hipLaunchKernelGGL(wpria_convert, dim3(s_blocks), blk, 0, stream, jobs32, jobs64, s_tab);
        hipLaunchKernelGGL(wide_promote<PriaJob<double>>, dim3(grid_for(count)), blk, 0, stream,
                           jobs64, count);
    }
    if (build_lists(stream, 1) || step(stream, 1, 0)) return kErrorDevice;
    hipLaunchKernelGGL(wpria_copy, dim3(s_blocks), blk, 0, stream, jobs32, jobs64, s_tab,
                       round_action, pending, list_at, 1);
    hipLaunchKernelGGL(wpria_sums, dim3(s_blocks + x_blocks), blk, 0, stream, jobs32, jobs64,
                       slists, s_blocks, xlists, x_blocks, lcount, rstep, first, partial);
    hipLaunchKernelGGL(wpria_logic, dim3(count), blk, 0, stream, jobs32, jobs64, first, s_blocks,
                       partial, round_action, 1);
    hipLaunchKernelGGL(wpria_copy, dim3(s_blocks), blk, 0, stream, jobs32, jobs64, s_tab,
                       round_action, pending, list_at, 2);
    hipLaunchKernelGGL(wide_count<PriaJob<double>>, dim3(1), blk, 0, stream, jobs64, count,
                       active);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

int WidePria::expire(hipStream_t stream, const int* mask) {
    const dim3 blk(kWideBlock);
    hipLaunchKernelGGL(wide_expire<PriaJob<double>>, dim3(grid_for(count)), blk, 0, stream,
                       jobs64, count, mask);
    hipLaunchKernelGGL(wpria_copy, dim3(s_blocks), blk, 0, stream, jobs32, jobs64, s_tab,
                       round_action, pending, list_at, 2);
    hipLaunchKernelGGL(wide_count<PriaJob<double>>, dim3(1), blk, 0, stream, jobs64, count,
                       active);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace reform
