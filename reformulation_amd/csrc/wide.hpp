// The batched device tier: many instances of one family advanced together by
// wide kernels.
//
// Every launch covers the work of every instance: a block table maps each
// block to (instance, local work unit) -- PRAQ: 256 node lanes or 256 edges;
// PRIA: one x tile (item, column range) or one slice -- and each block reads
// its instance's scalar state to decide what to do.  Instances therefore sit
// at different points of the Halpern state machine in the same launch: some
// take reflected steps, some vanilla warm-phase steps, some a restart step,
// finished and stalled ones skip.  One round is
//
//   pre           per instance: the round's action, pending anchor copies
//   anchor copy   x0 <- x, y0 <- y where pending
//   op_steps x    the step kernels (node + edge / x + y-z)
//   check sums    per block partial rows of the check sums
//   logic         one workgroup per instance: fixed-order reduction of its
//                 rows, then the scalar logic of halpern.hpp on one thread
//   restart       the step kernels again, for instances restarting (one
//                 uncounted vanilla step), anchor copy, norm sums, logic
//   finalize      (sx, sy) -> (x, y) where the answer is the step point;
//                 count the instances still active
//
// and the rounds are captured once into a hipGraph and replayed until no
// instance is active.  A round launches the largest check interval's worth
// of step kernels; an instance with a shorter interval skips the rest.  Per instance the arithmetic is that of the device
// tier (same device functions, same per-instance summation order of the
// check sums), so an instance's result does not depend on its batch.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "praq_host.hpp"
#include "pria_host.hpp"

namespace reform {

// Round actions beyond halpern.hpp's: the uncounted first vanilla step of a
// cold start.
constexpr int kActInit = 5;
constexpr int kActSkip = -1;

struct WideStats {
    int rounds = 0;
    double seconds = 0.0;
};

// What every wide engine holds besides its block tables: per-instance round
// control, the graph of one round and the host side of the active count.
struct WideControl {
    int count = 0, op_steps = 2;
    int* first = nullptr;          // [4 count]: two (first row, rows) ranges
    double* partial = nullptr;     // partial check-sum rows
    int* round_action = nullptr;   // [count] the action this round takes
    int* last_action = nullptr;    // [count] the last action executed
    int* pending = nullptr;        // [count] anchor copy due this round
    int* active = nullptr;         // [1] instances neither finished nor stalled
    int* host_active = nullptr;    // [2] pinned
    hipEvent_t polled[2] = {nullptr, nullptr};
    void* arena = nullptr;
    size_t bytes = 0;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;

    ~WideControl();
    // arena: the caller's tables (table_bytes) followed by the control arrays
    int allocate(size_t table_bytes, size_t partial_rows, int partial_width);
};

// PRAQ: first = (node block range, edge block range) per instance; partial
// rows (kPraqSums each) per node block, then per edge block.
template <class Real>
struct WidePraq : WideControl {
    PraqJob<Real>* jobs = nullptr;      // device, launch order
    int2* node_tab = nullptr;           // (instance, local block) per node block
    int2* edge_tab = nullptr;
    int node_blocks = 0, edge_blocks = 0;

    int setup(const std::vector<PraqJob<Real>>& host_jobs, PraqJob<Real>* dev_jobs);
    int start(hipStream_t stream, bool resume);
    int enqueue_round(hipStream_t stream);
    int expire(hipStream_t stream);
    int solve(hipStream_t stream, bool resume, double time_limit, WideStats* stats);
};

// PRIA: first = (slice range, column-tile range) per instance; partial rows
// (kPriaSums wide) per slice, then one per column tile in (item, tile) order.
template <class Real>
struct WidePria : WideControl {
    int max_T = 2;
    PriaJob<Real>* jobs = nullptr;
    int4* x_tab = nullptr;              // (instance, item, t0, column-tile row) per x tile
    int2* s_tab = nullptr;              // (instance, slice) per slice
    int x_blocks = 0, s_blocks = 0;

    int setup(const std::vector<PriaJob<Real>>& host_jobs, PriaJob<Real>* dev_jobs);
    int start(hipStream_t stream, bool resume);
    int enqueue_round(hipStream_t stream);
    int expire(hipStream_t stream);
    int solve(hipStream_t stream, bool resume, double time_limit, WideStats* stats);
};

}  // namespace reform
