// The batched device tier: many instances of one family advanced together by
// wide kernels, each instance in its own precision.
//
// Every launch covers the work of every instance: a block table maps each
// block to (instance, local work unit) -- PRAQ: 256 node lanes or 256 edges;
// PRIA: one x tile (item, column range) or one slice -- and each block reads
// its instance's scalar state to decide what to do.  Instances therefore sit
// at different points of the Halpern state machine in the same launch: some
// take reflected steps, some vanilla warm-phase steps, some a restart step,
// finished ones skip.  Each block also reads its instance's precision and
// runs the FP32 or the FP64 body on that instance's FP32 or FP64 image.  One
// round is
//
//   pre           per instance: the round's action, pending anchor copies
//   copies        start copies of instances starting, anchors where pending
//   op_steps x    the step kernels (node + edge / x + y-z)
//   check sums    per block partial rows of the check sums
//   logic         one workgroup per instance: fixed-order reduction of its
//                 rows, then the scalar logic of halpern.hpp on one thread
//   handover      FP32 instances whose logic asked for FP64 (gap below the
//                 switch gap, or stalled): state converted, precision flipped
//   restart       the step kernels again, for instances restarting (one
//                 uncounted vanilla step), anchor copy, norm sums, logic
//   finalize      (sx, sy) -> (x, y) where the answer is the step point;
//                 count the instances still active
//
// and the rounds are captured once into a hipGraph and replayed until no
// instance is active.  A round launches the largest check interval's worth
// of step kernels; an instance with a shorter interval skips the rest.  Per
// instance the arithmetic is that of the device tier (same device functions,
// summation order fixed by the instance's own blocks), so an instance's
// result does not depend on its batch, and its precision switch happens at
// the same check as in the plain solver.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <vector>

#include "praq_host.hpp"
#include "pria_host.hpp"

namespace reform {

// Actions beyond halpern.hpp's: kActNew marks an instance admitted but not
// started; its first round is kActInit (the start copies, the uncounted
// first vanilla step, its check); kActSkip is a round that does nothing.
constexpr int kActInit = 5;
constexpr int kActNew = 6;
constexpr int kActSkip = -1;

// XCDs of the device: a launch's workgroups go to them round robin.
constexpr int kXcds = 8;

// x tile width of the batched PRIA kernels, both precisions (one tile table
// serves an instance before and after its switch).
constexpr int kWideTW = 32;
// LDS budget of one FP64 x tile of 32 columns: two such tiles share a CU
constexpr size_t kPriaX64MaxLds = 80 * 1024;

struct WideStats {
    int rounds = 0;
    int shrinks = 0;
    double seconds = 0.0;
};

// What every wide engine holds besides its block tables: per-instance round
// control, the graph of one round and the host side of the active count.
// The instances' lasting state -- Halpern scalars, precision, last action,
// step counts -- lives in their LogicState records, so the tables can be
// rebuilt for a new set of instances (setup() again) between rounds.
struct WideControl {
    int count = 0, op_steps = 2;
    int* first = nullptr;          // [4 count]: two (first row, rows) ranges
    double* partial = nullptr;     // partial check-sum rows
    int* round_action = nullptr;   // [count] the action this round takes
    int* pending = nullptr;        // [count] anchor copy due this round
    int* active = nullptr;         // [1] instances not finished
    int* expire_mask = nullptr;    // [count] instances stopped by their own time limit
    int* live = nullptr;           // [count] instances not finished (for shrink)
    int* host_active = nullptr;    // [2] pinned
    hipEvent_t polled[2] = {nullptr, nullptr};
    void* arena = nullptr;
    size_t bytes = 0;
    bool has32 = false;            // FP32 images exist (FP32 iterates first)
    bool keep_has32 = true;        // the last shrink kept an instance still in FP32
    bool use_graph = true;         // replay one captured round (else launch directly)
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;

    // Launch timing of the two step kernels of the family (PRAQ node, edge;
    // PRIA x, yz): with `timing` on, rounds are launched directly rather
    // than replayed, every step launch bracketed by a pair of HIP events on
    // its stream; two event sets alternate so one is read while the other's
    // round runs.
    static constexpr int kKinds = 2;
    bool timing = false;
    std::vector<hipEvent_t> marks[3];   // round sets 0, 1; set 2: the start
    std::vector<int> mark_kinds[3];     // kind of each (begin, end) pair
    int marks_used[3] = {0, 0, 0};
    int mark_set = 0;
    double kernel_ms[kKinds] = {0.0, 0.0};
    int64_t launches[kKinds] = {0, 0};
    // algorithmic bytes of one step per instance: [kind][precision][instance]
    std::vector<double> step_bytes[kKinds][2];
    // Per-instance time limits, by full-table position: the clock reading
    // (seconds, on the owner's clock) past which the instance is stopped at
    // its current point; +inf for none, empty when no instance has a limit.
    // `expired` latches the instances already stopped.
    std::vector<double> deadlines;
    std::vector<int> expired;
    std::vector<int> mask_host;    // the latches in table order, copied to expire_mask

    // Compaction.  setup() builds the tables over a full set of instances;
    // shrink() rebuilds them over a subset (full-table positions, ascending)
    // between rounds, restore() over the full set again.  An instance's
    // arithmetic depends only on its own blocks, so its results do not
    // change.  table_index[q] is the full-table position of table position q.
    int full_count = 0;
    std::vector<int> table_index;
    bool shrunk() const { return count != full_count; }
    virtual int shrink(hipStream_t stream, const std::vector<int>& keep) = 0;
    virtual int restore() = 0;
    // live[q] = 1 for the table positions not finished (on the stream)
    virtual int live_flags(hipStream_t stream) = 0;
    // the table positions still active -> shrink (after the stream has drained)
    int shrink_to_live(hipStream_t stream);

    virtual ~WideControl();
    // every instance of the tables cold-starts at its next round (states set
    // to kActNew); then the first active count
    virtual int start(hipStream_t stream) = 0;
    // one round's launches (captured into the graph)
    virtual int enqueue_round(hipStream_t stream) = 0;
    // out of time: every active instance (mask null) or those with a
    // nonzero mask entry (device, [count]) end at their current point
    virtual int expire(hipStream_t stream, const int* mask = nullptr) = 0;
    // the instances whose deadline lies before `now` and that have not been
    // stopped yet: stopped (expire with the latched mask) on the stream;
    // returns < 0 on error, else how many
    int expire_due(hipStream_t stream, double now);
    // the instance at full-table position f stops at the next expire_due, as
    // if its deadline had passed (a cancelled request); with no work of this
    // engine's in flight
    void stop_position(int f);
    // set the deadlines (full-table order) and clear the latches; with no
    // work of this engine's in flight
    void set_deadlines(std::vector<double> d);
    // one round: the captured graph replayed (captured at the first call
    // after a setup), or launched directly when use_graph is off
    int launch_round(hipStream_t stream);
    // start, then rounds until no instance is active
    int solve(hipStream_t stream, double time_limit, WideStats* stats);
    // arena: the caller's tables (table_bytes) followed by the control arrays;
    // drops the previous tables and graph
    int allocate(size_t table_bytes, size_t partial_rows, int partial_width);
    void release();
    // the events that bracket one step launch of `kind`, null unless timing:
    // *begin goes to its first kernel's start, *end to its last kernel's
    // stop (hipExtLaunchKernelGGL), so the interval is the kernels' own
    int mark_events(int kind, hipEvent_t* begin, hipEvent_t* end);
    // add the event pairs of a completed set to kernel_ms / launches
    int collect(int set);
    // turn timing on or off (the totals of the last solve stay readable)
    void set_timing(bool on);
    void clear_timing();
    // algorithmic bytes moved by each step kernel, from the instances' step
    // counts (states in table order)
    void algorithmic_bytes(const std::vector<LogicState>& states, double (&out)[kKinds]) const;
};

// One engine's solve advanced in steps, so that one host thread can keep
// several engines on different streams busy: at most two groups of rounds
// are queued, each followed by a copy of the active count into its pinned
// slot and an event; advance() retires the groups whose event has
// completed, without blocking, and queues more while instances remain.
struct WideRun {
    WideControl* w = nullptr;
    hipStream_t stream = nullptr;
    double time_limit = 0.0;
    bool finished = true;
    bool shrink = true;            // compact the tables as instances finish
    int queued = 0, retired = 0;   // groups of rounds
    std::chrono::steady_clock::time_point t0;
    WideStats stats;

    int begin(WideControl* engine, hipStream_t s, double limit);
    // returns < 0 on error, else whether anything changed
    int advance();
};

// Device arrays for a subset's job records (gathered from the full set's by
// position) and the positions themselves; grown on demand, freed with the
// engine.
template <class J32, class J64>
struct WideSubset {
    void* mem = nullptr;
    int cap = 0;
    J32* jobs32 = nullptr;
    J64* jobs64 = nullptr;
    int* keep = nullptr;
    ~WideSubset() {
        if (mem) (void)hipFree(mem);
    }
    int reserve(int n);
};

// What a step block needs of its instance besides the job record's arrays,
// for the steps a work-list build enqueues: mode 0 the round's steps, mode
// 1 the restart step.
struct RoundStep {
    int64_t r0;       // restart counter at the round's first step
    double omega, eta;
    int act;          // the round's action (mode 0), kActRestart (mode 1)
    int steps;        // steps taken: op_steps, 1 for kActInit and a restart
};

// PRAQ: first = (node block range, edge block range) per instance; partial
// rows (kPraqSums each) per node block, then per edge block.
struct WidePraq : WideControl {
    const PraqJob<float>* jobs32 = nullptr;   // device, launch order (null: FP64 batch)
    const PraqJob<double>* jobs64 = nullptr;
    int2* node_tab = nullptr;                 // (instance, local block) per node block
    int2* edge_tab = nullptr;
    int node_blocks = 0, edge_blocks = 0;
    // the step kernels' work lists of the round: (instance, local block,
    // precision) of the node and edge blocks of instances taking steps,
    // their lengths, the instances' step scalars and list offsets
    int4* node_list = nullptr;
    int4* edge_list = nullptr;
    int* lcount = nullptr;
    RoundStep* rstep = nullptr;
    int* list_at = nullptr;

    // the full set's records: host copies, device arrays as setup() got them,
    // and the subset arrays shrink() gathers into
    std::vector<PraqJob<double>> full64;
    const PraqJob<float>* full_dev32 = nullptr;
    const PraqJob<double>* full_dev64 = nullptr;
    WideSubset<PraqJob<float>, PraqJob<double>> subset;

    // tables for the instances of host64 (device records dev32 / dev64 in the
    // same order; dev32 null when FP32 images do not exist)
    int setup(std::vector<PraqJob<double>> host64, const PraqJob<float>* dev32,
              const PraqJob<double>* dev64);
    int step(hipStream_t stream, int mode, int j);
    // the work lists of mode 0 (steps of the round) or 1 (restart steps)
    int build_lists(hipStream_t stream, int mode);
    int start(hipStream_t stream) override;
    int enqueue_round(hipStream_t stream) override;
    int expire(hipStream_t stream, const int* mask = nullptr) override;
    int shrink(hipStream_t stream, const std::vector<int>& keep) override;
    int restore() override;
    int live_flags(hipStream_t stream) override;

  private:
    int build_tables(const std::vector<PraqJob<double>>& host64, const PraqJob<float>* dev32,
                     const PraqJob<double>* dev64);
};

// PRIA: first = (slice range, column-tile range) per instance; partial rows
// (kPriaSums wide) per slice, then one per column tile in (item, tile) order.
struct WidePria : WideControl {
    int max_T = 2;
    const PriaJob<float>* jobs32 = nullptr;
    const PriaJob<double>* jobs64 = nullptr;
    int4* x_tab = nullptr;              // (instance, item, t0, column-tile row) per x tile
    int2* s_tab = nullptr;              // (instance, slice) per slice
    int x_blocks = 0, s_blocks = 0;
    size_t lds32 = 0, lds64 = 0;        // dynamic LDS of the x kernels
    bool x64_full = true;               // FP64 x tiles of 32 columns (else two halves)
    // shapes from the tuning table (read by setup): FP32 x-tile threads,
    // whether FP32 x tiles may hold x1 in registers, whether the y/z pass may
    // take the compact register shape; and what the tables allow of them
    int x32_threads = 256;
    bool x32_hold_knob = true, yz_compact_knob = true;
    bool x32_hold = false, compact = false;
    bool step32 = false;   // FP32 step launches: FP32 images, some instance still in FP32
    // work lists of the round: table entries of x tiles (FP32, then FP64)
    // and of slices (FP32, then FP64), their lengths, the listed instances'
    // step scalars and their ranges' list offsets
    int4* xlists = nullptr;
    int2* slists = nullptr;
    int* lcount = nullptr;
    RoundStep* rstep = nullptr;
    int* list_at = nullptr;
    int* x_range = nullptr;             // [2 count]: (first x-table entry, entries)

    std::vector<PriaJob<double>> full64;
    const PriaJob<float>* full_dev32 = nullptr;
    const PriaJob<double>* full_dev64 = nullptr;
    WideSubset<PriaJob<float>, PriaJob<double>> subset;

    int setup(std::vector<PriaJob<double>> host64, const PriaJob<float>* dev32,
              const PriaJob<double>* dev64);
    int step(hipStream_t stream, int mode, int j);
    const void* x32_kernel() const;
    // the work lists of mode 0 (steps of the round) or 1 (restart steps)
    int build_lists(hipStream_t stream, int mode);
    int start(hipStream_t stream) override;
    int enqueue_round(hipStream_t stream) override;
    int expire(hipStream_t stream, const int* mask = nullptr) override;
    int shrink(hipStream_t stream, const std::vector<int>& keep) override;
    int restore() override;
    int live_flags(hipStream_t stream) override;

  private:
    int build_tables(const std::vector<PriaJob<double>>& host64, const PriaJob<float>* dev32,
                     const PriaJob<double>* dev64);
};

}  // namespace reform
