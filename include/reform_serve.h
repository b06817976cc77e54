/* Continuous serving of PRAQ and PRIA solves on one MI355X.
 *
 * No counterpart in the reference (its solvers take one instance per call,
 * qubo/praq-cuda/src/praq.cu:35-79, jrp/pria-cuda/src/pria.cu:15-44).  A
 * server owns a device memory pool (the HBM budget) and a serving thread:
 * requests submitted from any thread are built on host threads, admitted as
 * soon as their images fit in the pool, join the batched tier's rounds (each
 * instance with its own Halpern state, precision and termination), and are
 * retired as they converge -- their slots and memory backfilled by the next
 * admissions.  Every request is solved with the same arithmetic as alone in a
 * batch, so its result does not depend on what else was in flight.
 *
 * Requests are admitted in submission order; one whose images exceed the
 * whole budget is rejected (status -2).  On an FP32 server a request takes
 * pool memory for its FP32 image at admission and for its FP64 image only
 * at its handover (one that converges in FP32 never does); a request at
 * its handover waits between rounds, its state intact, while the pool has
 * no room.  Admission keeps the pool's largest free block at least the
 * largest FP64 image a request in flight may still ask for, so one waiting
 * handover can always go ahead and every request finishes.  A request's options.time_limit
 * (seconds, > 0) counts from its admission: past it the request stops at its
 * current point between two rounds and finishes with status 2, as the plain
 * solvers do at their limits.
 *
 * A request whose inputs hold a NaN or an infinity is rejected (status -2)
 * by its builder: the plain solvers iterate on such an instance to their
 * iteration cap (qubo/praq-jax/praq/solver.py:375-381), which on a server
 * would hold pool memory for nothing.  reform_server_cancel stops a request:
 * one not yet admitted finishes at once with status -6 and no answer; one in
 * flight stops at its current point at its family's next round, as at a time
 * limit, and finishes with status 3 and that point as its answer; one that
 * converges or reaches a limit first keeps that status.
 *
 * Return codes: >= 0 success (submit returns the ticket), -1 argument,
 * -2 instance, -3 device, -4 unknown or unfinished ticket, -5 timeout.
 */
#ifndef REFORMULATION_REFORM_SERVE_H
#define REFORMULATION_REFORM_SERVE_H

#include <stdint.h>

#include "praq_hip.h"
#include "pria_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct reform_server reform_server;

typedef struct reform_server_config {
    int use_float32;        /* FP32 iterates (FP64 reductions, FP64 finish), else FP64 */
    int host_threads;       /* host builders of instance images (0: 4) */
    int64_t hbm_budget;     /* bytes of device memory for instance images (0: 200e9) */
    int64_t max_in_flight;  /* instances on the device at once (0: 16384) */
    int borrow_inputs;      /* 1: the server reads the caller's input arrays, which stay
                               valid and unchanged until the request has finished;
                               0: they are copied at submission */
    int build_ahead;        /* built requests waiting for device memory, at most (0: 32) */
} reform_server_config;

typedef struct reform_request_info {
    int64_t ticket;
    int family;             /* 0 PRAQ, 1 PRIA */
    int status;             /* 0 converged, 1 iteration limit, 2 time limit,
                               3 cancelled in flight, -1 not finished,
                               -2 instance rejected, -3 device,
                               -6 cancelled before admission (no answer) */
    int64_t iterations;
    int64_t restarts;
    int64_t switch_iteration;   /* FP32 -> FP64 handover, -1 if none */
    double lower, upper, relative_gap;
    double submitted, admitted, completed;  /* seconds since the server started */
    int64_t device_bytes;   /* pool bytes the request's images took at the end */
    double bracket;         /* upper - lower: bounds the answer's objective error
                               (NaN until finished) */
} reform_request_info;

typedef struct reform_server_counters {
    int64_t submitted, admitted, completed, rejected;
    int64_t rounds, rebuilds;
    int64_t in_flight;
    int64_t bytes_in_use, bytes_peak, bytes_budget;
    double retire_seconds;  /* retire threads (summed): copying answers out, freeing */
    double rebuild_seconds; /* serving thread: rebuilding the block tables */
    int64_t in_flight_peak; /* most requests on the device at once */
    int64_t handovers;      /* FP64 images placed at FP32 requests' handovers */
    int64_t handover_waits; /* rounds a request at its handover waited for room */
    int64_t cancelled;      /* requests finished by a cancellation (status 3 or -6) */
} reform_server_counters;

reform_server_config reform_server_default_config(void);
reform_server* reform_server_create(const reform_server_config* config);
/* Stops the serving thread (requests still in flight are dropped). */
void reform_server_destroy(reform_server* server);

/* Copies the instance and returns its ticket at once; the work happens on the
 * server's threads. */
int64_t reform_server_submit_praq(reform_server* server, const PraqInstance* instance,
                                  const PraqParameters* parameters, const PraqOptions* options);
int64_t reform_server_submit_pria(reform_server* server, int64_t n, int64_t t,
                                  const double* order_costs, double joint_cost,
                                  const double* holding_costs, const pria_params* params,
                                  const pria_options* options);

/* Tickets finished since the last poll (at most max); returns the count. */
int64_t reform_server_poll(reform_server* server, int64_t* tickets, int64_t max);
/* Blocks until every submitted request has finished, or the timeout. */
int reform_server_wait_idle(reform_server* server, double timeout_seconds);

int reform_server_info(reform_server* server, int64_t ticket, reform_request_info* info);
/* Results of a finished request, shapes as reform_batch_*_result.  Outputs
 * may be NULL. */
int reform_server_praq_result(reform_server* server, int64_t ticket, double* x, double* y);
int reform_server_pria_result(reform_server* server, int64_t ticket, double* x, double* y,
                              double* z);
/* Cancels an unfinished request (see above); -4 if the ticket is unknown,
 * finished or already cancelled.  Its record stays until released. */
int reform_server_cancel(reform_server* server, int64_t ticket);
/* Drops a finished request's record and results. */
int reform_server_release(reform_server* server, int64_t ticket);
int reform_server_stats(reform_server* server, reform_server_counters* stats);

#ifdef __cplusplus
}
#endif
#endif
