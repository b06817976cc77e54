/* Many PRAQ and PRIA instances of uneven size on one MI355X.
 *
 * No counterpart in the reference beyond pria_solve_batch_* (same-shape PRIA
 * only, a thread fan-out of solo solves: jrp/pria-cuda/include/pria/pria.h:85-99,
 * jrp/pria-cuda/src/pria.cu:15-44).  Here a batch holds instances of any size
 * of both families, packed ragged in one device allocation, and solves them
 *   REFORM_BATCHED     all together through block tables: every kernel
 *                      launch covers the work of every instance, each
 *                      instance keeps its own restart state and termination
 *                      flag on the device, and one check interval's launches
 *                      are a hipGraph replayed until all have finished;
 *   REFORM_SEQUENTIAL  one at a time with the plain solver's tier choice
 *                      (praq_solve_ex / pria_solve_ex on resident inputs):
 *                      the baseline the batched mode is measured against;
 *   REFORM_WORKGROUPS  one persistent launch per family in which each
 *                      instance is one workgroup running its whole solve.
 * An instance's batched result does not depend on the other instances of the
 * batch: its arithmetic and summation order are those of its own blocks.
 * An instance's options.time_limit (seconds, > 0) counts from the start of
 * each solve in every mode: past it the instance stops at its current point
 * (between two rounds in the batched mode) with status 2.
 *
 * Usage: create, add_* (host), upload (host build + one H2D copy), then
 * solve any number of times from the resident inputs, then read results.
 * Return codes: >= 0 success (add_* return the instance's index within its
 * family), -1 argument, -2 instance, -3 device.
 */
#ifndef REFORMULATION_REFORM_BATCH_H
#define REFORMULATION_REFORM_BATCH_H

#include <stdint.h>

#include "praq_hip.h"
#include "pria_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct reform_batch reform_batch;

enum { REFORM_BATCHED = 0, REFORM_SEQUENTIAL = 1, REFORM_WORKGROUPS = 2 };

/* use_float32: iterate precision of every instance of the batch
 * (reductions and scalar logic stay FP64). */
reform_batch* reform_batch_create(int use_float32);
void reform_batch_destroy(reform_batch* batch);

int64_t reform_batch_add_praq(reform_batch* batch, const PraqInstance* instance,
                              const PraqParameters* parameters, const PraqOptions* options);
int64_t reform_batch_add_pria(reform_batch* batch, int64_t n, int64_t t, const double* order_costs,
                              double joint_cost, const double* holding_costs,
                              const pria_params* params, const pria_options* options);

/* Builds every instance's device image on the host (threads) and copies the
 * inputs to the GPU in one transfer. */
int reform_batch_upload(reform_batch* batch);

/* Solves every instance from its resident inputs; synchronous. */
int reform_batch_solve(reform_batch* batch, int mode);

/* Workgroup sizes of the batched launches (0 keeps the default); a tuner
 * sets these. */
int reform_batch_set_shape(reform_batch* batch, int praq_block, int pria_block);

int64_t reform_batch_count(const reform_batch* batch, int family /* 0 PRAQ, 1 PRIA */);

/* Text of the library's last failed HIP call (call, source place, HIP error
 * string), "" if none; valid until the calling thread's next call. */
const char* reform_last_error(void);

/* Durations of the last batched solve's two families, measured with HIP
 * events on the streams they ran on (0 for an empty family or a sequential
 * solve). */
int reform_batch_kernel_ms(const reform_batch* batch, double* praq_ms, double* pria_ms);
int64_t reform_batch_device_bytes(const reform_batch* batch);
/* Host seconds the last sequential solve spent on instance `index` of the
 * family (from its first launch to its synchronized end, the FP64 handover
 * included), -1 if it has none: the plain solver's service time per
 * instance, inputs already resident. */
double reform_batch_instance_seconds(const reform_batch* batch, int family, int64_t index);

/* Launch timing of the batched mode's two step kernels per family (PRAQ:
 * node, edge; PRIA: x, y/z).  With timing on, the rounds are launched
 * directly instead of replayed from their graph, and every step launch is
 * bracketed by HIP events on its stream; the figures cover the last batched
 * solve.  bytes: algorithmic HBM bytes those launches moved (per step, per
 * instance: PRAQ node n (4 + 7b) + 2m (4 + b), edge m (8 + 7b); PRIA x 7b ND,
 * y/z 8b ND; b = bytes per value in the instance's precision at that step). */
typedef struct reform_kernel_stats {
    char name[64];       /* kernel symbol, e.g. "wpria_yz" */
    int64_t launches;
    double ms;           /* sum of the launches' durations */
    double bytes;        /* sum of their algorithmic bytes */
} reform_kernel_stats;
int reform_batch_set_timing(reform_batch* batch, int on);
/* Batched mode: rebuild the block tables over the instances still running
 * whenever at most half of a table's instances are (default on).  Results do
 * not change; later rounds launch only the running instances' blocks. */
int reform_batch_set_compaction(reform_batch* batch, int on);
int reform_batch_kernel_stats(const reform_batch* batch, int family,
                              reform_kernel_stats out[2]);

/* Results of the last solve.  Outputs may be NULL.  PRAQ y is [m][2] in
 * instance units; PRIA x, y, z are dense [n][t][t], y and z normalised. */
int reform_batch_praq_result(reform_batch* batch, int64_t index, double* x, double* y,
                             PraqInfo* info);
int reform_batch_pria_result(reform_batch* batch, int64_t index, double* x, double* y, double* z,
                             pria_info* info);

#ifdef __cplusplus
}
#endif
#endif
