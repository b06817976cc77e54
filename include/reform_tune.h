/* Shape tuning (no counterpart in the reference, whose tuners tune restart
 * parameters through praq.h, qubo/harness/praq_cuda_library.py:90-139).
 *
 * A process-wide table maps keys to integer shape choices; the autotuner
 * (reformulation_amd/autotune.py) times candidates with
 * reform_time_praq_steps on the GPU and writes a table, which the Python
 * package loads at import.  Unset keys keep the built-in rules.
 *
 * Keys (a value of 0 leaves the built-in choice):
 *   "praq.lanes.<b>", b = 0..7: lanes per node (1, 2, ..., 64) of
 *   single-instance device-tier PRAQ solves whose mean degree 2m/n lies in
 *   [2^b, 2^(b+1)) (b = 0 also below 1, b = 7 also above 256).  Lanes set
 *   the CSR gather's lane groups, and hubs (degree above 64 lanes) take the
 *   chunked gather.
 *   "batch.praq.lanes.<b>": the same for the instances of a batch
 *   (reform_batch.h) and of a server (reform_serve.h); built-in: at most 8
 *   half-edges per lane at the mean degree.
 *   "batch.pria.x_threads": 256 or 512 threads per FP32 x-tile workgroup of
 *   the batched and served PRIA step (8 or 16 lanes per column); built-in
 *   256.
 *   "batch.pria.x_hold": 1 FP32 x-tile threads keep x1 in registers for the
 *   write phase (when every T' <= 320), 2 they read it again; built-in 1.
 *   "batch.pria.yz_shape": 1 the compact y/z register shape when every T'
 *   <= 320, 2 always the full one; built-in 1.
 * A batch takes its instances' lanes when it uploads them and the PRIA
 * shapes when its first batched solve starts; a server takes a request's
 * lanes when it builds the request and the PRIA shapes whenever it
 * rebuilds its tables, so change the table only while no server runs.
 * x_hold and yz_shape give the same results bit for bit; lanes and
 * x_threads change the order of FP32/FP64 sums (x_threads only FP32's),
 * not the algorithm.
 */
#ifndef REFORMULATION_REFORM_TUNE_H
#define REFORMULATION_REFORM_TUNE_H

#include <stdint.h>

#include "praq_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* 0, or -1 for an unknown key or a value out of range. */
int reform_tune_set(const char* key, int64_t value);
/* The value, 0 if unset, -1 for an unknown key. */
int64_t reform_tune_get(const char* key);
void reform_tune_clear(void);

/* Lanes per node of a device-tier PRAQ solve of an instance of this shape. */
int reform_praq_lanes_for(int64_t n, int64_t m);
/* Lanes per node of the same instance in a batch or a server. */
int reform_batch_praq_lanes_for(int64_t n, int64_t m);

/* Mean milliseconds of one device-tier PDHG step of `instance` (hub chunks,
 * node, hub and edge kernels; a reflected Halpern step) with `lanes` lanes
 * per node (0: the library's choice), FP32 or FP64 iterates, measured with
 * HIP events over `steps` steps after `warmup` untimed ones.  Returns 0, -1
 * argument, -2 instance, -3 device. */
int reform_time_praq_steps(const PraqInstance* instance, int use_float32, int lanes, int warmup,
                           int steps, double* ms);

#ifdef __cplusplus
}
#endif
#endif
