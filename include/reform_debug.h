/* Fault injection for tests of the library's error paths.
 *
 * No counterpart in the reference, whose error paths are reached only by
 * invalid arguments (qubo/praq-cuda/tests/test_solver.cu:210-236).  Every
 * device allocation the library makes (solve images, batch arenas, the
 * serving pool and its tables) goes through one wrapper; armed here, the
 * wrapper makes one of them fail as hipMalloc does when the device is out
 * of memory, so each entry point's handling of that failure (an error code,
 * nothing leaked, the next call unaffected) can be tested from the host.
 */
#ifndef REFORMULATION_REFORM_DEBUG_H
#define REFORMULATION_REFORM_DEBUG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* after >= 0: the device allocations from now on succeed `after` times, the
 * next one fails with hipErrorOutOfMemory, and the injector disarms itself.
 * after < 0: disarm.  Process-wide; returns 0. */
int reform_debug_fail_device_alloc(int64_t after);

/* Device allocations the library has made (or been refused) so far in this
 * process; the difference across a call is how many that call makes. */
int64_t reform_debug_device_allocs(void);

#ifdef __cplusplus
}
#endif
#endif
