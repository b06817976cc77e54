/* PRAQ on MI355X: the QUBO roof-dual LP by restarted reflected Halpern PDHG,
 * as hand-written gfx950 HIP kernels behind a C ABI.
 *
 * The first block of declarations is the reference's praq-cuda ABI,
 * unchanged in layout and meaning, so its callers (the harness's ctypes
 * `Library`, the file CLI, the tuner) bind this library as they bind
 * libpraq.so:
 *   PraqInstance / PraqParameters / PraqResult   qubo/praq-cuda/include/praq/praq.h:20-75
 *   praq_default_parameters                      qubo/praq-cuda/include/praq/praq.h:77-79
 *   praq_check_interval                          qubo/praq-cuda/include/praq/praq.h:81-84
 *   praq_solve                                   qubo/praq-cuda/include/praq/praq.h:86-95
 * Status codes: 0 converged, 1 iteration limit, -1 argument, -2 instance,
 * -3 device (qubo/praq-cuda/src/solver.cuh:121-124).
 *
 * The second block extends it (no counterpart in the reference): options the
 * praq-jax Python API has and praq.h lacks (use_halpern, warm_steps, the
 * dense-build degree, a warm start), the dual output, and a choice of
 * execution tier.
 */
#ifndef REFORMULATION_PRAQ_HIP_H
#define REFORMULATION_PRAQ_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int64_t n;              /* binary variables                              */
    int64_t m;              /* off-diagonal nonzeros (edges)                 */
    const double* diag;     /* [n]                                           */
    const int64_t* rows;    /* [m]                                           */
    const int64_t* cols;    /* [m]                                           */
    const double* vals;     /* [m], W_ij = W_ji                              */
} PraqInstance;

typedef struct {
    double tolerance;       /* relative duality gap at termination           */
    int64_t max_iter;       /* <= 0 for no cap                               */
    int use_float32;        /* 0 = float64 iterates; 1 = float32 iterates
                             * (reductions and scalar logic stay float64),
                             * switching to float64 if the gap stalls        */
    double omega;
    double beta_sufficient;
    double beta_necessary;
    double beta_artificial;
    double theta_initial;
    double theta_target;
    double theta_decay;
    int64_t op_steps;       /* <= 0: praq_check_interval(n)                  */
    double interval_offset;
    double interval_scale;
    double interval_power;
} PraqParameters;

typedef struct {
    double lower;
    double upper;
    int64_t iterations;
    int status;             /* 0 converged, 1 iteration limit, < 0 error     */
} PraqResult;

PraqParameters praq_default_parameters(void);
int64_t praq_check_interval(int64_t n, const PraqParameters* parameters);
int praq_solve(const PraqInstance* instance, const PraqParameters* parameters,
               double* x, PraqResult* result);

/* ---- extensions -------------------------------------------------------- */

/* Execution tier. */
enum {
    PRAQ_TIER_AUTO = 0,     /* by size: one workgroup below a threshold      */
    PRAQ_TIER_WORKGROUP = 1,/* the whole solve in one workgroup, one launch  */
    PRAQ_TIER_DEVICE = 2    /* every workgroup of the GPU per iteration      */
};

typedef struct {
    int use_halpern;        /* 1 (default) restarted Halpern; 0 vanilla PDHG */
    int64_t warm_steps;     /* vanilla op_steps blocks before Halpern (0)    */
    int dense_degree;       /* 1: degree n-1 on every node, praq-jax's dense
                             * build; 0 (default): the true degree           */
    int tier;               /* PRAQ_TIER_*                                   */
    double eta;             /* step size; 0 means 0.999                      */
    double time_limit;      /* seconds; <= 0: none (the workgroup tier still
                             * stops after 600 s so no launch runs forever)  */
} PraqOptions;

typedef struct {
    double lower, upper, relative_gap;
    int64_t iterations, restarts, checks;
    double omega;           /* final primal weight                           */
    double scale;           /* objective scale                               */
    int status;             /* 0 converged, 1 iteration limit, 2 time limit  */
    int tier;               /* tier that ran                                 */
    int64_t switch_iteration; /* FP32 iterates handed over to FP64 at this
                               * iteration (gap below 3e-5, stall, or a
                               * tolerance below 3e-5 met); -1 if they did not */
    int64_t device_bytes;   /* device memory the solve held at its peak: its
                             * images (both precisions when FP32 handed over)
                             * and work areas                               */
    double bracket;         /* upper - lower, instance units (NaN on error).
                             * A solve that finished in FP64 (FP64 iterates,
                             * or FP32 ones after switch_iteration; every
                             * FP32 solve with a tolerance below 3e-5) brackets
                             * the LP optimum: lower <= OPT <= upper, upper
                             * the objective at x, so x's objective error is
                             * at most this.  One that finished in FP32
                             * (switch_iteration -1) brackets the optimum of
                             * the instance as its FP32 image holds it (each
                             * cost rounded to FP32), within 2^-24 (sum |c_i|
                             * + sum |q_ij|) of OPT.                        */
} PraqInfo;

PraqOptions praq_default_options(void);

/* praq_solve with the extensions.  warm_x [n] and warm_y [m][2] (instance
 * units) may be NULL for the cold start; y [m][2] receives the dual in
 * instance units and may be NULL. */
int praq_solve_ex(const PraqInstance* instance, const PraqParameters* parameters,
                  const PraqOptions* options, const double* warm_x, const double* warm_y,
                  double* x, double* y, PraqInfo* info);

/* praq_solve_ex over device memory (no counterpart in the reference, whose
 * entry points take host arrays, qubo/praq-cuda/src/praq.cu:35-79): the
 * instance's four arrays and x [n], y [m][2] (y may be NULL) are device
 * pointers -- caller-owned allocations such as torch tensors' data_ptr() --
 * and every launch and copy is ordered on `stream` (a hipStream_t; NULL: the
 * default stream).  The COO arrays are copied to the host for the O(n + m)
 * half-edge build (the reference also builds on the host,
 * qubo/praq-cuda/src/solver.cuh:552-560); the answer is written on the
 * device.  Returns when the answer is in place; codes as praq_solve_ex. */
int praq_solve_device(const PraqInstance* instance, const PraqParameters* parameters,
                      const PraqOptions* options, double* x, double* y, PraqInfo* info,
                      void* stream);

#ifdef __cplusplus
}
#endif
#endif
