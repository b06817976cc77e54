/* PRIA on MI355X: the joint-replenishment LP relaxation by restarted
 * reflected Halpern PDHG, as hand-written gfx950 HIP kernels behind a C ABI.
 *
 * The first block is the reference's pria-cuda ABI, unchanged in layout and
 * meaning:
 *   pria_params / pria_result_info      jrp/pria-cuda/include/pria/pria.h:34-61
 *   pria_params_auto                    jrp/pria-cuda/include/pria/pria.h:51-52
 *   pria_solve_f64 / pria_solve_f32     jrp/pria-cuda/include/pria/pria.h:63-83
 *   pria_solve_batch_f64 / _f32         jrp/pria-cuda/include/pria/pria.h:85-99
 *   pria_objective_f64                  jrp/pria-cuda/include/pria/pria.h:101-106
 * Inputs: H dense row-major [n][t][t] (upper triangle read), K [n], K0.
 * Outputs x, y, z dense [n][t][t], upper triangular; y and z normalised by
 * K_i and K0 (0 where the budget is 0).  Returns 0 on success, nonzero on a
 * device error (and -1 for an invalid argument).
 *
 * The second block extends it (no counterpart in the reference): a solve
 * taking FP64 inputs with the iterate precision as a flag, a tier choice, a
 * time limit, and the full solver record.
 */
#ifndef REFORMULATION_PRIA_HIP_H
#define REFORMULATION_PRIA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pria_params {
    double omega;
    double eta;
    double beta_sufficient;
    double beta_artificial;
    double beta_necessary;
    double theta_initial;
    double theta_target;
    double theta_decay;
    double tolerance;
    int64_t max_iter;        /* -1 for unlimited */
    int use_halpern;
    int warm_steps;
    int op_steps;
    int verbose;
} pria_params;

pria_params pria_params_auto(int64_t max_iter, double tolerance);

typedef struct pria_result_info {
    int64_t iterations;
    double primal_objective;  /* upper bound at the final iterate (padded problem) */
    double dual_objective;    /* lower bound at the final iterate (padded problem) */
    double relative_gap;
} pria_result_info;

int pria_solve_f64(int64_t n, int64_t t, const double* order_costs, double joint_cost,
                   const double* holding_costs, const pria_params* params,
                   const double* warm_x, const double* warm_y, const double* warm_z,
                   double* x, double* y, double* z, pria_result_info* info);

int pria_solve_f32(int64_t n, int64_t t, const float* order_costs, float joint_cost,
                   const float* holding_costs, const pria_params* params,
                   const float* warm_x, const float* warm_y, const float* warm_z,
                   float* x, float* y, float* z, pria_result_info* info);

int pria_solve_batch_f64(int64_t count, int64_t n, int64_t t, const double* order_costs,
                         const double* joint_costs, const double* holding_costs,
                         const pria_params* params, double* x, double* y, double* z,
                         pria_result_info* infos);

int pria_solve_batch_f32(int64_t count, int64_t n, int64_t t, const float* order_costs,
                         const float* joint_costs, const float* holding_costs,
                         const pria_params* params, float* x, float* y, float* z,
                         pria_result_info* infos);

double pria_objective_f64(int64_t n, int64_t t, const double* order_costs, double joint_cost,
                          const double* holding_costs, const double* x);

/* ---- extensions -------------------------------------------------------- */

enum {
    PRIA_TIER_AUTO = 0,       /* by size                                      */
    PRIA_TIER_WORKGROUP = 1,  /* the whole solve in one workgroup             */
    PRIA_TIER_DEVICE = 2      /* every workgroup of the GPU per iteration     */
};

typedef struct {
    int tier;                 /* PRIA_TIER_*                                  */
    double time_limit;        /* seconds; <= 0: none (the workgroup tier still
                               * stops after 600 s)                           */
} pria_options;

typedef struct {
    int64_t iterations, restarts, checks;
    double lower;             /* solver's own lower bound (padded problem)    */
    double upper;             /* primal objective of x on the original problem */
    double upper_iterate;     /* upper bound the termination test used        */
    double relative_gap;      /* gap the termination test saw last            */
    double omega, scale;
    int status;               /* 0 converged, 1 iteration limit, 2 time limit */
    int tier;
    int64_t switch_iteration; /* FP32 iterates handed over to FP64 at this
                               * iteration (gap below 3e-5, stall, or a
                               * tolerance below 3e-5 met); -1 if they did not */
    int64_t device_bytes;     /* device memory the solve held at its peak     */
    double bracket;           /* upper - lower, instance units (NaN on error).
                               * upper is the objective of x on the original
                               * instance.  A solve that finished in FP64
                               * (FP64 iterates, or FP32 ones after
                               * switch_iteration; every FP32 solve with a
                               * tolerance below 3e-5) has lower <= OPT <=
                               * upper to FP64 rounding (the padded period of
                               * an odd t costs nothing, so padding leaves
                               * OPT alone): x's objective error is at most
                               * this.  One that finished in FP32
                               * (switch_iteration -1) brackets the optimum
                               * of the instance as its FP32 image holds it
                               * (costs rounded to FP32), and x's columns sum
                               * to 1 to FP32 rounding: both bounds within
                               * 2^-24 (sum |h| + t sum K_i + t K0) of OPT.   */
} pria_info;

pria_options pria_default_options(void);

/* One solve with FP64 inputs/outputs; use_float32 selects FP32 iterates
 * (reductions and scalar logic stay FP64).  Returns 0 (converged), 1 (limit),
 * -1 (argument), -3 (device). */
int pria_solve_ex(int64_t n, int64_t t, const double* order_costs, double joint_cost,
                  const double* holding_costs, const pria_params* params,
                  const pria_options* options, int use_float32, const double* warm_x,
                  const double* warm_y, const double* warm_z, double* x, double* y, double* z,
                  pria_info* info);

/* pria_solve_ex over device memory, cold start (no counterpart in the
 * reference, whose entry points take host arrays, jrp/pria-cuda/src/pria.cu:72-108):
 * order_costs [n], holding_costs [n][t][t] and the outputs x, y, z
 * [n][t][t] (each may be NULL) are device pointers -- caller-owned
 * allocations such as torch tensors' data_ptr() -- and every launch and
 * copy is ordered on `stream` (a hipStream_t; NULL: the default stream).
 * The scale, the slice-order image, the dense answers and the upper bound
 * are computed on the device; only scalars cross PCIe.  Returns when the
 * answer is in place; codes as pria_solve_ex. */
int pria_solve_device(int64_t n, int64_t t, const double* order_costs, double joint_cost,
                      const double* holding_costs, const pria_params* params,
                      const pria_options* options, int use_float32, double* x, double* y,
                      double* z, pria_info* info, void* stream);

/* Bytes of the device array pria_solve_device_noalloc works in, per GPU. */
#define PRIA_NOALLOC_BYTES (64ll << 20)

/* pria_solve_device that allocates nothing (no counterpart in the reference:
 * it is for callers that cannot allocate device memory -- code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments, results and codes as pria_solve_device, and the same kernels
 * give the same bits.  Every scratch byte of the solve (scale partials,
 * objective rows, x when the caller takes none, the iterate images of
 * both precisions) lives in one static device array of PRIA_NOALLOC_BYTES
 * per GPU, held from the library's load until the process ends.  An
 * instance that needs more (about 140 bytes per entry N T'(T'+1)/2 with
 * FP64 iterates, 210 with FP32 ones and their FP64 handover) is rejected
 * with -1 and an explanation in reform_last_error().  Calls on one GPU
 * share the array: within a process they take turns (a call waits for the
 * one before it to return); the array keeps the last call's contents, which
 * no call reads. */
int pria_solve_device_noalloc(int64_t n, int64_t t, const double* order_costs,
                              double joint_cost, const double* holding_costs,
                              const pria_params* params, const pria_options* options,
                              int use_float32, double* x, double* y, double* z,
                              pria_info* info, void* stream);

/* Test hook: one PDHG step x2 = P_X(x1 - tau (h + y1 + z1)), r = sigma
 * (2 x2 - x1), y2 = P_Y(y1 + r), z2 = P_Z(z1 + r) on an even-T instance whose
 * h, K, K0 are already scaled; dense [n][t][t] in and out.  tier 1: one
 * workgroup; 2: the device-wide kernels. */
int pria_debug_step(int64_t n, int64_t t, const double* k, double k0, const double* h,
                    const double* x1, const double* y1, const double* z1, double omega,
                    double eta, int use_float32, int tier, double* x2, double* y2, double* z2);

#ifdef __cplusplus
}
#endif
#endif
