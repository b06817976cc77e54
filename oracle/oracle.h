/* The oracle: a plain-C, FP64 restatement of the reference's PRAQ and PRIA
 * solve semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped package links, loads or
 * calls this code; only tests/, the smoke check and bench.py's cpu_baseline
 * leg do, and only as the checker (or as the timed CPU baseline).
 *
 * PRAQ follows qubo/praq-jax/praq/solver.py (the spec) with the sparse
 * half-edge build of qubo/praq-cuda/src/solver.cuh:524-568; `dense_degree`
 * selects praq-jax's dense-build degree (n - 1 for every node), which is the
 * dense build exactly, since absent edges carry weight 0 and stay inert.
 * PRIA follows jrp/pria-jax/pria/solver.py on plain [N][T'][T'] arrays.
 *
 * Every function is deterministic: reductions are split into a fixed number
 * of chunks whose partial sums are added in a fixed order, so the result
 * does not depend on how many OpenMP threads ran it.
 */
#ifndef REFORMULATION_ORACLE_H
#define REFORMULATION_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Parameters of both solvers: praq-jax / pria-jax `Parameters` fields. */
typedef struct {
    double omega;
    double eta;
    double beta_sufficient;
    double beta_necessary;
    double beta_artificial;
    double theta_initial;
    double theta_target;
    double theta_decay;
    double tolerance;
    int64_t max_iter;      /* < 0: no cap                                     */
    int64_t op_steps;      /* PRAQ: <= 0 uses check_interval; PRIA: >= 2      */
    int64_t warm_steps;
    int64_t use_halpern;
    int64_t dense_degree;  /* PRAQ only: 1 = praq-jax dense-build degree n-1  */
    double interval_offset, interval_scale, interval_power;
} OracleParams;

typedef struct {
    int64_t iterations;
    int64_t restarts;      /* restarts taken (not counted in iterations)      */
    int64_t checks;        /* Halpern checks evaluated                        */
    double lower;          /* solver's own lower bound, original units        */
    double upper;          /* upper bound reported to the caller              */
    double upper_iterate;  /* upper bound the termination test used           */
    double relative_gap;   /* gap the termination test saw last               */
    double omega;          /* final primal weight                             */
    double scale;          /* objective scale                                 */
    int64_t status;        /* 0 converged, 1 iteration limit, <0 error        */
} OracleInfo;

/* ---- PRAQ --------------------------------------------------------------- */

int64_t oracle_praq_check_interval(int64_t n, double offset, double scale, double power);

/* Closed-form projection onto the dual set of one edge (projection.py:16-42);
 * v is [m][2] in place. */
void oracle_praq_project_dual(int64_t m, double* v, const double* w, const uint8_t* positive);

/* Solve.  y_out (optional) is [m][2] in instance units (the dual times the
 * objective scale), as solve_qubo_problem returns it.  warm_x / warm_y
 * (optional) are in instance units, as solve_qubo_problem_warm_start takes
 * them. */
int oracle_praq_solve(int64_t n, int64_t m, const double* diag, const int64_t* rows,
                      const int64_t* cols, const double* vals, const OracleParams* p,
                      const double* warm_x, const double* warm_y,
                      double* x_out, double* y_out, OracleInfo* info);

/* relaxation_bounds (solver.py:348-362) from a primal [n] and a dual [m][2]
 * in instance units, with the given degree mode irrelevant (bounds do not
 * use degree). */
void oracle_praq_bounds(int64_t n, int64_t m, const double* diag, const int64_t* rows,
                        const int64_t* cols, const double* vals, const double* x,
                        const double* y, double* lower, double* upper);

/* One raw step from (x1, y1) (scaled units, y [m][2]) and the canonical norm
 * of (dx, dy); the scaled edge weights / covering flags / scale. */
int oracle_praq_raw_step(int64_t n, int64_t m, const double* diag, const int64_t* rows,
                         const int64_t* cols, const double* vals, int dense_degree, double omega,
                         double eta, const double* x1, const double* y1, double* x2, double* y2);
double oracle_praq_canonical(int64_t n, int64_t m, const double* diag, const int64_t* rows,
                             const int64_t* cols, const double* vals, int dense_degree,
                             const double* dx, const double* dy, double eta, double omega);
int oracle_praq_structure(int64_t n, int64_t m, const double* diag, const int64_t* rows,
                          const int64_t* cols, const double* vals, double* w_out,
                          uint8_t* pos_out, double* scale_out);

/* ---- PRIA --------------------------------------------------------------- */

/* Euclidean projection of v[len] onto {u >= 0, sum u = target}; target <= 0
 * gives exactly 0 (projection.py:71-130, sort-based). */
void oracle_simplex_project(int64_t len, double* v, double target);

/* Solve.  h is [n][t][t] (upper triangle read), k [n].  Outputs x, y, z
 * (optional) are [n][t][t], y and z normalised by K_i and K0 (0 where the
 * budget is 0), unpadded, as solve_jrp_problem returns them. */
int oracle_pria_solve(int64_t n, int64_t t, const double* k, double k0, const double* h,
                      const OracleParams* p, const double* warm_x, const double* warm_y,
                      const double* warm_z, double* x_out, double* y_out, double* z_out,
                      OracleInfo* info);

/* One raw step on an even-T instance whose h, k, k0 are already scaled;
 * dense [n][t][t] in and out. */
int oracle_pria_raw_step(int64_t n, int64_t t, const double* k, double k0, const double* h,
                         const double* x1, const double* y1, const double* z1, double omega,
                         double eta, double* x2, double* y2, double* z2);

/* primal_objective (jrp_problem.py:143-161) at x [n][t][t]. */
double oracle_pria_primal_objective(int64_t n, int64_t t, const double* k, double k0,
                                    const double* h, const double* x);

/* relaxation_bounds (solver.py:740-790) lower bound from normalised y, z. */
double oracle_pria_dual_objective(int64_t n, int64_t t, const double* k, double k0,
                                  const double* h, const double* y, const double* z);

/* Roof dual (= the PRAQ relaxation's value) by max-flow with int64
 * capacities round(a scale) for each posiform term a (maxflow.c); exact for
 * weights on the 1 / scale grid.  NaN on bad arguments or memory. */
double oracle_praq_roof_dual(int64_t n, const double* diag, int64_t m, const int64_t* rows,
                             const int64_t* cols, const double* vals, double scale,
                             int64_t* flow_out);

#ifdef __cplusplus
}
#endif
#endif
