/* PRAQ oracle: FP64 restatement of the QUBO roof-dual PDHG solve.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows, function by function:
 *   HalfEdge.build / gather / scatter / shift / centroid / scaled
 *       qubo/praq-jax/praq/half_edge.py:42-99 (sparse edge set as in
 *       qubo/praq-cuda/src/solver.cuh:524-568)
 *   objective_scale          qubo/praq-jax/praq/qubo_problem.py:113-124
 *   project_primal/dual      qubo/praq-jax/praq/projection.py:5-42
 *   canonical_norm           qubo/praq-jax/praq/halpern_operator.py:7-35
 *   _raw_step                qubo/praq-jax/praq/solver.py:259-273
 *   _reflect_step            qubo/praq-jax/praq/solver.py:276-292
 *   _upper/_lower_bound_raw  qubo/praq-jax/praq/solver.py:320-337
 *   _obj_gap / termination   qubo/praq-jax/praq/solver.py:365-381
 *   _halpern_step            qubo/praq-jax/praq/solver.py:397-469
 *   solve_*_warm_start       qubo/praq-jax/praq/solver.py:483-561
 *   check_interval           qubo/praq-jax/praq/solver.py:199-225
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NCHUNK 64

static int64_t chunk_lo(int64_t len, int c) { return (len * c) / NCHUNK; }

#define SCALE_CHUNK 4096

static double sumsq_blocked(const double *v, int64_t len, int weights) {
    double total = 0.0;
    for (int64_t c0 = 0; c0 < len; c0 += SCALE_CHUNK) {
        int64_t c1 = c0 + SCALE_CHUNK < len ? c0 + SCALE_CHUNK : len;
        double part = 0.0;
        for (int64_t k = c0; k < c1; ++k) part += weights ? 4.0 * v[k] * v[k] : v[k] * v[k];
        total += part;
    }
    return total;
}

/* ---- structure ---------------------------------------------------------- */

typedef struct {
    int64_t n, m;
    int64_t *e0, *e1;          /* ends of edge e                                  */
    double *w;                 /* w_e = 2|W_ij| / scale                           */
    uint8_t *pos;              /* W_ij > 0                                        */
    double *deg;               /* degree, floored at 1                            */
    double *d;                 /* diag / scale                                    */
    int64_t *ptr, *half;       /* CSR over half-edges, half = 2e + slot           */
    double scale;
} Praq;

static void praq_free(Praq* s) {
    free(s->e0); free(s->e1); free(s->w); free(s->pos); free(s->deg); free(s->d);
    free(s->ptr); free(s->half);
}

static int praq_build(Praq* s, int64_t n, int64_t m, const double* diag, const int64_t* rows,
                      const int64_t* cols, const double* vals, int dense_degree) {
    memset(s, 0, sizeof(*s));
    if (n <= 0 || m < 0 || !diag || (m > 0 && (!rows || !cols || !vals))) return -2;
    for (int64_t e = 0; e < m; ++e) {
        if (rows[e] < 0 || rows[e] >= n || cols[e] < 0 || cols[e] >= n || rows[e] == cols[e])
            return -2;
    }
    s->n = n; s->m = m;
    s->e0 = malloc(sizeof(int64_t) * (m + 1));
    s->e1 = malloc(sizeof(int64_t) * (m + 1));
    s->w = malloc(sizeof(double) * (m + 1));
    s->pos = malloc(m + 1);
    s->deg = malloc(sizeof(double) * n);
    s->d = malloc(sizeof(double) * n);
    s->ptr = calloc(n + 1, sizeof(int64_t));
    s->half = malloc(sizeof(int64_t) * (2 * m + 1));

    /* objective_scale: ||(diag; 2|W_ij|)||_2, 1 if it is 0.  The squares in
     * a fixed blocked order (chunks of 4096 terms in order, the chunk sums in
     * order, diagonal and weights as two such sums), the order the solver
     * builds in (the reference's jnp.linalg.norm reduces in an order of its
     * own), so the oracle's scale has the solver's bits. */
    double sq = sumsq_blocked(diag, n, 0) + sumsq_blocked(vals, m, 1);
    s->scale = sq > 0.0 ? sqrt(sq) : 1.0;

    for (int64_t e = 0; e < m; ++e) {
        s->e0[e] = rows[e];
        s->e1[e] = cols[e];
        s->w[e] = 2.0 * fabs(vals[e]) / s->scale;
        s->pos[e] = vals[e] > 0.0;
    }
    for (int64_t i = 0; i < n; ++i) s->d[i] = diag[i] / s->scale;

    for (int64_t e = 0; e < m; ++e) { ++s->ptr[s->e0[e] + 1]; ++s->ptr[s->e1[e] + 1]; }
    for (int64_t i = 0; i < n; ++i) s->ptr[i + 1] += s->ptr[i];
    int64_t* fill = malloc(sizeof(int64_t) * (n + 1));
    memcpy(fill, s->ptr, sizeof(int64_t) * n);
    for (int64_t e = 0; e < m; ++e) {
        s->half[fill[s->e0[e]]++] = 2 * e;
        s->half[fill[s->e1[e]]++] = 2 * e + 1;
    }
    free(fill);
    for (int64_t i = 0; i < n; ++i) {
        double incident = dense_degree ? (double)(n - 1) : (double)(s->ptr[i + 1] - s->ptr[i]);
        s->deg[i] = incident > 1.0 ? incident : 1.0;
    }
    return 0;
}

/* scatter S^T y at node i */
static double scatter_at(const Praq* s, const double* y, int64_t i) {
    double acc = 0.0;
    for (int64_t k = s->ptr[i]; k < s->ptr[i + 1]; ++k) acc += y[s->half[k]];
    return acc;
}

/* ---- projections -------------------------------------------------------- */

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

static void project_edge(double* v, double w, int positive) {
    if (positive) {
        double mu = clampd(0.5 * (v[0] + v[1]), -w, 0.0);
        v[0] = mu;
        v[1] = mu;
    } else {
        double line = v[0] + 0.5 * (w - v[0] - v[1]);
        double first = clampd(line, 0.0, w);
        v[0] = first;
        v[1] = w - first;
    }
}

void oracle_praq_project_dual(int64_t m, double* v, const double* w, const uint8_t* positive) {
    for (int64_t e = 0; e < m; ++e) project_edge(v + 2 * e, w[e], positive[e]);
}

/* ---- one PDHG step (_raw_step) ----------------------------------------- */

static void raw_step(const Praq* s, const double* x1, const double* y1, double omega, double eta,
                     double* x2, double* y2, double* refl) {
    const int64_t n = s->n, m = s->m;
    const double tau = eta / omega, sigma = eta * omega;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double g = s->d[i] - scatter_at(s, y1, i);
        x2[i] = clampd(x1[i] - tau * (g / s->deg[i]), 0.0, 1.0);
        refl[i] = 2.0 * x2[i] - x1[i];
    }
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < m; ++e) {
        double v[2];
        v[0] = y1[2 * e] + sigma * ((s->pos[e] ? 1.0 : 0.0) - refl[s->e0[e]]);
        v[1] = y1[2 * e + 1] + sigma * (0.0 - refl[s->e1[e]]);
        project_edge(v, s->w[e], s->pos[e]);
        y2[2 * e] = v[0];
        y2[2 * e + 1] = v[1];
    }
}

/* ---- reductions --------------------------------------------------------- */

/* canonical_norm of (x - sx, y - sy) */
static double canonical(const Praq* s, const double* x, const double* sx, const double* y,
                        const double* sy, double eta, double omega) {
    double pn[NCHUNK], dn[NCHUNK], cr[NCHUNK];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < NCHUNK; ++c) {
        double a = 0.0, b = 0.0, q = 0.0;
        for (int64_t i = chunk_lo(s->n, c); i < chunk_lo(s->n, c + 1); ++i) {
            double dx = x[i] - sx[i];
            a += s->deg[i] * dx * dx;
        }
        for (int64_t e = chunk_lo(s->m, c); e < chunk_lo(s->m, c + 1); ++e) {
            double d0 = y[2 * e] - sy[2 * e], d1 = y[2 * e + 1] - sy[2 * e + 1];
            b += d0 * d0 + d1 * d1;
            q += (x[s->e0[e]] - sx[s->e0[e]]) * d0 + (x[s->e1[e]] - sx[s->e1[e]]) * d1;
        }
        pn[c] = a; dn[c] = b; cr[c] = q;
    }
    double a = 0.0, b = 0.0, q = 0.0;
    for (int c = 0; c < NCHUNK; ++c) { a += pn[c]; b += dn[c]; q += cr[c]; }
    double val = (omega / eta) * a + (1.0 / (eta * omega)) * b + 2.0 * q;
    return sqrt(val > 0.0 ? val : 0.0);
}

/* (upper, lower) bounds at (bx, by), original scale */
static void bounds(const Praq* s, const double* bx, const double* by, double* upper, double* lower) {
    double up[NCHUNK], lo[NCHUNK];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < NCHUNK; ++c) {
        double u = 0.0, l = 0.0;
        for (int64_t i = chunk_lo(s->n, c); i < chunk_lo(s->n, c + 1); ++i) {
            u += s->d[i] * bx[i];
            double rc = s->d[i] - scatter_at(s, by, i);
            l += rc < 0.0 ? rc : 0.0;
        }
        for (int64_t e = chunk_lo(s->m, c); e < chunk_lo(s->m, c + 1); ++e) {
            double bi = bx[s->e0[e]], bj = bx[s->e1[e]];
            if (s->pos[e]) {
                double t = bi + bj - 1.0;
                u += s->w[e] * (t > 0.0 ? t : 0.0);
                l += by[2 * e];
            } else {
                u += -s->w[e] * (bi < bj ? bi : bj);
            }
        }
        up[c] = u; lo[c] = l;
    }
    double u = 0.0, l = 0.0;
    for (int c = 0; c < NCHUNK; ++c) { u += up[c]; l += lo[c]; }
    *upper = u * s->scale;
    *lower = l * s->scale;
}

static double rel_gap(double upper, double lower) {
    return (upper - lower) / (1.0 + fabs(upper) + fabs(lower));
}

/* sqrt(sum deg (a - b)^2), ||c - d|| */
static void drift(const Praq* s, const double* a, const double* b, const double* c,
                  const double* d, double* primal, double* dual) {
    double pp[NCHUNK], dd[NCHUNK];
#pragma omp parallel for schedule(static)
    for (int ch = 0; ch < NCHUNK; ++ch) {
        double u = 0.0, v = 0.0;
        for (int64_t i = chunk_lo(s->n, ch); i < chunk_lo(s->n, ch + 1); ++i) {
            double t = a[i] - b[i];
            u += s->deg[i] * t * t;
        }
        for (int64_t e = chunk_lo(2 * s->m, ch); e < chunk_lo(2 * s->m, ch + 1); ++e) {
            double t = c[e] - d[e];
            v += t * t;
        }
        pp[ch] = u; dd[ch] = v;
    }
    double u = 0.0, v = 0.0;
    for (int ch = 0; ch < NCHUNK; ++ch) { u += pp[ch]; v += dd[ch]; }
    *primal = sqrt(u);
    *dual = sqrt(v);
}

int64_t oracle_praq_check_interval(int64_t n, double offset, double scale, double power) {
    if (n <= 0) return 2;
    double u = log((double)n / 50000.0);
    double b = power == 0.0 ? u : expm1(power * u) / power;
    double v = floor(offset + scale * b + 0.5);
    if (!(v >= 2.0)) return 2;
    if (v > 1e9) return 1000000000;
    return (int64_t)v;
}

/* ---- the solve ---------------------------------------------------------- */

typedef struct {
    double *x, *y, *sx, *sy, *x0, *y0, *refl;
    double omega, theta, fp_error, check_error;
    int64_t restart_iter, iteration;
} State;

static void blend(int64_t len, double* cur, const double* step, const double* anchor, double lw, double rw) {
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < len; ++k) cur[k] = lw * (2.0 * step[k] - cur[k]) + rw * anchor[k];
}

/* _reflect_step: reflect = Halpern combination, else vanilla (x1 = step) */
static void reflect_step(const Praq* s, State* t, double eta, int reflect) {
    if (reflect) {
        double k = (double)(t->restart_iter + 1);
        double lw = k / (k + 1.0), rw = 1.0 / (k + 1.0);
        blend(s->n, t->x, t->sx, t->x0, lw, rw);
        blend(2 * s->m, t->y, t->sy, t->y0, lw, rw);
    } else {
        memcpy(t->x, t->sx, sizeof(double) * s->n);
        memcpy(t->y, t->sy, sizeof(double) * 2 * s->m);
    }
    raw_step(s, t->x, t->y, t->omega, eta, t->sx, t->sy, t->refl);
    t->restart_iter += 1;
    t->iteration += 1;
}

static int within_max_iter(int64_t iteration, int64_t max_iter) {
    return iteration < max_iter || max_iter < 0;
}

int oracle_praq_solve(int64_t n, int64_t m, const double* diag, const int64_t* rows,
                      const int64_t* cols, const double* vals, const OracleParams* pin,
                      const double* warm_x, const double* warm_y,
                      double* x_out, double* y_out, OracleInfo* info) {
    if (!pin || !x_out) return -1;
    OracleParams p = *pin;
    Praq s;
    int st = praq_build(&s, n, m, diag, rows, cols, vals, (int)p.dense_degree);
    if (st) { praq_free(&s); if (info) info->status = st; return st; }

    if (p.op_steps <= 0)
        p.op_steps = oracle_praq_check_interval(n, p.interval_offset, p.interval_scale, p.interval_power);
    if (p.op_steps < 2) p.op_steps = 2;
    if (p.warm_steps < 0) p.warm_steps = 0;

    const double eta = p.eta;
    State t;
    memset(&t, 0, sizeof(t));
    double** arrays[] = {&t.x, &t.sx, &t.x0, &t.refl};
    for (int a = 0; a < 4; ++a) *arrays[a] = malloc(sizeof(double) * n);
    double** duals[] = {&t.y, &t.sy, &t.y0};
    for (int a = 0; a < 3; ++a) *duals[a] = malloc(sizeof(double) * (2 * m + 1));
    double* px = malloc(sizeof(double) * n);           /* projected warm start */
    double* py = malloc(sizeof(double) * (2 * m + 1));

    /* warm start (instance units), projected; cold start = centroid */
    for (int64_t i = 0; i < n; ++i) px[i] = clampd(warm_x ? warm_x[i] : 0.5, 0.0, 1.0);
    for (int64_t e = 0; e < m; ++e) {
        if (warm_y) {
            py[2 * e] = warm_y[2 * e] / s.scale;
            py[2 * e + 1] = warm_y[2 * e + 1] / s.scale;
        } else {
            /* centroid: -W_ij on both slots, scaled; w/2 packing, -w/2 covering */
            double half = 0.5 * (s.pos[e] ? -s.w[e] : s.w[e]);
            py[2 * e] = half;
            py[2 * e + 1] = half;
        }
        project_edge(py + 2 * e, s.w[e], s.pos[e]);
    }
    memcpy(t.x, px, sizeof(double) * n);
    memcpy(t.x0, px, sizeof(double) * n);
    memcpy(t.y, py, sizeof(double) * 2 * m);
    memcpy(t.y0, py, sizeof(double) * 2 * m);
    t.omega = p.omega;
    t.theta = p.theta_initial;
    raw_step(&s, t.x, t.y, t.omega, eta, t.sx, t.sy, t.refl);
    t.fp_error = 0.0;
    t.check_error = INFINITY;

    double upper = 0.0, lower = 0.0, gap = INFINITY;
    int64_t restarts = 0, checks = 0;

    /* the warm (vanilla) phase: lax.while_loop(_check_termination, _op_steps) */
    int64_t warm_max = p.use_halpern ? p.op_steps * p.warm_steps : p.max_iter;
    for (;;) {
        bounds(&s, t.x, t.y, &upper, &lower);
        gap = rel_gap(upper, lower);
        if (!(within_max_iter(t.iteration, warm_max) && !(gap <= p.tolerance))) break;
        for (int64_t j = 0; j < p.op_steps; ++j) reflect_step(&s, &t, eta, 0);
    }

    if (p.use_halpern) {
        double pn, dn;
        drift(&s, t.x, px, t.y, py, &pn, &dn);
        double om = dn / pn;
        t.omega = (isnan(om) || om <= 0.0) ? p.omega : om;
        t.theta = p.theta_initial;

        t.fp_error = canonical(&s, t.x, t.sx, t.y, t.sy, eta, t.omega);
        t.check_error = INFINITY;
        bounds(&s, t.x, t.y, &upper, &lower);
        gap = rel_gap(upper, lower);
        int terminated = !(within_max_iter(t.iteration, p.max_iter) && !(gap <= p.tolerance));
        memcpy(t.x0, t.x, sizeof(double) * n);
        memcpy(t.y0, t.y, sizeof(double) * 2 * m);
        t.restart_iter = 0;

        while (!terminated) {
            /* _halpern_step */
            double fp_prev = t.fp_error, check_prev = t.check_error;
            for (int64_t j = 0; j < p.op_steps; ++j) reflect_step(&s, &t, eta, 1);
            ++checks;
            double new_err = canonical(&s, t.x, t.sx, t.y, t.sy, eta, t.omega);
            bounds(&s, t.sx, t.sy, &upper, &lower);
            gap = rel_gap(upper, lower);
            terminated = !(within_max_iter(t.iteration, p.max_iter) && !(gap <= p.tolerance));

            int sufficient = new_err <= p.beta_sufficient * fp_prev;
            int necessary = (new_err <= p.beta_necessary * fp_prev) && (new_err > check_prev);
            int artificial = (double)t.restart_iter >= p.beta_artificial * (double)t.iteration;
            int restart = sufficient || necessary || artificial;

            if (terminated) {
                /* the returned point is the vanilla point the gap was measured at */
                memcpy(t.x, t.sx, sizeof(double) * n);
                memcpy(t.y, t.sy, sizeof(double) * 2 * m);
                break;
            }
            if (restart) {
                double new_omega = t.omega, new_theta = t.theta;
                drift(&s, t.x, t.x0, t.y, t.y0, &pn, &dn);
                if (pn > 1e-16 && dn > 1e-16) {
                    new_omega = exp(t.theta * (log(dn) - log(pn)) + (1.0 - t.theta) * log(t.omega));
                    new_theta = p.theta_target + p.theta_decay * (t.theta - p.theta_target);
                }
                t.omega = new_omega;
                t.theta = new_theta;
                /* one uncounted vanilla step from the step point: its start is the new anchor */
                int64_t ri = t.restart_iter, it = t.iteration;
                reflect_step(&s, &t, eta, 0);
                t.restart_iter = 0;
                t.iteration = it;
                (void)ri;
                memcpy(t.x0, t.x, sizeof(double) * n);
                memcpy(t.y0, t.y, sizeof(double) * 2 * m);
                t.fp_error = canonical(&s, t.x, t.sx, t.y, t.sy, eta, t.omega);
                t.check_error = INFINITY;
                ++restarts;
            } else {
                t.check_error = new_err;
            }
        }
    }

    memcpy(x_out, t.x, sizeof(double) * n);
    if (y_out) for (int64_t e = 0; e < 2 * m; ++e) y_out[e] = t.y[e] * s.scale;
    if (info) {
        info->iterations = t.iteration;
        info->restarts = restarts;
        info->checks = checks;
        info->lower = lower;
        info->upper = upper;
        info->upper_iterate = upper;
        info->relative_gap = gap;
        info->omega = t.omega;
        info->scale = s.scale;
        info->status = gap <= p.tolerance ? 0 : 1;
    }
    int status = gap <= p.tolerance ? 0 : 1;
    for (int a = 0; a < 4; ++a) free(*arrays[a]);
    for (int a = 0; a < 3; ++a) free(*duals[a]);
    free(px); free(py);
    praq_free(&s);
    return status;
}

void oracle_praq_bounds(int64_t n, int64_t m, const double* diag, const int64_t* rows,
                        const int64_t* cols, const double* vals, const double* x,
                        const double* y, double* lower, double* upper) {
    Praq s;
    if (praq_build(&s, n, m, diag, rows, cols, vals, 0)) {
        *lower = NAN; *upper = NAN; praq_free(&s); return;
    }
    /* bounds in instance units: unscale the structure, keep y as given */
    double* ys = malloc(sizeof(double) * (2 * m + 1));
    for (int64_t e = 0; e < 2 * m; ++e) ys[e] = y[e] / s.scale;
    bounds(&s, x, ys, upper, lower);
    free(ys);
    praq_free(&s);
}

/* One raw step and the canonical norm, exposed for the metric tests
 * (nonexpansiveness fixes the sign of the cross term). */
int oracle_praq_raw_step(int64_t n, int64_t m, const double* diag, const int64_t* rows,
                         const int64_t* cols, const double* vals, int dense_degree, double omega,
                         double eta, const double* x1, const double* y1, double* x2, double* y2) {
    Praq s;
    int st = praq_build(&s, n, m, diag, rows, cols, vals, dense_degree);
    if (st) { praq_free(&s); return st; }
    double* refl = malloc(sizeof(double) * n);
    raw_step(&s, x1, y1, omega, eta, x2, y2, refl);
    free(refl);
    praq_free(&s);
    return 0;
}

double oracle_praq_canonical(int64_t n, int64_t m, const double* diag, const int64_t* rows,
                             const int64_t* cols, const double* vals, int dense_degree,
                             const double* dx, const double* dy, double eta, double omega) {
    Praq s;
    if (praq_build(&s, n, m, diag, rows, cols, vals, dense_degree)) { praq_free(&s); return NAN; }
    double* zx = calloc(n, sizeof(double));
    double* zy = calloc(2 * m + 1, sizeof(double));
    double v = canonical(&s, dx, zx, dy, zy, eta, omega);
    free(zx); free(zy);
    praq_free(&s);
    return v;
}

/* The scaled structure's weights and covering flags, for tests that build
 * points of the dual set. */
int oracle_praq_structure(int64_t n, int64_t m, const double* diag, const int64_t* rows,
                          const int64_t* cols, const double* vals, double* w_out,
                          uint8_t* pos_out, double* scale_out) {
    Praq s;
    int st = praq_build(&s, n, m, diag, rows, cols, vals, 0);
    if (!st) {
        memcpy(w_out, s.w, sizeof(double) * m);
        memcpy(pos_out, s.pos, m);
        *scale_out = s.scale;
    }
    praq_free(&s);
    return st;
}
