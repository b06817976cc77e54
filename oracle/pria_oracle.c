/* PRIA oracle: FP64 restatement of the JRP LP PDHG solve on plain
 * [N][T'][T'] arrays (T' = T padded to even).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows, function by function:
 *   padding, scaling, warm start, outputs
 *                            jrp/pria-jax/pria/solver.py:608-724
 *   double_simplex_projection (sort, cumsum, count-based rho, target <= 0
 *   gives 0)                 jrp/pria-jax/pria/projection.py:71-130
 *   x / y / z projections    jrp/pria-jax/pria/solver.py:405-406,
 *                            projection.py:133-190 (columns s <= t per (i,t);
 *                            rows t >= s per (i,s); per-s groups {(i,t>=s)})
 *   _raw_step                jrp/pria-jax/pria/solver.py:409-425
 *   _reflect_step            jrp/pria-jax/pria/solver.py:428-447
 *   _upper_bound             jrp/pria-jax/pria/solver.py:347-365
 *   _lower_bound             jrp/pria-jax/pria/solver.py:388-402
 *   canonical_norm (- cross) jrp/pria-jax/pria/halpern_operator.py:5-19
 *   _halpern_step            jrp/pria-jax/pria/solver.py:514-594
 *   primal_objective         jrp/pria-jax/pria/jrp_problem.py:143-161
 *   relaxation_bounds        jrp/pria-jax/pria/solver.py:740-790
 * The objective scale falls back to 1 for an all-zero instance, as the CUDA
 * port does (jrp/pria-cuda/tests/ref_impl.hpp:247-257); pria-jax divides by 0.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NCHUNK 64

typedef struct {
    int64_t n, t;          /* items, padded periods                            */
    int64_t sz;            /* n t t                                            */
    double *h, *k, k0;     /* scaled                                           */
    double scale;
} Pria;

static inline int64_t G(const Pria* s, int64_t i, int64_t a, int64_t b) {
    return (i * s->t + a) * s->t + b;
}

static int cmp_desc(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    return (x < y) - (x > y);
}

void oracle_simplex_project(int64_t len, double* v, double target) {
    if (len <= 0) return;
    double* u = malloc(sizeof(double) * len);
    memcpy(u, v, sizeof(double) * len);
    qsort(u, (size_t)len, sizeof(double), cmp_desc);
    /* rho is the COUNT of k with u_k k > cumsum_k - target, and theta uses
     * the prefix sum at rho - 1 (projection.py:113-128, count-based): with
     * entries tied at the threshold the test can hold off the prefix, and the
     * count still lands theta on the tie. */
    double* cum = malloc(sizeof(double) * len);
    double running = 0.0;
    int64_t rho = 0;
    for (int64_t k = 0; k < len; ++k) {
        running += u[k];
        cum[k] = running;
        if (u[k] * (double)(k + 1) > running - target) ++rho;
    }
    if (rho == 0) {
        for (int64_t k = 0; k < len; ++k) v[k] = 0.0;
    } else {
        double theta = (cum[rho - 1] - target) / (double)rho;
        for (int64_t k = 0; k < len; ++k) {
            double d = v[k] - theta;
            v[k] = d > 0.0 ? d : 0.0;
        }
    }
    free(cum);
    free(u);
}

static void proj_x(const Pria* s, double* x) {
#pragma omp parallel
    {
        double* seg = malloc(sizeof(double) * (s->t + 1));
#pragma omp for schedule(static) collapse(2)
        for (int64_t i = 0; i < s->n; ++i)
            for (int64_t t = 0; t < s->t; ++t) {
                for (int64_t a = 0; a <= t; ++a) seg[a] = x[G(s, i, a, t)];
                oracle_simplex_project(t + 1, seg, 1.0);
                for (int64_t a = 0; a <= t; ++a) x[G(s, i, a, t)] = seg[a];
            }
        free(seg);
    }
}

static void proj_y(const Pria* s, double* y) {
#pragma omp parallel for schedule(static) collapse(2)
    for (int64_t i = 0; i < s->n; ++i)
        for (int64_t a = 0; a < s->t; ++a)
            oracle_simplex_project(s->t - a, y + G(s, i, a, a), s->k[i]);
}

static void proj_z(const Pria* s, double* z) {
#pragma omp parallel
    {
        double* seg = malloc(sizeof(double) * (s->n * s->t + 1));
#pragma omp for schedule(dynamic, 1)
        for (int64_t a = 0; a < s->t; ++a) {
            int64_t p = 0;
            for (int64_t i = 0; i < s->n; ++i)
                for (int64_t b = a; b < s->t; ++b) seg[p++] = z[G(s, i, a, b)];
            oracle_simplex_project(p, seg, s->k0);
            p = 0;
            for (int64_t i = 0; i < s->n; ++i)
                for (int64_t b = a; b < s->t; ++b) z[G(s, i, a, b)] = seg[p++];
        }
        free(seg);
    }
}

/* _raw_step: x2 = P_X(x1 - tau (h + y1 + z1)); r = sigma (2 x2 - x1);
 * y2 = P_Y(y1 + r); z2 = P_Z(z1 + r).  Entries below the diagonal stay 0. */
static void raw_step(const Pria* s, const double* x1, const double* y1, const double* z1,
                     double omega, double eta, double* x2, double* y2, double* z2) {
    const double tau = eta / omega, sigma = eta * omega;
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < s->sz; ++g) x2[g] = x1[g] - tau * (s->h[g] + z1[g] + y1[g]);
    proj_x(s, x2);
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < s->sz; ++g) {
        double r = sigma * (2.0 * x2[g] - x1[g]);
        y2[g] = y1[g] + r;
        z2[g] = z1[g] + r;
    }
    proj_y(s, y2);
    proj_z(s, z2);
}


/* canonical_norm, negative cross term */
static double canonical(const Pria* s, const double* x, const double* sx, const double* y,
                        const double* sy, const double* z, const double* sz, double eta,
                        double omega) {
    double pa[NCHUNK], pb[NCHUNK], pc[NCHUNK];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < NCHUNK; ++c) {
        double a0 = 0, a1 = 0, a2 = 0;
        for (int64_t g = (s->sz * c) / NCHUNK; g < (s->sz * (c + 1)) / NCHUNK; ++g) {
            double dx = x[g] - sx[g], dy = y[g] - sy[g], dz = z[g] - sz[g];
            a0 += dx * dx;
            a1 += dy * dy + dz * dz;
            a2 += dx * dy + dx * dz;
        }
        pa[c] = a0; pb[c] = a1; pc[c] = a2;
    }
    double a0 = 0, a1 = 0, a2 = 0;
    for (int c = 0; c < NCHUNK; ++c) { a0 += pa[c]; a1 += pb[c]; a2 += pc[c]; }
    double v = (omega / eta) * a0 + (1.0 / (eta * omega)) * a1 - 2.0 * a2;
    return sqrt(v > 0.0 ? v : 0.0);
}

/* (upper, lower) on the padded problem at (x, y, z), original scale */
static void bounds(const Pria* s, const double* x, const double* y, const double* z,
                   double* up, double* lo) {
    const int64_t n = s->n, T = s->t;
    double* row = malloc(sizeof(double) * n * T);   /* per (i,s) */
    double* col = malloc(sizeof(double) * n * T);   /* per (i,t) */
    double* slc = malloc(sizeof(double) * T);       /* per s     */
#pragma omp parallel for schedule(static) collapse(2)
    for (int64_t i = 0; i < n; ++i)
        for (int64_t a = 0; a < T; ++a) {
            double m = -INFINITY;
            for (int64_t b = a; b < T; ++b) {
                double v = s->k[i] * x[G(s, i, a, b)];
                m = v > m ? v : m;
            }
            row[i * T + a] = m;
            double mn = INFINITY;
            for (int64_t q = 0; q <= a; ++q) {  /* column t = a */
                int64_t g = G(s, i, q, a);
                double v = s->h[g] + y[g] + z[g];
                mn = v < mn ? v : mn;
            }
            col[i * T + a] = mn;
        }
#pragma omp parallel for schedule(static)
    for (int64_t a = 0; a < T; ++a) {
        double m = -INFINITY;
        for (int64_t i = 0; i < n; ++i)
            for (int64_t b = a; b < T; ++b) {
                double v = s->k0 * x[G(s, i, a, b)];
                m = v > m ? v : m;
            }
        slc[a] = m;
    }
    double hx[NCHUNK];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < NCHUNK; ++c) {
        double acc = 0.0;
        for (int64_t g = (s->sz * c) / NCHUNK; g < (s->sz * (c + 1)) / NCHUNK; ++g)
            acc += x[g] * s->h[g];
        hx[c] = acc;
    }
    double t1 = 0, t2 = 0, t3 = 0, d = 0;
    for (int64_t q = 0; q < n * T; ++q) { t1 += row[q]; d += col[q]; }
    for (int64_t a = 0; a < T; ++a) t2 += slc[a];
    for (int c = 0; c < NCHUNK; ++c) t3 += hx[c];
    *up = (t1 + t2 + t3) * s->scale;
    *lo = d * s->scale;
    free(row); free(col); free(slc);
}

static double rel_gap(double upper, double lower) {
    return (upper - lower) / (1.0 + fabs(lower) + fabs(upper));
}

static void drift(const Pria* s, const double* x, const double* x0, const double* y,
                  const double* y0, const double* z, const double* z0, double* pn, double* dn) {
    double pa[NCHUNK], pb[NCHUNK];
#pragma omp parallel for schedule(static)
    for (int c = 0; c < NCHUNK; ++c) {
        double a = 0, b = 0;
        for (int64_t g = (s->sz * c) / NCHUNK; g < (s->sz * (c + 1)) / NCHUNK; ++g) {
            double dx = x[g] - x0[g], dy = y[g] - y0[g], dz = z[g] - z0[g];
            a += dx * dx;
            b += dy * dy + dz * dz;
        }
        pa[c] = a; pb[c] = b;
    }
    double a = 0, b = 0;
    for (int c = 0; c < NCHUNK; ++c) { a += pa[c]; b += pb[c]; }
    *pn = sqrt(a);
    *dn = sqrt(b);
}

double oracle_pria_primal_objective(int64_t n, int64_t t, const double* k, double k0,
                                    const double* h, const double* x) {
    double obj = 0.0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t a = 0; a < t; ++a) {
            double m = -INFINITY;
            for (int64_t b = 0; b < t; ++b) {
                double v = x[(i * t + a) * t + b];
                m = v > m ? v : m;
            }
            obj += k[i] * m;
        }
    for (int64_t a = 0; a < t; ++a) {
        double m = -INFINITY;
        for (int64_t i = 0; i < n; ++i)
            for (int64_t b = 0; b < t; ++b) {
                double v = x[(i * t + a) * t + b];
                m = v > m ? v : m;
            }
        obj += k0 * m;
    }
    for (int64_t i = 0; i < n; ++i)
        for (int64_t a = 0; a < t; ++a)
            for (int64_t b = a; b < t; ++b)
                obj += x[(i * t + a) * t + b] * h[(i * t + a) * t + b];
    return obj;
}

double oracle_pria_dual_objective(int64_t n, int64_t t, const double* k, double k0,
                                  const double* h, const double* y, const double* z) {
    double total = 0.0;
    for (int64_t i = 0; i < n; ++i)
        for (int64_t b = 0; b < t; ++b) {
            double m = INFINITY;
            for (int64_t a = 0; a <= b; ++a) {
                int64_t g = (i * t + a) * t + b;
                double v = h[g] + k[i] * y[g] + k0 * z[g];
                m = v < m ? v : m;
            }
            total += m;
        }
    return total;
}

typedef struct {
    double *x, *y, *z, *sx, *sy, *sz, *x0, *y0, *z0;
    double omega, theta, fp_error, check_error;
    int64_t restart_iter, iteration;
} State;

static void blend(int64_t len, double* cur, const double* step, const double* anchor,
                  double lw, double rw) {
#pragma omp parallel for schedule(static)
    for (int64_t g = 0; g < len; ++g) cur[g] = lw * (2.0 * step[g] - cur[g]) + rw * anchor[g];
}

static void reflect_step(const Pria* s, State* t, double eta, int reflect) {
    if (reflect) {
        double k = (double)(t->restart_iter + 1);
        double lw = k / (k + 1.0), rw = 1.0 / (k + 1.0);
        blend(s->sz, t->x, t->sx, t->x0, lw, rw);
        blend(s->sz, t->y, t->sy, t->y0, lw, rw);
        blend(s->sz, t->z, t->sz, t->z0, lw, rw);
    } else {
        memcpy(t->x, t->sx, sizeof(double) * s->sz);
        memcpy(t->y, t->sy, sizeof(double) * s->sz);
        memcpy(t->z, t->sz, sizeof(double) * s->sz);
    }
    raw_step(s, t->x, t->y, t->z, t->omega, eta, t->sx, t->sy, t->sz);
    t->restart_iter += 1;
    t->iteration += 1;
}

static int within_max_iter(int64_t iteration, int64_t max_iter) {
    return iteration < max_iter || max_iter < 0;
}

int oracle_pria_solve(int64_t n, int64_t torig, const double* kin, double k0in,
                      const double* hin, const OracleParams* pin, const double* warm_x,
                      const double* warm_y, const double* warm_z, double* x_out,
                      double* y_out, double* z_out, OracleInfo* info) {
    if (!pin || n <= 0 || torig <= 0 || !kin || !hin) return -1;
    OracleParams p = *pin;
    if (p.op_steps < 2) p.op_steps = 2;
    if (p.warm_steps < 0) p.warm_steps = 0;
    const double eta = p.eta;

    Pria s;
    s.n = n;
    s.t = torig + (torig & 1);
    s.sz = n * s.t * s.t;

    double sq = k0in * k0in;
    for (int64_t i = 0; i < n; ++i) sq += kin[i] * kin[i];
    for (int64_t i = 0; i < n; ++i)
        for (int64_t a = 0; a < torig; ++a)
            for (int64_t b = a; b < torig; ++b) {
                double v = hin[(i * torig + a) * torig + b];
                sq += v * v;
            }
    s.scale = sq > 0.0 ? sqrt(sq) : 1.0;
    s.h = calloc(s.sz, sizeof(double));
    s.k = malloc(sizeof(double) * n);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t a = 0; a < torig; ++a)
            for (int64_t b = a; b < torig; ++b)
                s.h[G(&s, i, a, b)] = hin[(i * torig + a) * torig + b] / s.scale;
    for (int64_t i = 0; i < n; ++i) s.k[i] = kin[i] / s.scale;
    s.k0 = k0in / s.scale;

    State t;
    memset(&t, 0, sizeof(t));
    double** arrays[] = {&t.x, &t.y, &t.z, &t.sx, &t.sy, &t.sz, &t.x0, &t.y0, &t.z0};
    for (int a = 0; a < 9; ++a) *arrays[a] = calloc(s.sz, sizeof(double));
    double* px = calloc(s.sz, sizeof(double));
    double* py = calloc(s.sz, sizeof(double));
    double* pz = calloc(s.sz, sizeof(double));

    /* warm start: normalised duals scaled back onto the K_i / K0 simplices */
    for (int64_t i = 0; i < n; ++i)
        for (int64_t a = 0; a < torig; ++a)
            for (int64_t b = a; b < torig; ++b) {
                int64_t g = G(&s, i, a, b), go = (i * torig + a) * torig + b;
                if (warm_x) px[g] = warm_x[go];
                if (warm_y) py[g] = warm_y[go] * s.k[i];
                if (warm_z) pz[g] = warm_z[go] * s.k0;
            }
    proj_x(&s, px);
    proj_y(&s, py);
    proj_z(&s, pz);
    memcpy(t.x, px, sizeof(double) * s.sz);
    memcpy(t.y, py, sizeof(double) * s.sz);
    memcpy(t.z, pz, sizeof(double) * s.sz);
    memcpy(t.x0, px, sizeof(double) * s.sz);
    memcpy(t.y0, py, sizeof(double) * s.sz);
    memcpy(t.z0, pz, sizeof(double) * s.sz);
    t.omega = p.omega;
    t.theta = p.theta_initial;
    raw_step(&s, t.x, t.y, t.z, t.omega, eta, t.sx, t.sy, t.sz);
    t.check_error = INFINITY;

    double upper = 0, lower = 0, gap = INFINITY;
    int64_t restarts = 0, checks = 0;

    int64_t warm_max = p.use_halpern ? p.op_steps * p.warm_steps : p.max_iter;
    for (;;) {
        bounds(&s, t.x, t.y, t.z, &upper, &lower);
        gap = rel_gap(upper, lower);
        if (!(within_max_iter(t.iteration, warm_max) && !(gap <= p.tolerance))) break;
        for (int64_t j = 0; j < p.op_steps; ++j) reflect_step(&s, &t, eta, 0);
    }

    if (p.use_halpern) {
        double pn, dn;
        drift(&s, t.x, px, t.y, py, t.z, pz, &pn, &dn);
        double om = dn / pn;
        t.omega = isnan(om) ? p.omega : om;   /* PRIA falls back on NaN only */
        t.theta = p.theta_initial;
        t.fp_error = canonical(&s, t.x, t.sx, t.y, t.sy, t.z, t.sz, eta, t.omega);
        t.check_error = INFINITY;
        bounds(&s, t.x, t.y, t.z, &upper, &lower);
        gap = rel_gap(upper, lower);
        int terminated = !(within_max_iter(t.iteration, p.max_iter) && !(gap <= p.tolerance));
        memcpy(t.x0, t.x, sizeof(double) * s.sz);
        memcpy(t.y0, t.y, sizeof(double) * s.sz);
        memcpy(t.z0, t.z, sizeof(double) * s.sz);
        t.restart_iter = 0;

        while (!terminated) {
            double fp_prev = t.fp_error, check_prev = t.check_error;
            for (int64_t j = 0; j < p.op_steps; ++j) reflect_step(&s, &t, eta, 1);
            ++checks;
            double new_err = canonical(&s, t.x, t.sx, t.y, t.sy, t.z, t.sz, eta, t.omega);
            bounds(&s, t.sx, t.sy, t.sz, &upper, &lower);
            gap = rel_gap(upper, lower);
            terminated = !(within_max_iter(t.iteration, p.max_iter) && !(gap <= p.tolerance));

            if (getenv("ORACLE_TRACE"))
                fprintf(stderr, "[check] it=%lld r=%lld gap=%.6e up=%.12e lo=%.12e new_err=%.12e fp=%.12e check=%.12e omega=%.12e\n",
                        (long long)t.iteration, (long long)t.restart_iter, gap, upper, lower, new_err,
                        fp_prev, check_prev, t.omega);
            int sufficient = new_err <= p.beta_sufficient * fp_prev;
            int necessary = (new_err <= p.beta_necessary * fp_prev) && (new_err > check_prev);
            int artificial = (double)t.restart_iter >= p.beta_artificial * (double)t.iteration;
            int restart = sufficient || necessary || artificial;

            if (terminated) {
                memcpy(t.x, t.sx, sizeof(double) * s.sz);
                memcpy(t.y, t.sy, sizeof(double) * s.sz);
                memcpy(t.z, t.sz, sizeof(double) * s.sz);
                break;
            }
            if (restart) {
                double new_omega = t.omega, new_theta = t.theta;
                drift(&s, t.x, t.x0, t.y, t.y0, t.z, t.z0, &pn, &dn);
                if (pn > 1e-16 && dn > 1e-16) {
                    new_omega = exp(t.theta * (log(dn) - log(pn)) + (1.0 - t.theta) * log(t.omega));
                    new_theta = p.theta_target + p.theta_decay * (t.theta - p.theta_target);
                }
                t.omega = new_omega;
                t.theta = new_theta;
                int64_t it = t.iteration;
                reflect_step(&s, &t, eta, 0);
                t.restart_iter = 0;
                t.iteration = it;
                memcpy(t.x0, t.x, sizeof(double) * s.sz);
                memcpy(t.y0, t.y, sizeof(double) * s.sz);
                memcpy(t.z0, t.z, sizeof(double) * s.sz);
                t.fp_error = canonical(&s, t.x, t.sx, t.y, t.sy, t.z, t.sz, eta, t.omega);
                t.check_error = INFINITY;
                ++restarts;
            } else {
                t.check_error = new_err;
            }
        }
    }

    /* the solver's own lower bound at the returned duals (padded problem) */
    double lower_final, upper_padded;
    bounds(&s, t.x, t.y, t.z, &upper_padded, &lower_final);

    /* outputs: normalise duals, unpad */
    double* xo = malloc(sizeof(double) * n * torig * torig);
    for (int64_t i = 0; i < n; ++i)
        for (int64_t a = 0; a < torig; ++a)
            for (int64_t b = 0; b < torig; ++b) {
                int64_t g = G(&s, i, a, b), go = (i * torig + a) * torig + b;
                double xv = a <= b ? t.x[g] : 0.0;
                double yv = a <= b ? (s.k[i] > 0.0 ? t.y[g] / s.k[i] : 0.0) : 0.0;
                double zv = a <= b ? (s.k0 > 0.0 ? t.z[g] / s.k0 : 0.0) : 0.0;
                xo[go] = xv;
                if (x_out) x_out[go] = xv;
                if (y_out) y_out[go] = yv;
                if (z_out) z_out[go] = zv;
            }
    double upper_final = oracle_pria_primal_objective(n, torig, kin, k0in, hin, xo);
    free(xo);

    if (info) {
        info->iterations = t.iteration;
        info->restarts = restarts;
        info->checks = checks;
        info->lower = lower_final;
        info->upper = upper_final;
        info->upper_iterate = upper_padded;
        info->relative_gap = gap;
        info->omega = t.omega;
        info->scale = s.scale;
        info->status = gap <= p.tolerance ? 0 : 1;
    }
    int status = gap <= p.tolerance ? 0 : 1;
    for (int a = 0; a < 9; ++a) free(*arrays[a]);
    free(px); free(py); free(pz); free(s.h); free(s.k);
    (void)upper;
    return status;
}

/* One raw step on an even-T instance whose h, k, k0 are already scaled;
 * dense [n][t][t] in and out. */
int oracle_pria_raw_step(int64_t n, int64_t t, const double* k, double k0, const double* h,
                         const double* x1, const double* y1, const double* z1, double omega,
                         double eta, double* x2, double* y2, double* z2) {
    if (n <= 0 || t <= 0 || (t & 1)) return -1;
    Pria s;
    s.n = n;
    s.t = t;
    s.sz = n * t * t;
    s.scale = 1.0;
    s.h = (double*)h;
    s.k = (double*)k;
    s.k0 = k0;
    raw_step(&s, x1, y1, z1, omega, eta, x2, y2, z2);
    return 0;
}
