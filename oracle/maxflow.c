/* The roof dual of a QUBO by max-flow -- an independent value for PRAQ at
 * sizes HiGHS does not reach.  TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * The Hammer-Hansen-Simeone implication network: the objective
 * sum d_i x_i + sum_e 2 W_e x_i x_j is rewritten as a posiform (a constant
 * plus nonnegative terms a u v over literals u, v, where a term a u is
 * a u * 1): w x_i x_j with w < 0 becomes |w| x_i (1 - x_j) - |w| x_i, a
 * negative linear coefficient a x_i becomes a + |a| (1 - x_i).  Each term
 * a u v gives arcs (not u) -> v and (not v) -> u of capacity a / 2 (the
 * literal "1" is the sink, its negation the source), and the roof dual -- the
 * value of the PRAQ relaxation -- is the constant plus the minimum cut
 * (summary of the network at qubo/baselines/maxflow_backend.py:19-60; the
 * same construction as tests/lp_models.py, there on scipy's int32 flow).
 *
 * Capacities are integers: a / 2 * 2 scale rounded to int64, so the cut is
 * exact for weights on the 1 / scale grid and within (arcs in the cut) /
 * (4 scale) of the real cut otherwise.  The flow is Dinic's algorithm
 * (level graph by BFS, blocking flow by DFS with current-arc pointers,
 * iterative), single threaded.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int64_t n_nodes, n_arcs;
    int64_t* head;     /* [n_nodes + 1] first arc of each node (CSR) */
    int32_t* to;       /* [arcs] */
    int64_t* rev;      /* [arcs] index of the paired arc */
    int64_t* cap;      /* [arcs] residual capacity */
} Net;

typedef struct {
    int32_t u, v;
    int64_t c;
} Arc;

static void add_term(Arc* arcs, int64_t* k, int32_t nu, int32_t v, int32_t nv, int32_t u,
                     int64_t c) {
    if (c <= 0) return;
    arcs[(*k)++] = (Arc){nu, v, c};
    arcs[(*k)++] = (Arc){nv, u, c};
}

/* CSR with a reverse arc (capacity 0) for every arc. */
static int build(Net* g, int64_t nodes, const Arc* arcs, int64_t count) {
    g->n_nodes = nodes;
    g->n_arcs = 2 * count;
    g->head = calloc((size_t)nodes + 1, sizeof(int64_t));
    g->to = malloc(sizeof(int32_t) * (size_t)g->n_arcs);
    g->rev = malloc(sizeof(int64_t) * (size_t)g->n_arcs);
    g->cap = malloc(sizeof(int64_t) * (size_t)g->n_arcs);
    int64_t* fill = malloc(sizeof(int64_t) * (size_t)nodes);
    if (!g->head || !g->to || !g->rev || !g->cap || !fill) {
        free(fill);
        return -1;
    }
    for (int64_t a = 0; a < count; ++a) {
        g->head[arcs[a].u + 1]++;
        g->head[arcs[a].v + 1]++;
    }
    for (int64_t q = 0; q < nodes; ++q) g->head[q + 1] += g->head[q];
    memcpy(fill, g->head, sizeof(int64_t) * (size_t)nodes);
    for (int64_t a = 0; a < count; ++a) {
        const int64_t p = fill[arcs[a].u]++, q = fill[arcs[a].v]++;
        g->to[p] = arcs[a].v;
        g->cap[p] = arcs[a].c;
        g->rev[p] = q;
        g->to[q] = arcs[a].u;
        g->cap[q] = 0;
        g->rev[q] = p;
    }
    free(fill);
    return 0;
}

static void release(Net* g) {
    free(g->head);
    free(g->to);
    free(g->rev);
    free(g->cap);
}

static int64_t dinic(Net* g, int32_t s, int32_t t) {
    const int64_t N = g->n_nodes;
    int32_t* level = malloc(sizeof(int32_t) * (size_t)N);
    int32_t* queue = malloc(sizeof(int32_t) * (size_t)N);
    int64_t* cur = malloc(sizeof(int64_t) * (size_t)N);
    int32_t* stack_node = malloc(sizeof(int32_t) * (size_t)N);
    int64_t* stack_arc = malloc(sizeof(int64_t) * (size_t)N);
    int64_t total = 0;
    if (!level || !queue || !cur || !stack_node || !stack_arc) goto done;
    for (;;) {
        for (int64_t q = 0; q < N; ++q) level[q] = -1;
        int64_t qh = 0, qt = 0;
        level[s] = 0;
        queue[qt++] = s;
        while (qh < qt) {
            const int32_t u = queue[qh++];
            for (int64_t a = g->head[u]; a < g->head[u + 1]; ++a)
                if (g->cap[a] > 0 && level[g->to[a]] < 0) {
                    level[g->to[a]] = level[u] + 1;
                    queue[qt++] = g->to[a];
                }
        }
        if (level[t] < 0) break;
        memcpy(cur, g->head, sizeof(int64_t) * (size_t)N);
        /* blocking flow: walk s -> t along level-increasing arcs; at t push
         * the bottleneck back along the path and restart from the arc
         * that saturated */
        int64_t depth = 0;
        stack_node[0] = s;
        while (depth >= 0) {
            const int32_t u = stack_node[depth];
            if (u == t) {
                int64_t push = INT64_MAX;
                for (int64_t d = 0; d < depth; ++d)
                    if (g->cap[stack_arc[d]] < push) push = g->cap[stack_arc[d]];
                int64_t back = depth;
                for (int64_t d = depth - 1; d >= 0; --d) {
                    const int64_t a = stack_arc[d];
                    g->cap[a] -= push;
                    g->cap[g->rev[a]] += push;
                    if (g->cap[a] == 0) back = d;
                }
                total += push;
                depth = back;
                continue;
            }
            int64_t a = cur[u];
            for (; a < g->head[u + 1]; ++a)
                if (g->cap[a] > 0 && level[g->to[a]] == level[u] + 1) break;
            cur[u] = a;
            if (a == g->head[u + 1]) {  /* dead end: retreat */
                level[u] = -1;
                --depth;
                if (depth >= 0) cur[stack_node[depth]]++;
                continue;
            }
            stack_arc[depth] = a;
            stack_node[++depth] = g->to[a];
        }
    }
done:
    free(level);
    free(queue);
    free(cur);
    free(stack_node);
    free(stack_arc);
    return total;
}

double oracle_praq_roof_dual(int64_t n, const double* diag, int64_t m, const int64_t* rows,
                             const int64_t* cols, const double* vals, double scale,
                             int64_t* flow_out) {
    if (n <= 0 || n > (INT32_MAX - 2) / 2 || m < 0 || !(scale > 0.0)) return NAN;
    const int32_t src = (int32_t)(2 * n), snk = (int32_t)(2 * n + 1);
    /* literal k: x_k; k + n: its negation; the sink stands for "1" */
#define NEG(x) ((x) == snk ? src : ((x) < n ? (int32_t)((x) + n) : (int32_t)((x) - n)))
    Arc* arcs = malloc(sizeof(Arc) * (size_t)(2 * (m + n)));
    double* linear = malloc(sizeof(double) * (size_t)n);
    if (!arcs || !linear) {
        free(arcs);
        free(linear);
        return NAN;
    }
    memcpy(linear, diag, sizeof(double) * (size_t)n);
    double constant = 0.0;
    int64_t k = 0;
    for (int64_t e = 0; e < m; ++e) {
        const double w = 2.0 * vals[e];
        const int32_t i = (int32_t)rows[e], j = (int32_t)cols[e];
        if (w > 0.0) {         /* w x_i x_j */
            add_term(arcs, &k, NEG(i), j, NEG(j), i, llround(w * scale));
        } else if (w < 0.0) {  /* |w| x_i (1 - x_j) - |w| x_i */
            const int32_t nj = (int32_t)(j + n);
            add_term(arcs, &k, NEG(i), nj, NEG(nj), i, llround(-w * scale));
            linear[i] += w;
        }
    }
    for (int64_t q = 0; q < n; ++q) {
        const double a = linear[q];
        const int32_t x = (int32_t)q, nx = (int32_t)(q + n);
        if (a > 0.0) {         /* a x */
            add_term(arcs, &k, NEG(x), snk, NEG(snk), x, llround(a * scale));
        } else if (a < 0.0) {  /* a + |a| (1 - x) */
            constant += a;
            add_term(arcs, &k, NEG(nx), snk, NEG(snk), nx, llround(-a * scale));
        }
    }
#undef NEG
    free(linear);
    Net g;
    memset(&g, 0, sizeof(g));
    double value = NAN;
    if (build(&g, 2 * n + 2, arcs, k) == 0) {
        const int64_t flow = dinic(&g, src, snk);
        if (flow_out) *flow_out = flow;
        /* arcs carry a / 2 of each term, stored as a * scale: the cut in
         * original units is flow / (2 scale) */
        value = constant + (double)flow / (2.0 * scale);
    }
    release(&g);
    free(arcs);
    return value;
}
