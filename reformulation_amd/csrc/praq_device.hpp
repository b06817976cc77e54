// PRAQ device code: the per-node and per-edge halves of one PDHG step with
// the Halpern blend fused in, and the check-pass sums.  Used by both the
// one-workgroup whole-solve kernel and the device-wide kernels.
//
// Data layout in HBM (all arrays of one instance contiguous, 4-byte indices):
//   edges   e0[m], e1[m] int32 ends; w[m] signed scaled weight, 2 W_ij / s
//           (> 0 covering, < 0 packing, 0 inert)
//   nodes   diag[n], deg[n], xs[n] (projected start), x[n], sx[n], x0[n], refl[n]
//   CSR     ptr[n+1], half[2m]: node i's incident half-edges, each encoded as
//           slot * m + e, i.e. an index into a slot-major dual array
//   duals   ys[2m] (projected start), yc[2m], y[2m], sy[2m], y0[2m], slot-major (slot 0 of every edge,
//           then slot 1), so the edge pass reads both slots coalesced
//   hubs    heavy[nh] nodes of degree above heavy_degree, hchunk[nh + 1] their
//           first chunk, chunk_node[nc] the hub of each chunk of kPraqChunk
//           half-edges, chunk_sum[nc] FP64 chunk sums (device tier only)
//
// yc holds the Halpern-combined dual the NEXT step starts from, written by the
// edge pass as it finishes a step: the node pass's scatter, the one random
// access of the iteration, then reads one value per half-edge.  y and sy are
// written only on steps followed by a check or a restart.
//
// Reference semantics: _raw_step qubo/praq-jax/praq/solver.py:259-273,
// _reflect_step :276-292, project_dual qubo/praq-jax/praq/projection.py:16-42,
// bounds solver.py:320-337, canonical norm halpern_operator.py:7-35.
#pragma once

#include "common.hpp"
#include "halpern.hpp"

namespace reform {

template <class Real>
struct PraqArrays {
    int n, m;
    const int* e0;
    const int* e1;
    const Real* w;
    const int* ptr;
    const int* half;
    const Real* diag;
    const Real* deg;
    const Real* xs;   // projected start [n]
    const Real* ys;   // projected start [2m]
    Real* x;
    Real* sx;
    Real* x0;
    Real* refl;
    Real* yc;
    Real* y;
    Real* sy;
    Real* y0;
    // hubs: gathered by one wavefront per chunk in the device tier; nodes of
    // degree above heavy_degree are skipped by the lane-group passes (every
    // tier but the device tier binds heavy_degree = INT_MAX)
    int nh, nc, heavy_degree;
    const int* heavy;
    const int* hchunk;
    const int* chunk_node;
    double* chunk_sum;
};

// The device view of PraqArrays: the same arrays through global-memory
// pointers (see gptr); every device function works on a view.
template <class Real>
struct PraqView {
    int n, m;
    gptr<const int> e0, e1;
    gptr<const Real> w;
    gptr<const int> ptr, half;
    gptr<const Real> diag, deg, xs, ys;
    gptr<Real> x, sx, x0, refl, yc, y, sy, y0;
    int nh, nc, heavy_degree;
    gptr<const int> heavy, hchunk, chunk_node;
    gptr<double> chunk_sum;
};

template <class Real>
__device__ __forceinline__ PraqView<Real> praq_view(const PraqArrays<Real>& a) {
    PraqView<Real> v;
    v.n = a.n;
    v.m = a.m;
    v.e0 = to_global(a.e0);
    v.e1 = to_global(a.e1);
    v.w = to_global(a.w);
    v.ptr = to_global(a.ptr);
    v.half = to_global(a.half);
    v.diag = to_global(a.diag);
    v.deg = to_global(a.deg);
    v.xs = to_global(a.xs);
    v.ys = to_global(a.ys);
    v.x = to_global(a.x);
    v.sx = to_global(a.sx);
    v.x0 = to_global(a.x0);
    v.refl = to_global(a.refl);
    v.yc = to_global(a.yc);
    v.y = to_global(a.y);
    v.sy = to_global(a.sy);
    v.y0 = to_global(a.y0);
    v.nh = a.nh;
    v.nc = a.nc;
    v.heavy_degree = a.heavy_degree;
    v.heavy = to_global(a.heavy);
    v.hchunk = to_global(a.hchunk);
    v.chunk_node = to_global(a.chunk_node);
    v.chunk_sum = to_global(a.chunk_sum);
    return v;
}

// Half-edges per hub chunk: one wavefront, 16 per lane.
constexpr int kPraqChunk = 1024;

// The scalars of one step, computed once per step from the restart counter.
template <class Real>
struct PraqStep {
    Real lw, rw;      // Halpern weights of this step (reflected)
    Real lwn, rwn;    // weights the next step will use (for yc)
    Real tau, sigma;  // eta / omega, eta * omega
    bool reflect;     // Halpern step (else vanilla: start from sx, sy)
    bool store;       // write y, sy (a check or restart follows)
};

template <class Real>
__host__ __device__ inline PraqStep<Real> praq_step_scalars(int64_t restart_iter, double omega,
                                                            double eta, bool reflect, bool store) {
    PraqStep<Real> s;
    double lw, rw, lwn, rwn;
    halpern_weights(restart_iter, &lw, &rw);
    // after a reflected step the next one has counter + 1; after a vanilla
    // step the counter restarts at 0 with y1 as its anchor
    halpern_weights(reflect ? restart_iter + 1 : 0, &lwn, &rwn);
    s.lw = Real(lw);
    s.rw = Real(rw);
    s.lwn = Real(lwn);
    s.rwn = Real(rwn);
    Real re = Real(eta), ro = Real(omega);
    s.tau = re / ro;
    s.sigma = re * ro;
    s.reflect = reflect;
    s.store = store || !reflect;
    return s;
}

// Node i's update by a group of `lanes` consecutive lanes (a power of two,
// <= 64); all lanes of the group must call it (the sum is a shuffle).  The
// group's lanes stride the CSR list and meet in an xor butterfly, so the sum
// depends only on the instance and `lanes`.
template <class Real>
__device__ __forceinline__ void praq_node_update(const PraqView<Real>& a,
                                                 const PraqStep<Real>& st, int i, Real acc) {
    Real xc = a.x[i], sxi = a.sx[i];
    Real x1 = st.reflect ? st.lw * (Real(2) * sxi - xc) + st.rw * a.x0[i] : sxi;
    Real g = a.diag[i] - acc;
    Real x2 = clampr(x1 - st.tau * (g / a.deg[i]), Real(0), Real(1));
    a.x[i] = x1;
    a.sx[i] = x2;
    a.refl[i] = Real(2) * x2 - x1;
}

template <class Real>
__device__ __forceinline__ void praq_node(const PraqView<Real>& a, const PraqStep<Real>& st,
                                          int i, int lane, int lanes) {
    Real acc = Real(0);
    int begin = 0, end = 0;
    if (i < a.n) {
        begin = a.ptr[i];
        end = a.ptr[i + 1];
    }
    const bool hub = end - begin > a.heavy_degree;  // done by its chunks instead
    if (hub) end = begin;
    const gptr<const Real> src = st.reflect ? a.yc : a.sy;
    for (int k = begin + lane; k < end; k += lanes) acc += src[a.half[k]];
    acc = group_sum(acc, lanes);
    if (i >= a.n || lane != 0 || hub) return;
    praq_node_update(a, st, i, acc);
}

// One hub chunk by one wavefront (all 64 lanes call it): the FP64 sum of
// src over the chunk's half-edges, lanes striding, DPP reduction.
template <class Real>
__device__ __forceinline__ void praq_chunk(const PraqView<Real>& a, gptr<const Real> src, int c,
                                           int lane) {
    const int h = a.chunk_node[c], i = a.heavy[h];
    const int begin = a.ptr[i] + (c - a.hchunk[h]) * kPraqChunk;
    const int end = min(a.ptr[i + 1], begin + kPraqChunk);
    double acc = 0.0;
    for (int k = begin + lane; k < end; k += kWave) acc += double(src[a.half[k]]);
    acc = wave_sum(acc);
    if (lane == 0) a.chunk_sum[c] = acc;
}

// Hub h's total, its chunks added in order.
template <class Real>
__device__ __forceinline__ double praq_hub_total(const PraqView<Real>& a, int h) {
    double acc = 0.0;
    for (int c = a.hchunk[h]; c < a.hchunk[h + 1]; ++c) acc += a.chunk_sum[c];
    return acc;
}

// Edge e's update from its loaded values: projection of y1 + sigma (q -
// S refl), the next yc, and (y, sy) when storing; y2 (the projection) also
// into (o2a, o2b).
template <class Real>
__device__ __forceinline__ void praq_edge_update(const PraqView<Real>& a, const PraqStep<Real>& st,
                                                 int e, Real sw, Real y1a, Real y1b, Real ana,
                                                 Real anb, Real ra, Real rb, Real& o2a,
                                                 Real& o2b) {
    const int m = a.m;
    Real w = sw < Real(0) ? -sw : sw;
    bool covering = sw > Real(0);
    Real va = y1a + st.sigma * ((covering ? Real(1) : Real(0)) - ra);
    Real vb = y1b + st.sigma * (Real(0) - rb);
    Real y2a, y2b;
    if (covering) {
        Real mu = clampr(Real(0.5) * (va + vb), -w, Real(0));
        y2a = mu;
        y2b = mu;
    } else {
        Real line = va + Real(0.5) * (w - va - vb);
        y2a = clampr(line, Real(0), w);
        y2b = w - y2a;
    }
    if (st.store) {
        a.y[e] = y1a;
        a.y[m + e] = y1b;
        a.sy[e] = y2a;
        a.sy[m + e] = y2b;
    }
    a.yc[e] = st.lwn * (Real(2) * y2a - y1a) + st.rwn * ana;
    a.yc[m + e] = st.lwn * (Real(2) * y2b - y1b) + st.rwn * anb;
    o2a = y2a;
    o2b = y2b;
}

template <class Real>
__device__ __forceinline__ void praq_edge_update(const PraqView<Real>& a, const PraqStep<Real>& st,
                                                 int e, Real sw, Real y1a, Real y1b, Real ana,
                                                 Real anb, Real ra, Real rb) {
    Real o2a, o2b;
    praq_edge_update(a, st, e, sw, y1a, y1b, ana, anb, ra, rb, o2a, o2b);
}

// Edge e's update.
template <class Real>
__device__ __forceinline__ void praq_edge(const PraqView<Real>& a, const PraqStep<Real>& st,
                                          int e) {
    const int m = a.m;
    const gptr<const Real> y1 = st.reflect ? a.yc : a.sy;
    const gptr<const Real> an = st.reflect ? a.y0 : a.sy;
    praq_edge_update(a, st, e, a.w[e], y1[e], y1[m + e], an[e], an[m + e], a.refl[a.e0[e]],
                     a.refl[a.e1[e]]);
}

// K edges per thread, first, first + stride, ...: every load of the K edges
// (ends, then the reflected primal at both ends) is issued before any edge
// is updated, so a thread keeps its edges' latency chains in flight
// together; per edge the arithmetic of praq_edge.
template <class Real, int K>
__device__ __forceinline__ void praq_edges(const PraqView<Real>& a, const PraqStep<Real>& st,
                                           int first, int stride) {
    const int m = a.m;
    int e[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = first + k * stride;
        e[k] = q < m ? q : 0;   // past the end: edge 0, dropped
    }
    const gptr<const Real> y1 = st.reflect ? a.yc : a.sy;
    const gptr<const Real> an = st.reflect ? a.y0 : a.sy;
    int i0[K], i1[K];
    Real sw[K], ya[K], yb[K], aa[K], ab[K], ra[K], rb[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        i0[k] = a.e0[e[k]];
        i1[k] = a.e1[e[k]];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        sw[k] = a.w[e[k]];
        ya[k] = y1[e[k]];
        yb[k] = y1[m + e[k]];
        aa[k] = an[e[k]];
        ab[k] = an[m + e[k]];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        ra[k] = a.refl[i0[k]];
        rb[k] = a.refl[i1[k]];
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
        if (first + k * stride < m)
            praq_edge_update(a, st, e[k], sw[k], ya[k], yb[k], aa[k], ab[k], ra[k], rb[k]);
}

// The nine check sums, accumulated per thread.
enum PraqSum {
    kPrimalNorm = 0,   // sum deg (x - sx)^2
    kPrimalDrift = 1,  // sum deg (x - x0)^2
    kUpperNode = 2,    // sum diag bx
    kLowerNode = 3,    // sum min(diag - S^T by, 0)
    kDualNorm = 4,     // sum |y - sy|^2
    kCross = 5,        // sum <S (x - sx), y - sy>
    kDualDrift = 6,    // sum |y - y0|^2
    kUpperEdge = 7,    // sum w [covering: max(bi + bj - 1, 0); packing: -min(bi, bj)]
    kLowerEdge = 8,    // sum over covering edges of by slot 0
    kPraqSums = 9
};

// Node i's terms of the check sums given (S^T by)_i.
template <class Real>
__device__ __forceinline__ void praq_node_terms(const PraqView<Real>& a, int i, double scat,
                                                gptr<const Real> bx, double (&acc)[kPraqSums]) {
    double xi = a.x[i], sxi = a.sx[i], deg = a.deg[i], di = a.diag[i];
    double dx = xi - sxi, dr = xi - double(a.x0[i]);
    acc[kPrimalNorm] += deg * dx * dx;
    acc[kPrimalDrift] += deg * dr * dr;
    acc[kUpperNode] += di * double(bx[i]);
    double rc = di - scat;
    acc[kLowerNode] += rc < 0.0 ? rc : 0.0;
}

// Node part, by a lane group as praq_node; bounds at (bx, by).
template <class Real>
__device__ __forceinline__ void praq_node_sums(const PraqView<Real>& a, int i, int lane,
                                               int lanes, gptr<const Real> bx, gptr<const Real> by,
                                               double (&acc)[kPraqSums]) {
    double scat = 0.0;
    int begin = 0, end = 0;
    if (i < a.n) {
        begin = a.ptr[i];
        end = a.ptr[i + 1];
    }
    const bool hub = end - begin > a.heavy_degree;
    if (hub) end = begin;
    for (int k = begin + lane; k < end; k += lanes) scat += double(by[a.half[k]]);
    scat = group_sum(scat, lanes);
    if (i >= a.n || lane != 0 || hub) return;
    praq_node_terms(a, i, scat, bx, acc);
}

template <class Real>
__device__ __forceinline__ void praq_edge_sums(const PraqView<Real>& a, int e, gptr<const Real> bx,
                                               gptr<const Real> by, bool with_norms,
                                               double (&acc)[kPraqSums]) {
    const int m = a.m;
    int i = a.e0[e], j = a.e1[e];
    double sw = a.w[e];
    double w = fabs(sw);
    bool covering = sw > 0.0;
    if (with_norms) {
        double ya = a.y[e], yb = a.y[m + e];
        double da = ya - double(a.sy[e]), db = yb - double(a.sy[m + e]);
        acc[kDualNorm] += da * da + db * db;
        acc[kCross] += (double(a.x[i]) - double(a.sx[i])) * da
                     + (double(a.x[j]) - double(a.sx[j])) * db;
        double ra = ya - double(a.y0[e]), rb = yb - double(a.y0[m + e]);
        acc[kDualDrift] += ra * ra + rb * rb;
    }
    double bi = bx[i], bj = bx[j];
    if (covering) {
        double t = bi + bj - 1.0;
        acc[kUpperEdge] += w * (t > 0.0 ? t : 0.0);
        acc[kLowerEdge] += double(by[e]);
    } else {
        acc[kUpperEdge] += -w * (bi < bj ? bi : bj);
    }
}

__host__ __device__ inline CheckSums praq_check_sums(const double* s, double scale) {
    CheckSums c;
    c.pnorm = s[kPrimalNorm];
    c.dnorm = s[kDualNorm];
    c.cross = s[kCross];
    c.pdrift = s[kPrimalDrift];
    c.ddrift = s[kDualDrift];
    c.upper = (s[kUpperNode] + s[kUpperEdge]) * scale;
    c.lower = (s[kLowerNode] + s[kLowerEdge]) * scale;
    return c;
}

}  // namespace reform
