// PRAQ on gfx950: kernels, the two execution tiers and the C ABI of
// include/praq_hip.h.
//
//   workgroup tier  praq_wg_solve: one workgroup runs a whole solve -- steps,
//                   check sums and the restart/omega/termination logic -- in
//                   one launch, with only workgroup barriers.  One workgroup
//                   per instance is also how the batch runs many instances in
//                   one launch.
//   device tier     per step one node kernel and one edge kernel over the
//                   whole GPU; per check a two-stage FP64 reduction whose nine
//                   sums feed the same scalar logic on the host.
//
// Reference: qubo/praq-cuda/src/solver.cuh:227-800 (kernels and host loop),
// qubo/praq-cuda/src/praq.cu:8-79 (entry points and validation).
#include "praq_hip.h"
#include "praq_device.hpp"
#include "praq_host.hpp"

#include <algorithm>
#include <cstdint>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

namespace reform {

// ---------------------------------------------------------------------------
// workgroup tier
// ---------------------------------------------------------------------------

template <class Real>
__device__ void praq_wg_step(const PraqView<Real>& a, const PraqStep<Real>& st, int lanes) {
    const int tid = threadIdx.x, nthreads = blockDim.x;
    const int groups = nthreads / lanes, g = tid / lanes, lane = tid % lanes;
    for (int base = 0; base < a.n; base += groups) praq_node(a, st, base + g, lane, lanes);
    __syncthreads();
    for (int e = tid; e < a.m; e += nthreads) praq_edge(a, st, e);
    __syncthreads();
}

template <class Real>
__device__ void praq_wg_sums(const PraqView<Real>& a, int lanes, gptr<const Real> bx,
                             gptr<const Real> by, double* red, double (&out)[kPraqSums]) {
    const int tid = threadIdx.x, nthreads = blockDim.x;
    const int groups = nthreads / lanes, g = tid / lanes, lane = tid % lanes;
    double acc[kPraqSums];
#pragma unroll
    for (int k = 0; k < kPraqSums; ++k) acc[k] = 0.0;
    for (int base = 0; base < a.n; base += groups) praq_node_sums(a, base + g, lane, lanes, bx, by, acc);
    for (int e = tid; e < a.m; e += nthreads) praq_edge_sums(a, e, bx, by, true, acc);
    block_sums<kPraqSums>(acc, red);
#pragma unroll
    for (int k = 0; k < kPraqSums; ++k) out[k] = acc[k];
}

template <class Real>
__device__ void praq_wg_copy(gptr<Real> dx, gptr<const Real> sx, int n, gptr<Real> dy,
                             gptr<const Real> sy, int m2) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) dx[i] = sx[i];
    for (int e = threadIdx.x; e < m2; e += blockDim.x) dy[e] = sy[e];
    __syncthreads();
}

template <class Real>
__global__ void __launch_bounds__(1024) praq_wg_solve(const PraqJob<Real>* jobs) {
    __shared__ double red[kPraqSums * 16];
    __shared__ LogicState st;
    __shared__ PraqJob<Real> job;
    if (threadIdx.x == 0) job = jobs[blockIdx.x];
    __syncthreads();
    const PraqView<Real> a = praq_view(job.a);
    const LogicParams p = job.p;
    const int lanes = job.lanes;
    const long long t_start = wall_clock64();

    double sums[kPraqSums];
    if (job.resume) {
        if (threadIdx.x == 0) {
            st = *job.state;
            st.stalled = 0;
        }
        __syncthreads();
    } else {
        // state from the projected start
        praq_wg_copy(a.x0, a.xs, a.n, a.y0, a.ys, 2 * a.m);
        praq_wg_copy(a.x, a.xs, a.n, a.y, a.ys, 2 * a.m);
        praq_wg_copy(a.sx, a.xs, a.n, a.sy, a.ys, 2 * a.m);
        if (threadIdx.x == 0) logic_init(st, p);
        __syncthreads();
        // the start: one uncounted vanilla step, check at (x, y)
        praq_wg_step(a, praq_step_scalars<Real>(0, p.omega, p.eta, false, true), lanes);
        praq_wg_sums(a, lanes, a.x, a.y, red, sums);
        if (threadIdx.x == 0) logic_after_warm_check(st, p, praq_check_sums(sums, job.scale));
        __syncthreads();
    }

    for (;;) {
        const int action = st.action;
        const int anchor_pending = st.anchor_pending;
        const int stalled = st.stalled;
        const int64_t r0 = st.restart_iter;
        const double omega = st.omega;
        __syncthreads();
        if (stalled) break;  // the state is handed to a higher precision
        if (action == kActDoneX) break;
        if (action == kActDoneSX) {
            praq_wg_copy(a.x, a.sx, a.n, a.y, a.sy, 2 * a.m);
            break;
        }
        if (action == kActWarmBlock) {
            for (int64_t j = 0; j < p.op_steps; ++j)
                praq_wg_step(a, praq_step_scalars<Real>(r0 + j, omega, p.eta, false, true), lanes);
            praq_wg_sums(a, lanes, a.x, a.y, red, sums);
            if (threadIdx.x == 0) {
                st.restart_iter += p.op_steps;
                st.iteration += p.op_steps;
                logic_after_warm_check(st, p, praq_check_sums(sums, job.scale));
            }
        } else if (action == kActHalpernBlock) {
            if (anchor_pending) {
                // the last step was vanilla, whose yc already took the current
                // point as the anchor; only the anchor arrays need the copy
                praq_wg_copy(a.x0, a.x, a.n, a.y0, a.y, 2 * a.m);
                if (threadIdx.x == 0) st.anchor_pending = 0;
            }
            for (int64_t j = 0; j < p.op_steps; ++j)
                praq_wg_step(a, praq_step_scalars<Real>(r0 + j, omega, p.eta, true,
                                                        j + 1 == p.op_steps), lanes);
            praq_wg_sums(a, lanes, a.sx, a.sy, red, sums);
            if (threadIdx.x == 0) {
                st.restart_iter += p.op_steps;
                st.iteration += p.op_steps;
                logic_after_halpern_check(st, p, praq_check_sums(sums, job.scale));
            }
        } else {  // kActRestart
            praq_wg_step(a, praq_step_scalars<Real>(0, omega, p.eta, false, true), lanes);
            praq_wg_copy(a.x0, a.x, a.n, a.y0, a.y, 2 * a.m);
            praq_wg_sums(a, lanes, a.sx, a.sy, red, sums);
            if (threadIdx.x == 0) logic_after_restart(st, p, praq_check_sums(sums, job.scale));
        }
        __syncthreads();
        if (threadIdx.x == 0 && job.time_limit_ticks > 0 &&
            wall_clock64() - t_start > job.time_limit_ticks &&
            (st.action == kActHalpernBlock || st.action == kActRestart || st.action == kActWarmBlock)) {
            // stop at the point the last bounds were taken at
            st.action = action == kActWarmBlock ? kActDoneX : kActDoneSX;
            if (action == kActRestart) st.action = kActDoneX;
            st.status = 2;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *job.state = st;
}

// ---------------------------------------------------------------------------
// device tier
// ---------------------------------------------------------------------------

template <class Real>
__global__ void praq_node_kernel(PraqArrays<Real> a, PraqStep<Real> st, int lanes) {
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    praq_node(praq_view(a), st, gid / lanes, gid % lanes, lanes);
}

// hub chunks, one wavefront each: FP64 sums of src into chunk_sum
template <class Real>
__global__ void praq_chunk_kernel(PraqArrays<Real> a, const Real* src) {
    const int c = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (c < a.nc) praq_chunk(praq_view(a), to_global(src), c, threadIdx.x & (kWave - 1));
}

// hubs, one thread each: the chunk sums in order, then the node update
template <class Real>
__global__ void praq_hub_kernel(PraqArrays<Real> a, PraqStep<Real> st) {
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    const PraqView<Real> v = praq_view(a);
    if (h < a.nh) praq_node_update(v, st, a.heavy[h], Real(praq_hub_total(v, h)));
}

// the hubs' terms of the check sums into one more partial row
template <class Real>
__global__ void praq_hub_sums_kernel(PraqArrays<Real> a, const Real* bx, double* row) {
    __shared__ double red[kPraqSums * 16];
    double acc[kPraqSums];
#pragma unroll
    for (int k = 0; k < kPraqSums; ++k) acc[k] = 0.0;
    const PraqView<Real> v = praq_view(a);
    for (int h = threadIdx.x; h < a.nh; h += blockDim.x)
        praq_node_terms(v, a.heavy[h], praq_hub_total(v, h), to_global(bx), acc);
    block_sums<kPraqSums>(acc, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < kPraqSums; ++k) row[k] = acc[k];
}

// Edges per thread of the device tier's edge kernel (their loads in flight
// together).
constexpr int kEdgesPerThread = 4;

template <class Real>
__global__ void praq_edge_kernel(PraqArrays<Real> a, PraqStep<Real> st) {
    praq_edges<Real, kEdgesPerThread>(praq_view(a), st,
                                      blockIdx.x * blockDim.x * kEdgesPerThread + threadIdx.x,
                                      blockDim.x);
}

// (refl, x, sx) of every node in one record of four values aligned to
// its size, for the check step's edge kernel: one gather per edge end
// brings all three.
template <class Real>
__global__ void praq_quad_kernel(PraqArrays<Real> a, Real* quad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const PraqView<Real> v = praq_view(a);
    const gptr<Real> q = to_global(quad) + 4 * size_t(i);
    q[0] = v.refl[i];
    q[1] = v.x[i];
    q[2] = v.sx[i];
    q[3] = Real(0);
}

// The last step of a Halpern block with the edge terms of its check folded
// in: praq_edge_kernel's update of each edge (the reflected primal at its
// ends read from the node records), then, from the values the update holds
// and (x, sx) at both ends, the terms praq_edge_sums takes at the step point
// (sx, sy) -- one partial row per block.
template <class Real>
__global__ void __launch_bounds__(256)
praq_edge_check_kernel(PraqArrays<Real> a, PraqStep<Real> st, const Real* quad,
                       double* rows) {
    __shared__ double red[kPraqSums * 16];
    constexpr int K = kEdgesPerThread;
    const PraqView<Real> v = praq_view(a);
    const int m = a.m;
    const int first = blockIdx.x * blockDim.x * K + threadIdx.x, stride = blockDim.x;
    const gptr<const Real> qr = to_global(
        static_cast<const Real*>(__builtin_assume_aligned(quad, 4 * sizeof(Real))));
    int e[K], i0[K], i1[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int q = first + k * stride;
        e[k] = q < m ? q : 0;   // past the end: edge 0, dropped
    }
    Real sw[K], ya[K], yb[K], aa[K], ab[K], ra[K], rb[K], xa[K], sa[K], xb[K], sb[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        i0[k] = v.e0[e[k]];
        i1[k] = v.e1[e[k]];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        sw[k] = v.w[e[k]];
        ya[k] = v.yc[e[k]];
        yb[k] = v.yc[m + e[k]];
        aa[k] = v.y0[e[k]];
        ab[k] = v.y0[m + e[k]];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const gptr<const Real> qa = qr + 4 * size_t(i0[k]), qb = qr + 4 * size_t(i1[k]);
        ra[k] = qa[0];
        xa[k] = qa[1];
        sa[k] = qa[2];
        rb[k] = qb[0];
        xb[k] = qb[1];
        sb[k] = qb[2];
    }
    double acc[kPraqSums];
#pragma unroll
    for (int q = 0; q < kPraqSums; ++q) acc[q] = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (first + k * stride >= m) continue;
        Real y2a, y2b;
        praq_edge_update(v, st, e[k], sw[k], ya[k], yb[k], aa[k], ab[k], ra[k], rb[k], y2a, y2b);
        const double swd = sw[k];
        const double w = fabs(swd);
        const double da = double(ya[k]) - double(y2a), db = double(yb[k]) - double(y2b);
        acc[kDualNorm] += da * da + db * db;
        acc[kCross] += (double(xa[k]) - double(sa[k])) * da + (double(xb[k]) - double(sb[k])) * db;
        const double qa = double(ya[k]) - double(aa[k]), qb = double(yb[k]) - double(ab[k]);
        acc[kDualDrift] += qa * qa + qb * qb;
        const double bi = sa[k], bj = sb[k];
        if (swd > 0.0) {
            const double t = bi + bj - 1.0;
            acc[kUpperEdge] += w * (t > 0.0 ? t : 0.0);
            acc[kLowerEdge] += double(y2a);
        } else {
            acc[kUpperEdge] += -w * (bi < bj ? bi : bj);
        }
    }
    block_sums<kPraqSums>(acc, red);
    if (threadIdx.x == 0)
        for (int q = 0; q < kPraqSums; ++q) rows[size_t(blockIdx.x) * kPraqSums + q] = acc[q];
}

// Check sums over grid-stride node groups and (with_edges) edges, one
// partial row per block.
template <class Real>
__global__ void praq_sums_kernel(PraqArrays<Real> a, int lanes, const Real* bx, const Real* by,
                                 double* partial, bool with_edges) {
    __shared__ double red[kPraqSums * 16];
    double acc[kPraqSums];
#pragma unroll
    for (int k = 0; k < kPraqSums; ++k) acc[k] = 0.0;
    const int nthreads = gridDim.x * blockDim.x;
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    const int groups = nthreads / lanes;
    const PraqView<Real> v = praq_view(a);
    const gptr<const Real> gx = to_global(bx), gy = to_global(by);
    for (int base = 0; base < a.n; base += groups)
        praq_node_sums(v, base + gid / lanes, gid % lanes, lanes, gx, gy, acc);
    if (with_edges)
        for (int e = gid; e < a.m; e += nthreads) praq_edge_sums(v, e, gx, gy, true, acc);
    block_sums<kPraqSums>(acc, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < kPraqSums; ++k) partial[blockIdx.x * kPraqSums + k] = acc[k];
}

// one block adds the partial rows in a fixed order
__global__ void praq_sums_final(const double* partial, int rows, double* out) {
    __shared__ double red[kPraqSums * 16];
    double acc[kPraqSums];
#pragma unroll
    for (int k = 0; k < kPraqSums; ++k) acc[k] = 0.0;
    for (int r = threadIdx.x; r < rows; r += blockDim.x)
        for (int k = 0; k < kPraqSums; ++k) acc[k] += partial[size_t(r) * kPraqSums + k];
    block_sums<kPraqSums>(acc, red);
    if (threadIdx.x == 0)
        for (int k = 0; k < kPraqSums; ++k) out[k] = acc[k];
}

template <class Real>
__global__ void copy2_kernel(Real* dx, const Real* sx, int n, Real* dy, const Real* sy, int m2) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dx[k] = sx[k];
    if (k < m2) dy[k] = sy[k];
}

constexpr int kBlock = 256;
constexpr int kSumBlocks = 1024;   // check-sum blocks (rows), plus one row for the hubs

static int blocks_for(int64_t threads) { return int((threads + kBlock - 1) / kBlock); }

// hub chunk sums of src (nothing without hubs)
template <class Real>
static void device_chunks(const PraqArrays<Real>& a, const Real* src, hipStream_t stream) {
    if (a.nc > 0 && a.heavy_degree < std::numeric_limits<int>::max())
        hipLaunchKernelGGL(praq_chunk_kernel<Real>,
                           dim3((a.nc + kBlock / kWave - 1) / (kBlock / kWave)), dim3(kBlock), 0,
                           stream, a, src);
}

template <class Real>
static bool has_hubs(const PraqArrays<Real>& a) {
    return a.nh > 0 && a.heavy_degree < std::numeric_limits<int>::max();
}

template <class Real>
static int device_step(const PraqArrays<Real>& a, const PraqStep<Real>& st, int lanes,
                       hipStream_t stream) {
    device_chunks(a, st.reflect ? static_cast<const Real*>(a.yc) : a.sy, stream);
    hipLaunchKernelGGL(praq_node_kernel<Real>, dim3(blocks_for(int64_t(a.n) * lanes)),
                       dim3(kBlock), 0, stream, a, st, lanes);
    if (has_hubs(a))
        hipLaunchKernelGGL(praq_hub_kernel<Real>, dim3(blocks_for(a.nh)), dim3(kBlock), 0, stream,
                           a, st);
    if (a.m > 0)
        hipLaunchKernelGGL(praq_edge_kernel<Real>,
                           dim3(blocks_for((int64_t(a.m) + kEdgesPerThread - 1) / kEdgesPerThread)),
                           dim3(kBlock), 0, stream, a, st);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

template <class Real>
static int device_copy(Real* dx, const Real* sx, int n, Real* dy, const Real* sy, int m2,
                       hipStream_t stream) {
    int count = std::max(n, m2);
    hipLaunchKernelGGL(copy2_kernel<Real>, dim3(blocks_for(count)), dim3(kBlock), 0, stream, dx,
                       sx, n, dy, sy, m2);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

template <class Real>
static int device_sums(const PraqArrays<Real>& a, int lanes, const Real* bx, const Real* by,
                       double* partial, double* dsums, double* host, hipStream_t stream) {
    device_chunks(a, by, stream);
    hipLaunchKernelGGL(praq_sums_kernel<Real>, dim3(kSumBlocks), dim3(kBlock), 0, stream, a,
                       lanes, bx, by, partial, true);
    int rows = kSumBlocks;
    if (has_hubs(a)) {
        hipLaunchKernelGGL(praq_hub_sums_kernel<Real>, dim3(1), dim3(kBlock), 0, stream, a, bx,
                           partial + size_t(kSumBlocks) * kPraqSums);
        ++rows;
    }
    hipLaunchKernelGGL(praq_sums_final, dim3(1), dim3(1024), 0, stream, partial, rows, dsums);
    REFORM_HIP_CHECK(hipGetLastError());
    REFORM_HIP_CHECK(hipMemcpyAsync(host, dsums, sizeof(double) * kPraqSums,
                                    hipMemcpyDeviceToHost, stream));
    REFORM_HIP_CHECK(hipStreamSynchronize(stream));
    return 0;
}

// Blocks (and partial rows) of the edge kernels.
static int edge_blocks(int m) {
    return blocks_for((int64_t(m) + kEdgesPerThread - 1) / kEdgesPerThread);
}

// The last step of a Halpern block: as device_step, with the node records
// (refl, x, sx) written after the node pass and the edge pass taking its
// check terms into `rows`.
template <class Real>
static int device_check_step(const PraqArrays<Real>& a, const PraqStep<Real>& st, int lanes,
                             Real* quad, double* rows, hipStream_t stream) {
    device_chunks(a, static_cast<const Real*>(a.yc), stream);
    hipLaunchKernelGGL(praq_node_kernel<Real>, dim3(blocks_for(int64_t(a.n) * lanes)),
                       dim3(kBlock), 0, stream, a, st, lanes);
    if (has_hubs(a))
        hipLaunchKernelGGL(praq_hub_kernel<Real>, dim3(blocks_for(a.nh)), dim3(kBlock), 0, stream,
                           a, st);
    hipLaunchKernelGGL(praq_quad_kernel<Real>, dim3(blocks_for(a.n)), dim3(kBlock), 0, stream, a,
                       quad);
    if (a.m > 0)
        hipLaunchKernelGGL(praq_edge_check_kernel<Real>, dim3(edge_blocks(a.m)), dim3(kBlock), 0,
                           stream, a, st, static_cast<const Real*>(quad), rows);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

// The check after device_check_step, at (sx, sy): the node terms (CSR
// gather of sy) into rows [0, kSumBlocks), the hubs' into row kSumBlocks,
// the edge rows after them as the step left them; all rows added in order.
template <class Real>
static int device_check_sums(const PraqArrays<Real>& a, int lanes, double* partial, int eb,
                             double* dsums, double* host, hipStream_t stream) {
    device_chunks(a, static_cast<const Real*>(a.sy), stream);
    hipLaunchKernelGGL(praq_sums_kernel<Real>, dim3(kSumBlocks), dim3(kBlock), 0, stream, a,
                       lanes, static_cast<const Real*>(a.sx), static_cast<const Real*>(a.sy),
                       partial, false);
    double* hub_row = partial + size_t(kSumBlocks) * kPraqSums;
    if (has_hubs(a))
        hipLaunchKernelGGL(praq_hub_sums_kernel<Real>, dim3(1), dim3(kBlock), 0, stream, a,
                           static_cast<const Real*>(a.sx), hub_row);
    else
        REFORM_HIP_CHECK(hipMemsetAsync(hub_row, 0, sizeof(double) * kPraqSums, stream));
    const int rows = kSumBlocks + 1 + (a.m > 0 ? eb : 0);
    hipLaunchKernelGGL(praq_sums_final, dim3(1), dim3(1024), 0, stream, partial, rows, dsums);
    REFORM_HIP_CHECK(hipGetLastError());
    REFORM_HIP_CHECK(hipMemcpyAsync(host, dsums, sizeof(double) * kPraqSums,
                                    hipMemcpyDeviceToHost, stream));
    REFORM_HIP_CHECK(hipStreamSynchronize(stream));
    return 0;
}

// The host loop of the device tier: the same state machine as praq_wg_solve.
template <class Real>
int praq_device_solve(const PraqJob<Real>& job, LogicState& st, double* work,
                      double time_limit_seconds, hipStream_t stream, bool resume) {
    const LogicParams& p = job.p;
    const int lanes = job.lanes;
    // hubs by chunks in this tier
    PraqArrays<Real> a = job.a;
    a.heavy_degree = praq_heavy_degree(lanes);
    // work: partial rows (check sums, hubs, the check step's edge blocks),
    // the summed row, the node records (32-byte aligned)
    const int eb = edge_blocks(a.m);
    double* partial = work;
    double* edge_rows = partial + size_t(kSumBlocks + 1) * kPraqSums;
    double* dsums = partial + size_t(kSumBlocks + 1 + eb) * kPraqSums;
    auto* quad = reinterpret_cast<Real*>(
        (reinterpret_cast<uintptr_t>(dsums + kPraqSums) + 31) & ~uintptr_t(31));
    double sums[kPraqSums];
    auto now = [] { return std::chrono::steady_clock::now(); };
    auto t0 = now();
    LogicProgress progress;

    if (resume) {
        st.stalled = 0;
    } else {
        if (device_copy(a.x0, a.xs, a.n, a.y0, a.ys, 2 * a.m, stream)) return kErrorDevice;
        if (device_copy(a.x, a.xs, a.n, a.y, a.ys, 2 * a.m, stream)) return kErrorDevice;
        if (device_copy(a.sx, a.xs, a.n, a.sy, a.ys, 2 * a.m, stream)) return kErrorDevice;
        logic_init(st, p);
        if (device_step(a, praq_step_scalars<Real>(0, p.omega, p.eta, false, true), lanes, stream))
            return kErrorDevice;
        if (device_sums(a, lanes, a.x, a.y, partial, dsums, sums, stream)) return kErrorDevice;
        logic_after_warm_check(st, p, praq_check_sums(sums, job.scale));
    }

    for (;;) {
        const int action = st.action;
        const int64_t r0 = st.restart_iter;
        const double omega = st.omega;
        if (st.stalled) break;
        if (action == kActDoneX) break;
        if (action == kActDoneSX) {
            if (device_copy(a.x, a.sx, a.n, a.y, a.sy, 2 * a.m, stream)) return kErrorDevice;
            break;
        }
        if (action == kActWarmBlock) {
            for (int64_t j = 0; j < p.op_steps; ++j)
                if (device_step(a, praq_step_scalars<Real>(r0 + j, omega, p.eta, false, true),
                                lanes, stream))
                    return kErrorDevice;
            if (device_sums(a, lanes, a.x, a.y, partial, dsums, sums, stream)) return kErrorDevice;
            st.restart_iter += p.op_steps;
            st.iteration += p.op_steps;
            logic_after_warm_check(st, p, praq_check_sums(sums, job.scale));
        } else if (action == kActHalpernBlock) {
            if (st.anchor_pending) {
                if (device_copy(a.x0, a.x, a.n, a.y0, a.y, 2 * a.m, stream)) return kErrorDevice;
                st.anchor_pending = 0;
            }
            // the last step takes the edge terms of the check with it
            for (int64_t j = 0; j < p.op_steps; ++j) {
                const bool last = j + 1 == p.op_steps;
                const PraqStep<Real> sc =
                    praq_step_scalars<Real>(r0 + j, omega, p.eta, true, last);
                if (last ? device_check_step(a, sc, lanes, quad, edge_rows, stream)
                         : device_step(a, sc, lanes, stream))
                    return kErrorDevice;
            }
            if (device_check_sums(a, lanes, partial, eb, dsums, sums, stream))
                return kErrorDevice;
            st.restart_iter += p.op_steps;
            st.iteration += p.op_steps;
            logic_after_halpern_check(st, p, praq_check_sums(sums, job.scale));
        } else {
            if (device_step(a, praq_step_scalars<Real>(0, omega, p.eta, false, true), lanes, stream))
                return kErrorDevice;
            if (device_copy(a.x0, a.x, a.n, a.y0, a.y, 2 * a.m, stream)) return kErrorDevice;
            if (device_sums(a, lanes, a.sx, a.sy, partial, dsums, sums, stream))
                return kErrorDevice;
            logic_after_restart(st, p, praq_check_sums(sums, job.scale));
        }
        progress.maybe(st, "praq", std::chrono::duration<double>(now() - t0).count());
        if (time_limit_seconds > 0.0 &&
            std::chrono::duration<double>(now() - t0).count() > time_limit_seconds &&
            st.action != kActDoneX && st.action != kActDoneSX) {
            st.action = action == kActWarmBlock || action == kActRestart ? kActDoneX : kActDoneSX;
            st.status = 2;
        }
    }
    REFORM_HIP_CHECK(hipStreamSynchronize(stream));
    return 0;
}

template <class Real>
int praq_time_steps(const PraqJob<Real>& job, int warmup, int steps, hipStream_t stream,
                    double* ms) {
    PraqArrays<Real> a = job.a;
    a.heavy_degree = praq_heavy_degree(job.lanes);
    if (device_copy(a.x0, a.xs, a.n, a.y0, a.ys, 2 * a.m, stream) ||
        device_copy(a.x, a.xs, a.n, a.y, a.ys, 2 * a.m, stream) ||
        device_copy(a.sx, a.xs, a.n, a.sy, a.ys, 2 * a.m, stream) ||
        device_copy(a.x, a.xs, a.n, a.yc, a.ys, 2 * a.m, stream))
        return kErrorDevice;
    hipEvent_t ev[2];
    REFORM_HIP_CHECK(hipEventCreate(&ev[0]));
    REFORM_HIP_CHECK(hipEventCreate(&ev[1]));
    int st = 0;
    for (int j = 0; j < warmup + steps && !st; ++j) {
        if (j == warmup) st = hipEventRecord(ev[0], stream) == hipSuccess ? 0 : kErrorDevice;
        if (!st)
            st = device_step(a, praq_step_scalars<Real>(j, job.p.omega, job.p.eta, true, false),
                             job.lanes, stream);
    }
    float t = 0.0f;
    if (!st && (hipEventRecord(ev[1], stream) != hipSuccess ||
                hipEventSynchronize(ev[1]) != hipSuccess ||
                hipEventElapsedTime(&t, ev[0], ev[1]) != hipSuccess))
        st = kErrorDevice;
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    if (!st) *ms = double(t) / double(steps > 0 ? steps : 1);
    return st;
}

template int praq_time_steps<float>(const PraqJob<float>&, int, int, hipStream_t, double*);
template int praq_time_steps<double>(const PraqJob<double>&, int, int, hipStream_t, double*);

template int praq_device_solve<float>(const PraqJob<float>&, LogicState&, double*, double,
                                      hipStream_t, bool);
template int praq_device_solve<double>(const PraqJob<double>&, LogicState&, double*, double,
                                       hipStream_t, bool);

// ---------------------------------------------------------------------------
// precision switch: the whole Halpern state, FP32 -> FP64
// ---------------------------------------------------------------------------

__device__ inline void praq_convert_one(const PraqView<float>& f, const PraqView<double>& d,
                                        int k) {
    if (k < f.n) {
        d.x[k] = f.x[k];
        d.sx[k] = f.sx[k];
        d.x0[k] = f.x0[k];
    }
    if (k < 2 * f.m) {
        d.yc[k] = f.yc[k];
        d.y[k] = f.y[k];
        d.sy[k] = f.sy[k];
        d.y0[k] = f.y0[k];
    }
}

__global__ void praq_convert_kernel(PraqArrays<float> f, PraqArrays<double> d) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    praq_convert_one(praq_view(f), praq_view(d), k);
}

// one workgroup per job, grid-stride inside the instance
__global__ void praq_convert_jobs_kernel(const PraqJob<float>* from, const PraqJob<double>* to) {
    const PraqView<float> f = praq_view(from[blockIdx.x].a);
    const PraqView<double> d = praq_view(to[blockIdx.x].a);
    const int count = f.n > 2 * f.m ? f.n : 2 * f.m;
    for (int k = threadIdx.x; k < count; k += blockDim.x) praq_convert_one(f, d, k);
}

int praq_convert_state(const PraqArrays<float>& from, const PraqArrays<double>& to,
                       hipStream_t stream) {
    int count = std::max(from.n, 2 * from.m);
    hipLaunchKernelGGL(praq_convert_kernel, dim3(blocks_for(count)), dim3(kBlock), 0, stream, from,
                       to);
    return hipGetLastError() == hipSuccess ? 0 : kErrorDevice;
}

int praq_convert_jobs(const PraqJob<float>* from, const PraqJob<double>* to, int count,
                      hipStream_t stream) {
    if (count <= 0) return 0;
    hipLaunchKernelGGL(praq_convert_jobs_kernel, dim3(count), dim3(256), 0, stream, from, to);
    return hipGetLastError() == hipSuccess ? 0 : kErrorDevice;
}

size_t praq_device_work_bytes(int64_t n, int64_t m) {
    const int64_t eb = (m / kEdgesPerThread + 1) / kBlock + 1;   // >= edge_blocks(m)
    return sizeof(double) * (size_t(kSumBlocks + 2 + eb) * kPraqSums + 4 * size_t(n) + 4);
}

template <class Real>
int praq_launch_wg(const PraqJob<Real>* jobs, int count, int block, hipStream_t stream) {
    if (count <= 0) return 0;
    hipLaunchKernelGGL(praq_wg_solve<Real>, dim3(count), dim3(block), 0, stream, jobs);
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

template int praq_launch_wg<float>(const PraqJob<float>*, int, int, hipStream_t);
template int praq_launch_wg<double>(const PraqJob<double>*, int, int, hipStream_t);

}  // namespace reform
