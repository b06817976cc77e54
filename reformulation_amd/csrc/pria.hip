// PRIA on gfx950: kernels and the two execution tiers.
//
//   workgroup tier  pria_wg_solve: one 1024-thread workgroup runs a whole
//                   solve (x tiles, y/z slices, checks, scalar logic) in one
//                   launch; one workgroup per instance is also how a batch of
//                   instances runs in one launch.
//   device tier     per step an x kernel over (item, column tile) and a y/z
//                   kernel over slices; per check a slice kernel, a column
//                   kernel and a fixed-order final reduction.
//
// Reference: jrp/pria-cuda/src/solver.cuh:61-1253 (kernels, device logic,
// whole-solve graph), jrp/pria-cuda/src/projection.cuh:298-538 (Michelot).
#include "pria_host.hpp"

#include <algorithm>
#include <chrono>
#include <vector>

namespace reform {

// ---------------------------------------------------------------------------
// workgroup tier
// ---------------------------------------------------------------------------

template <class Real>
__device__ void pria_wg_step(const PriaView<Real>& a, const PriaStep<Real>& st, Real* lds,
                             double* red) {
    constexpr int TW = PriaShape<Real>::TW;
    const int tiles = (a.T + TW - 1) / TW;
    for (int i = 0; i < a.N; ++i)
        for (int q = 0; q < tiles; ++q) pria_x_tile<Real, TW>(a, st, i, q * TW, lds);
    for (int s = 0; s < a.T; ++s)
        pria_yz_slice<Real, PriaShape<Real>::VR, PriaShape<Real>::VZ>(a, st, s, red);
}

template <class Real>
__device__ void pria_wg_sums(const PriaView<Real>& a, gptr<const Real> bx, gptr<const Real> by,
                             gptr<const Real> bz, double* red, double (&out)[kPriaSums]) {
    constexpr int TW = PriaShape<Real>::TW;
    double acc[kPriaSums];
    for (int q = 0; q < kPriaSums; ++q) acc[q] = 0.0;
    for (int s = 0; s < a.T; ++s) pria_slice_sums(a, s, bx, red, acc);
    const int tiles = (a.T + TW - 1) / TW;
    for (int i = 0; i < a.N; ++i)
        for (int q = 0; q < tiles; ++q) pria_column_sums<Real, TW>(a, i, q * TW, by, bz, red, acc);
    for (int q = 0; q < kPriaSums; ++q) out[q] = acc[q];
}

template <class Real>
__device__ void pria_wg_copy3(gptr<Real> d1, gptr<const Real> s1, gptr<Real> d2,
                              gptr<const Real> s2, gptr<Real> d3, gptr<const Real> s3, int nd) {
    for (int e = threadIdx.x; e < nd; e += blockDim.x) {
        d1[e] = s1[e];
        d2[e] = s2[e];
        d3[e] = s3[e];
    }
    __syncthreads();
}

template <class Real>
__global__ void __launch_bounds__(1024) pria_wg_solve(const PriaJob<Real>* jobs) {
    extern __shared__ char dyn[];
    Real* lds = reinterpret_cast<Real*>(dyn);
    __shared__ double red[kPriaSums * 16 > kPriaYZRed ? kPriaSums * 16 : kPriaYZRed];
    __shared__ LogicState st;
    __shared__ PriaJob<Real> job;
    if (threadIdx.x == 0) job = jobs[blockIdx.x];
    __syncthreads();
    const PriaView<Real> a = pria_view(job.a);
    LogicState* const state = job.state;
    const LogicParams p = job.p;
    const long long t_start = wall_clock64();

    double sums[kPriaSums];
    if (job.resume) {
        if (threadIdx.x == 0) {
            st = *state;
            st.stalled = 0;
        }
        __syncthreads();
    } else {
        pria_wg_copy3(a.x0, a.xs, a.y0, a.ys, a.z0, a.zs, a.ND);
        pria_wg_copy3(a.x, a.xs, a.y, a.ys, a.z, a.zs, a.ND);
        pria_wg_copy3(a.sx, a.xs, a.sy, a.ys, a.sz, a.zs, a.ND);
        pria_reset_thresholds(a, threadIdx.x, blockDim.x);
        if (threadIdx.x == 0) logic_init(st, p);
        __syncthreads();
        pria_wg_step(a, pria_step_scalars<Real>(0, p.omega, p.eta, false, true), lds, red);
        pria_wg_sums(a, a.x, a.y, a.z, red, sums);
        if (threadIdx.x == 0) logic_after_warm_check(st, p, pria_check_sums(sums, job.scale));
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
            pria_wg_copy3(a.x, a.sx, a.y, a.sy, a.z, a.sz, a.ND);
            break;
        }
        if (action == kActWarmBlock) {
            for (int64_t j = 0; j < p.op_steps; ++j)
                pria_wg_step(a, pria_step_scalars<Real>(r0 + j, omega, p.eta, false, true), lds,
                             red);
            pria_wg_sums(a, a.x, a.y, a.z, red, sums);
            if (threadIdx.x == 0) {
                st.restart_iter += p.op_steps;
                st.iteration += p.op_steps;
                logic_after_warm_check(st, p, pria_check_sums(sums, job.scale));
            }
        } else if (action == kActHalpernBlock) {
            if (anchor_pending) {
                pria_wg_copy3(a.x0, a.x, a.y0, a.y, a.z0, a.z, a.ND);
                if (threadIdx.x == 0) st.anchor_pending = 0;
            }
            for (int64_t j = 0; j < p.op_steps; ++j)
                pria_wg_step(a, pria_step_scalars<Real>(r0 + j, omega, p.eta, true,
                                                        j + 1 == p.op_steps),
                             lds, red);
            pria_wg_sums(a, a.sx, a.sy, a.sz, red, sums);
            if (threadIdx.x == 0) {
                st.restart_iter += p.op_steps;
                st.iteration += p.op_steps;
                logic_after_halpern_check(st, p, pria_check_sums(sums, job.scale));
            }
        } else {  // kActRestart
            pria_wg_step(a, pria_step_scalars<Real>(0, omega, p.eta, false, true), lds, red);
            pria_wg_copy3(a.x0, a.x, a.y0, a.y, a.z0, a.z, a.ND);
            pria_wg_sums(a, a.sx, a.sy, a.sz, red, sums);
            if (threadIdx.x == 0) logic_after_restart(st, p, pria_check_sums(sums, job.scale));
        }
        __syncthreads();
        if (threadIdx.x == 0 && job.time_limit_ticks > 0 &&
            wall_clock64() - t_start > job.time_limit_ticks && st.action != kActDoneX &&
            st.action != kActDoneSX) {
            st.action = (action == kActWarmBlock || action == kActRestart) ? kActDoneX : kActDoneSX;
            st.status = 2;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *state = st;
}

// ---------------------------------------------------------------------------
// device tier
// ---------------------------------------------------------------------------

constexpr int kPriaXBlock = 512;
constexpr int kPriaYZBlock = 512;
constexpr int kPriaCheckBlock = 256;

template <class Real>
__global__ void __launch_bounds__(512) pria_x_kernel(PriaArrays<Real> a, PriaStep<Real> st) {
    extern __shared__ char dyn[];
    constexpr int TW = PriaShape<Real>::TW;
    const int tiles = (a.T + TW - 1) / TW;
    // heaviest tiles (largest t) first
    const int b = blockIdx.x;
    const int q = tiles - 1 - (b % tiles), i = b / tiles;
    pria_x_tile<Real, TW>(pria_view(a), st, i, q * TW, reinterpret_cast<Real*>(dyn));
}

// The device tier's y rows: kPriaYRows rows of one slice per 256-thread
// block (rows are independent, so a large instance's rows fill the chip),
// then the z slices one 512-thread block each; per entry the arithmetic of
// the fused pass.
constexpr int kPriaYRows = 16;

// Shape by T' (the compact register shape up to T' = 320; the sums do not
// depend on the shape).
template <class Real, bool compact>
struct PriaPick {
    static constexpr int VR = compact ? PriaCompactShape<Real>::VR : PriaShape<Real>::VR;
    static constexpr int VZ = compact ? PriaCompactShape<Real>::VZ : PriaShape<Real>::VZ;
};

template <class Real, bool compact>
__global__ void __launch_bounds__(256) pria_y_kernel(PriaArrays<Real> a, PriaStep<Real> st,
                                                     int chunks) {
    const int s = blockIdx.x / chunks, i_lo = (blockIdx.x % chunks) * kPriaYRows;
    pria_yz_slice<Real, PriaPick<Real, compact>::VR, PriaPick<Real, compact>::VZ>(
        pria_view(a), st, s, nullptr, 1, i_lo, i_lo + kPriaYRows);
}

template <class Real, bool compact>
__global__ void __launch_bounds__(512) pria_yz_kernel(PriaArrays<Real> a, PriaStep<Real> st) {
    __shared__ double red[kPriaYZRed];
    extern __shared__ double zcache[];   // kPriaZCacheBytes
    pria_yz_slice<Real, PriaPick<Real, compact>::VR, PriaPick<Real, compact>::VZ>(
        pria_view(a), st, blockIdx.x, red, 2, 0, 1 << 30, reinterpret_cast<Real*>(zcache),
        kPriaZCacheBytes / int(sizeof(Real)));
}

template <class Real>
__global__ void __launch_bounds__(256) pria_check_slice_kernel(PriaArrays<Real> a, const Real* bx,
                                                               double* partial) {
    __shared__ double red[kPriaSums * 16];
    double acc[kPriaSums];
    for (int q = 0; q < kPriaSums; ++q) acc[q] = 0.0;
    pria_slice_sums(pria_view(a), blockIdx.x, to_global(bx), red, acc);
    if (threadIdx.x == 0)
        for (int q = 0; q < kPriaSums; ++q) partial[blockIdx.x * kPriaSums + q] = acc[q];
}

template <class Real>
__global__ void __launch_bounds__(256) pria_check_column_kernel(PriaArrays<Real> a, const Real* by,
                                                                const Real* bz, double* partial) {
    __shared__ double red[16];
    constexpr int TW = PriaShape<Real>::TW;
    const int tiles = (a.T + TW - 1) / TW;
    double acc[kPriaSums];
    for (int q = 0; q < kPriaSums; ++q) acc[q] = 0.0;
    pria_column_sums<Real, TW>(pria_view(a), blockIdx.x / tiles, (blockIdx.x % tiles) * TW,
                               to_global(by), to_global(bz), red, acc);
    if (threadIdx.x == 0) partial[blockIdx.x] = acc[kPLower];
}

// adds rows in a fixed order: slices first, then column tiles
__global__ void pria_sums_final(const double* slices, int nslices, const double* cols, int ncols,
                                double* out) {
    __shared__ double red[kPriaSums * 16];
    double acc[kPriaSums];
    for (int q = 0; q < kPriaSums; ++q) acc[q] = 0.0;
    for (int s = threadIdx.x; s < nslices; s += blockDim.x)
        for (int q = 0; q < kPLower; ++q) acc[q] += slices[s * kPriaSums + q];
    for (int c = threadIdx.x; c < ncols; c += blockDim.x) acc[kPLower] += cols[c];
    block_sums<kPriaSums>(acc, red);
    if (threadIdx.x == 0)
        for (int q = 0; q < kPriaSums; ++q) out[q] = acc[q];
}

template <class Real>
__global__ void pria_reset_thresholds_kernel(PriaArrays<Real> a) {
    pria_reset_thresholds(pria_view(a), blockIdx.x * blockDim.x + threadIdx.x,
                          gridDim.x * blockDim.x);
}

template <class Real>
__global__ void copy3_kernel(Real* d1, const Real* s1, Real* d2, const Real* s2, Real* d3,
                             const Real* s3, int nd) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < nd) {
        d1[e] = s1[e];
        d2[e] = s2[e];
        d3[e] = s3[e];
    }
}

// The cold start, on the device: the projection of zero onto each simplex is
// its centre, x = 1 / (t + 1) per column, y = K_i / (T' - s) per row,
// z = K0 / (N (T' - s)) per slice (0 where the budget is 0).
template <class Real>
__global__ void pria_cold_start_kernel(PriaArrays<Real> a, Real* xs, Real* ys, Real* zs) {
    const int s = blockIdx.x, len = a.T - s, L = a.N * len, base = pria_off(a.N, a.T, s);
    const double k0 = double(a.k0);
    for (int e = threadIdx.x; e < L; e += blockDim.x) {
        int i = e / len, t = s + e % len;
        double ki = double(a.k[i]);
        xs[base + e] = Real(1.0 / double(t + 1));
        ys[base + e] = Real(ki > 0.0 ? ki / double(len) : 0.0);
        zs[base + e] = Real(k0 > 0.0 ? k0 / double(L) : 0.0);
    }
}

template <class Real>
int pria_cold_start(const PriaArrays<Real>& a, hipStream_t stream) {
    hipLaunchKernelGGL(pria_cold_start_kernel<Real>, dim3(a.T), dim3(256), 0, stream, a,
                       const_cast<Real*>(a.xs), const_cast<Real*>(a.ys), const_cast<Real*>(a.zs));
    return hipGetLastError() == hipSuccess ? 0 : kPriaErrorDevice;
}

template int pria_cold_start<float>(const PriaArrays<float>&, hipStream_t);
template int pria_cold_start<double>(const PriaArrays<double>&, hipStream_t);

// ---- the device tier's z pass over chunks -----------------------------------
//
// For a large instance (ND >= 2^22 entries with slices longer than the z
// cache of one workgroup) one workgroup per slice leaves most of the chip
// idle and re-reads each long slice on every Michelot round.  There the z
// projection runs over chunks of kPriaZChunk entries of a slice instead,
// one 256-thread workgroup each: a pass sums its chunk's values (the first
// pass also those above the slice's last threshold), a combine (one thread
// per slice) adds the slice's chunk sums in chunk order and takes the
// Michelot step -- the same test and update as pria_yz_slice's rounds -- and
// kPriaZRounds such rounds run over all chunks; a slice whose active set
// has not settled by then finishes its rounds in one workgroup, and a write
// pass over the chunks applies each slice's threshold.  Every sum is in a
// fixed order, so the result does not depend on scheduling; it is theta of
// the final active set, as every tier's.
constexpr int kPriaZChunk = 4096;
constexpr int kPriaZBlock = 256;
constexpr int kPriaZRounds = 4;
enum { kZTh = 0, kZC = 1, kZFirst = 2, kZDone = 3, kZState = 4 };

// Whether the device tier runs the z pass over chunks for an instance.
inline bool pria_zsplit(int N, int T, size_t real_bytes) {
    const int64_t nd = int64_t(N) * T * (T + 1) / 2;
    return nd >= (int64_t(1) << 22) &&
           int64_t(N) * T > int64_t(kPriaZCacheBytes) / int64_t(real_bytes);
}

__host__ __device__ inline int pria_zchunks_of(int N, int T, int s) {
    const int L = N * (T - s);
    return (L + kPriaZChunk - 1) / kPriaZChunk;
}

inline size_t pria_zchunks(int N, int T) {
    size_t n = 0;
    for (int s = 0; s < T; ++s) n += size_t(pria_zchunks_of(N, T, s));
    return n;
}

// zfirst[s] = the first chunk of slice s (zfirst[T] = all chunks); zchunk[k]
// = the slice of chunk k.  One thread: once per solve.
__global__ void pria_zchunk_table_kernel(int N, int T, int* zfirst, int* zchunk) {
    int k = 0;
    for (int s = 0; s < T; ++s) {
        zfirst[s] = k;
        const int n = pria_zchunks_of(N, T, s);
        for (int c = 0; c < n; ++c) zchunk[k++] = s;
    }
    zfirst[T] = k;
}

// Round 0: per chunk (sum, count) of all u = z1 + r and of those above the
// slice's last threshold; round > 0: (sum, count) of those above the
// slice's current threshold, for slices not settled.
template <class Real>
__global__ void __launch_bounds__(kPriaZBlock)
pria_zpass_kernel(PriaArrays<Real> a, PriaStep<Real> st, const int* zchunk, const int* zfirst,
                  const double* zst, double* zp, int round) {
    __shared__ double red[4 * 16];
    const PriaView<Real> v = pria_view(a);
    if (double(v.k0) <= 0.0) return;   // z = 0: no sums (the combine settles it)
    const int k = blockIdx.x, s = zchunk[k], c = k - zfirst[s];
    if (round > 0 && zst[s * kZState + kZDone] != 0.0) return;
    const int L = a.N * (a.T - s), base = pria_off(a.N, a.T, s);
    const int e0 = c * kPriaZChunk, e1 = min(L, e0 + kPriaZChunk);
    const gptr<const Real> srcz = st.reflect ? v.zc : v.sz;
    const double th = round == 0 ? v.thz[s] : zst[s * kZState + kZTh];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    constexpr int P = kPriaZChunk / kPriaZBlock;   // entries per thread
    Real u[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int e = e0 + int(threadIdx.x) + kPriaZBlock * q;
        const int g = base + (e < e1 ? e : e0);   // past the chunk: its first entry, dropped
        u[q] = srcz[g] + v.r[g];
    }
#pragma unroll
    for (int q = 0; q < P; ++q) {
        const int e = e0 + int(threadIdx.x) + kPriaZBlock * q;
        if (e >= e1) continue;
        const double x = double(u[q]);
        if (round == 0) {
            acc[0] += x;
            acc[1] += 1.0;
            if (x > th) {
                acc[2] += x;
                acc[3] += 1.0;
            }
        } else if (x > th) {
            acc[0] += x;
            acc[1] += 1.0;
        }
    }
    block_sums<4>(acc, red);
    if (threadIdx.x == 0)
        for (int q = 0; q < 4; ++q) zp[size_t(k) * 4 + q] = acc[q];
}

// One wavefront per slice: its chunk sums (lane l adds chunks l, l + 64,
// ... in order, then a DPP sum over the lanes), then the first threshold
// (round 0) or one Michelot step (round > 0); a settled slice keeps its
// threshold for the next step's start.
__global__ void __launch_bounds__(256)
pria_zcombine_kernel(const double* zp, const int* zfirst, double* zst, double* thz, double target,
                     int T, int round) {
    const int s = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, lane = threadIdx.x % kWave;
    if (s >= T) return;   // wave-uniform
    double* z = zst + size_t(s) * kZState;
    if (target <= 0.0) {
        if (round == 0 && lane == 0) {
            z[kZTh] = 0.0;
            z[kZDone] = 1.0;
        }
        return;
    }
    if (round > 0 && z[kZDone] != 0.0) return;
    double S = 0.0, C = 0.0, Sw = 0.0, Cw = 0.0;
    for (int k = zfirst[s] + lane; k < zfirst[s + 1]; k += kWave) {
        S += zp[size_t(k) * 4];
        C += zp[size_t(k) * 4 + 1];
        Sw += zp[size_t(k) * 4 + 2];
        Cw += zp[size_t(k) * 4 + 3];
    }
    S = wave_sum(S);
    C = wave_sum(C);
    Sw = wave_sum(Sw);
    Cw = wave_sum(Cw);
    if (lane != 0) return;
    if (round == 0) {
        const bool first = Cw > 0.0;
        const double Cn = first ? Cw : C;
        z[kZTh] = ((first ? Sw : S) - target) / Cn;
        z[kZC] = Cn;
        z[kZFirst] = first ? 1.0 : 0.0;
        z[kZDone] = 0.0;
        return;
    }
    const double C0 = z[kZC];
    const bool first = z[kZFirst] != 0.0;
    if (C == C0 || C == 0.0 || (C > C0 && !first)) {
        z[kZDone] = 1.0;
        thz[s] = z[kZTh];
    } else {
        z[kZC] = C;
        z[kZTh] = (S - target) / C;
        z[kZFirst] = 0.0;
    }
}

// Slices not settled after the chunk rounds: the rest of their rounds in
// one workgroup (values re-read each round), as pria_yz_slice's memory path.
template <class Real>
__global__ void __launch_bounds__(512)
pria_zfinish_kernel(PriaArrays<Real> a, PriaStep<Real> st, double* zst) {
    __shared__ double red[kPriaYZRed];
    const PriaView<Real> v = pria_view(a);
    const int s = blockIdx.x;
    double* z = zst + size_t(s) * kZState;
    const double target = double(v.k0);
    if (target <= 0.0 || z[kZDone] != 0.0) return;
    const int L = a.N * (a.T - s), base = pria_off(a.N, a.T, s);
    const gptr<const Real> srcz = st.reflect ? v.zc : v.sz;
    double th = z[kZTh], C = z[kZC];
    bool first = z[kZFirst] != 0.0;
    int parity = 0;
    for (int round = 0; round <= L; ++round) {
        double w2[2] = {0.0, 0.0};
        for (int e = threadIdx.x; e < L; e += blockDim.x) {
            const double x = double(Real(srcz[base + e] + v.r[base + e]));
            if (x > th) {
                w2[0] += x;
                w2[1] += 1.0;
            }
        }
        block_sums_alt<2>(w2, red, parity);
        if (w2[1] == C || w2[1] == 0.0 || (w2[1] > C && !first)) break;
        C = w2[1];
        th = (w2[0] - target) / w2[1];
        first = false;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        z[kZTh] = th;
        z[kZDone] = 1.0;
        v.thz[s] = th;
    }
}

// Per chunk: z2 from its slice's threshold, the next zc (and z, sz when
// storing), with pria_yz_slice's arithmetic.
template <class Real>
__global__ void __launch_bounds__(kPriaZBlock)
pria_zwrite_kernel(PriaArrays<Real> a, PriaStep<Real> st, const int* zchunk, const int* zfirst,
                   const double* zst) {
    const PriaView<Real> v = pria_view(a);
    const int k = blockIdx.x, s = zchunk[k], c = k - zfirst[s];
    const int L = a.N * (a.T - s), base = pria_off(a.N, a.T, s);
    const int e0 = c * kPriaZChunk, e1 = min(L, e0 + kPriaZChunk);
    const double target = double(v.k0);
    const Real thr = Real(zst[s * kZState + kZTh]);
    const gptr<const Real> srcz = st.reflect ? v.zc : v.sz;
    constexpr int U = 4;
    for (int b0 = e0 + int(threadIdx.x); b0 < e1; b0 += kPriaZBlock * U) {
        int g[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int e = b0 + kPriaZBlock * q;
            g[q] = base + (e < e1 ? e : e0);
        }
        Real v1[U], rr[U], an[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            v1[q] = srcz[g[q]];
            rr[q] = v.r[g[q]];
        }
#pragma unroll
        for (int q = 0; q < U; ++q) an[q] = st.reflect ? Real(v.z0[g[q]]) : v1[q];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int e = b0 + kPriaZBlock * q;
            if (e < e1) {
                const Real d = (v1[q] + rr[q]) - thr;
                const Real z2 = target > 0.0 ? (d > Real(0) ? d : Real(0)) : Real(0);
                v.zc[g[q]] = st.lwn * (Real(2) * z2 - v1[q]) + st.rwn * an[q];
                if (st.store) {
                    v.z[g[q]] = v1[q];
                    v.sz[g[q]] = z2;
                }
            }
        }
    }
}

size_t pria_device_work_bytes(int N, int T) {
    // the narrower precision's tiles bound the column rows
    constexpr int TW = PriaShape<float>::TW < PriaShape<double>::TW ? PriaShape<float>::TW
                                                                    : PriaShape<double>::TW;
    size_t tiles = size_t(N) * ((T + TW - 1) / TW);
    size_t bytes = sizeof(double) * (size_t(T) * kPriaSums + tiles + kPriaSums + 8);
    // the chunked z pass: slice states, chunk sums, chunk table (FP64's
    // test is the looser one: its z cache holds fewer values)
    if (pria_zsplit(N, T, sizeof(double))) {
        const size_t chunks = pria_zchunks(N, T);
        bytes += sizeof(double) * (size_t(T) * kZState + chunks * 4) +
                 sizeof(int) * (size_t(T) + 1 + chunks);
    }
    return bytes;
}

template <class Real>
struct PriaDevice {
    const PriaArrays<Real>& a;
    hipStream_t stream;
    double* slices;
    double* cols;
    double* dsums;
    // the chunked z pass (zst null: one workgroup per slice)
    double* zst = nullptr;
    double* zp = nullptr;
    int* zfirst = nullptr;
    int* zchunk = nullptr;
    int nchunks = 0;

    // Places the chunked z pass's arrays after the check rows (`end` = the
    // work area's first free double) when the instance takes it, and builds
    // the chunk table.
    int place_zsplit(double* end) {
        if (!pria_zsplit(a.N, a.T, sizeof(Real))) return 0;
        nchunks = int(pria_zchunks(a.N, a.T));
        zst = end;
        zp = zst + size_t(a.T) * kZState;
        zfirst = reinterpret_cast<int*>(zp + size_t(nchunks) * 4);
        zchunk = zfirst + a.T + 1;
        hipLaunchKernelGGL(pria_zchunk_table_kernel, dim3(1), dim3(1), 0, stream, a.N, a.T, zfirst,
                           zchunk);
        return hipGetLastError() == hipSuccess ? 0 : kPriaErrorDevice;
    }

    // The z part of a step over chunks (place_zsplit done).
    void z_chunked(const PriaStep<Real>& st) {
        const double target = double(a.k0);
        const int cb = (a.T * kWave + 255) / 256;
        for (int round = 0; round <= kPriaZRounds; ++round) {
            hipLaunchKernelGGL(pria_zpass_kernel<Real>, dim3(nchunks), dim3(kPriaZBlock), 0,
                               stream, a, st, zchunk, zfirst, zst, zp, round);
            hipLaunchKernelGGL(pria_zcombine_kernel, dim3(cb), dim3(256), 0, stream, zp, zfirst,
                               zst, a.thz, target, a.T, round);
        }
        hipLaunchKernelGGL(pria_zfinish_kernel<Real>, dim3(a.T), dim3(512), 0, stream, a, st, zst);
        hipLaunchKernelGGL(pria_zwrite_kernel<Real>, dim3(nchunks), dim3(kPriaZBlock), 0, stream,
                           a, st, zchunk, zfirst, zst);
    }

    int reset_thresholds() {
        const int n = a.N * a.T;
        hipLaunchKernelGGL(pria_reset_thresholds_kernel<Real>, dim3((n + 255) / 256), dim3(256), 0,
                           stream, a);
        return hipGetLastError() == hipSuccess ? 0 : kPriaErrorDevice;
    }
    int step(const PriaStep<Real>& st) {
        constexpr int TW = PriaShape<Real>::TW;
        const int tiles = (a.T + TW - 1) / TW;
        hipLaunchKernelGGL(pria_x_kernel<Real>, dim3(a.N * tiles), dim3(kPriaXBlock),
                           pria_x_lds_bytes<Real>(a.T), stream, a, st);
        const int chunks = (a.N + kPriaYRows - 1) / kPriaYRows;
        if (a.T <= kPriaCompactT) {
            hipLaunchKernelGGL((pria_y_kernel<Real, true>), dim3(a.T * chunks), dim3(256), 0,
                               stream, a, st, chunks);
            if (!zst)
                hipLaunchKernelGGL((pria_yz_kernel<Real, true>), dim3(a.T), dim3(kPriaYZBlock),
                                   kPriaZCacheBytes, stream, a, st);
        } else {
            hipLaunchKernelGGL((pria_y_kernel<Real, false>), dim3(a.T * chunks), dim3(256), 0,
                               stream, a, st, chunks);
            if (!zst)
                hipLaunchKernelGGL((pria_yz_kernel<Real, false>), dim3(a.T), dim3(kPriaYZBlock),
                                   kPriaZCacheBytes, stream, a, st);
        }
        if (zst) z_chunked(st);
        return hipGetLastError() == hipSuccess ? 0 : kPriaErrorDevice;
    }
    int copy3(Real* d1, const Real* s1, Real* d2, const Real* s2, Real* d3, const Real* s3) {
        hipLaunchKernelGGL(copy3_kernel<Real>, dim3((a.ND + 255) / 256), dim3(256), 0, stream, d1,
                           s1, d2, s2, d3, s3, a.ND);
        return hipGetLastError() == hipSuccess ? 0 : kPriaErrorDevice;
    }
    int sums(const Real* bx, const Real* by, const Real* bz, double* host) {
        constexpr int TW = PriaShape<Real>::TW;
        const int tiles = (a.T + TW - 1) / TW;
        hipLaunchKernelGGL(pria_check_slice_kernel<Real>, dim3(a.T), dim3(kPriaCheckBlock), 0,
                           stream, a, bx, slices);
        hipLaunchKernelGGL(pria_check_column_kernel<Real>, dim3(a.N * tiles), dim3(kPriaCheckBlock),
                           0, stream, a, by, bz, cols);
        hipLaunchKernelGGL(pria_sums_final, dim3(1), dim3(256), 0, stream, slices, a.T, cols,
                           a.N * tiles, dsums);
        if (hipGetLastError() != hipSuccess) return kPriaErrorDevice;
        if (hipMemcpyAsync(host, dsums, sizeof(double) * kPriaSums, hipMemcpyDeviceToHost,
                           stream) != hipSuccess)
            return kPriaErrorDevice;
        return hipStreamSynchronize(stream) == hipSuccess ? 0 : kPriaErrorDevice;
    }
};

template <class Real>
int pria_device_solve(const PriaJob<Real>& job, LogicState& st, double* work,
                      double time_limit_seconds, hipStream_t stream, bool resume) {
    const PriaArrays<Real>& a = job.a;
    const LogicParams& p = job.p;
    constexpr int TW = PriaShape<Real>::TW;
    const int tiles = (a.T + TW - 1) / TW;
    double* const dsums = work + size_t(a.T) * kPriaSums + size_t(a.N) * tiles;
    PriaDevice<Real> d{a, stream, work, work + size_t(a.T) * kPriaSums, dsums};
    if (d.place_zsplit(dsums + kPriaSums + 8)) return kPriaErrorDevice;
    if (pria_x_lds_bytes<Real>(a.T) > 65536) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(pria_x_kernel<Real>),
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                int(pria_x_lds_bytes<Real>(a.T))) != hipSuccess)
            return kPriaErrorDevice;
    }
    double sums[kPriaSums];
    auto t0 = std::chrono::steady_clock::now();
    LogicProgress progress;

    if (resume) {
        st.stalled = 0;
    } else {
        if (d.copy3(a.x0, a.xs, a.y0, a.ys, a.z0, a.zs) ||
            d.copy3(a.x, a.xs, a.y, a.ys, a.z, a.zs) ||
            d.copy3(a.sx, a.xs, a.sy, a.ys, a.sz, a.zs) || d.reset_thresholds())
            return kPriaErrorDevice;
        logic_init(st, p);
        if (d.step(pria_step_scalars<Real>(0, p.omega, p.eta, false, true))) return kPriaErrorDevice;
        if (d.sums(a.x, a.y, a.z, sums)) return kPriaErrorDevice;
        logic_after_warm_check(st, p, pria_check_sums(sums, job.scale));
    }

    for (;;) {
        const int action = st.action;
        const int64_t r0 = st.restart_iter;
        const double omega = st.omega;
        if (st.stalled) break;
        if (action == kActDoneX) break;
        if (action == kActDoneSX) {
            if (d.copy3(a.x, a.sx, a.y, a.sy, a.z, a.sz)) return kPriaErrorDevice;
            break;
        }
        if (action == kActWarmBlock) {
            for (int64_t j = 0; j < p.op_steps; ++j)
                if (d.step(pria_step_scalars<Real>(r0 + j, omega, p.eta, false, true)))
                    return kPriaErrorDevice;
            if (d.sums(a.x, a.y, a.z, sums)) return kPriaErrorDevice;
            st.restart_iter += p.op_steps;
            st.iteration += p.op_steps;
            logic_after_warm_check(st, p, pria_check_sums(sums, job.scale));
        } else if (action == kActHalpernBlock) {
            if (st.anchor_pending) {
                if (d.copy3(a.x0, a.x, a.y0, a.y, a.z0, a.z)) return kPriaErrorDevice;
                st.anchor_pending = 0;
            }
            for (int64_t j = 0; j < p.op_steps; ++j)
                if (d.step(pria_step_scalars<Real>(r0 + j, omega, p.eta, true,
                                                   j + 1 == p.op_steps)))
                    return kPriaErrorDevice;
            if (d.sums(a.sx, a.sy, a.sz, sums)) return kPriaErrorDevice;
            st.restart_iter += p.op_steps;
            st.iteration += p.op_steps;
            logic_after_halpern_check(st, p, pria_check_sums(sums, job.scale));
            if (p.verbose) logic_trace(st, "check");
        } else {
            if (d.step(pria_step_scalars<Real>(0, omega, p.eta, false, true))) return kPriaErrorDevice;
            if (d.copy3(a.x0, a.x, a.y0, a.y, a.z0, a.z)) return kPriaErrorDevice;
            if (d.sums(a.sx, a.sy, a.sz, sums)) return kPriaErrorDevice;
            logic_after_restart(st, p, pria_check_sums(sums, job.scale));
            if (p.verbose) logic_trace(st, "restart");
        }
        progress.maybe(st, "pria",
                       std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
        if (time_limit_seconds > 0.0 &&
            std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >
                time_limit_seconds &&
            st.action != kActDoneX && st.action != kActDoneSX) {
            st.action = (action == kActWarmBlock || action == kActRestart) ? kActDoneX : kActDoneSX;
            st.status = 2;
        }
    }
    return hipStreamSynchronize(stream) == hipSuccess ? 0 : kPriaErrorDevice;
}

template int pria_device_solve<float>(const PriaJob<float>&, LogicState&, double*, double,
                                      hipStream_t, bool);

// One vanilla step from (sx, sy, sz), in place, on the device tier (tier 2)
// or by one workgroup (tier 1); a test hook for the step kernels.
template <class Real>
__global__ void __launch_bounds__(1024) pria_wg_step_kernel(PriaArrays<Real> a, PriaStep<Real> st) {
    extern __shared__ char dyn[];
    __shared__ double red[kPriaYZRed];
    pria_wg_step(pria_view(a), st, reinterpret_cast<Real*>(dyn), red);
}

template <class Real>
int pria_one_step(const PriaArrays<Real>& a, double omega, double eta, int tier, hipStream_t stream) {
    PriaStep<Real> st = pria_step_scalars<Real>(0, omega, eta, false, true);
    {
        const int n = a.N * a.T;
        hipLaunchKernelGGL(pria_reset_thresholds_kernel<Real>, dim3((n + 255) / 256), dim3(256), 0,
                           stream, a);
    }
    if (tier == 1) {
        size_t lds = pria_x_lds_bytes<Real>(a.T);
        if (lds > 65536 &&
            hipFuncSetAttribute(reinterpret_cast<const void*>(pria_wg_step_kernel<Real>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)) != hipSuccess)
            return kPriaErrorDevice;
        hipLaunchKernelGGL(pria_wg_step_kernel<Real>, dim3(1), dim3(1024), lds, stream, a, st);
        return hipGetLastError() == hipSuccess ? 0 : kPriaErrorDevice;
    }
    PriaDevice<Real> d{a, stream, nullptr, nullptr, nullptr};
    if (pria_x_lds_bytes<Real>(a.T) > 65536 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(pria_x_kernel<Real>),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            int(pria_x_lds_bytes<Real>(a.T))) != hipSuccess)
        return kPriaErrorDevice;
    return d.step(st);
}

template int pria_one_step<float>(const PriaArrays<float>&, double, double, int, hipStream_t);
template int pria_one_step<double>(const PriaArrays<double>&, double, double, int, hipStream_t);
template int pria_device_solve<double>(const PriaJob<double>&, LogicState&, double*, double,
                                       hipStream_t, bool);

// precision switch: the whole Halpern state, FP32 -> FP64
__device__ inline void pria_convert_one(const PriaView<float>& f, const PriaView<double>& d,
                                        int e) {
    if (e < f.N * f.T) {
        d.thx[e] = f.thx[e];
        d.thy[e] = f.thy[e];
    }
    if (e < f.T) d.thz[e] = f.thz[e];
    d.x0[e] = f.x0[e];
    d.y0[e] = f.y0[e];
    d.z0[e] = f.z0[e];
    d.xc[e] = f.xc[e];
    d.yc[e] = f.yc[e];
    d.zc[e] = f.zc[e];
    d.x[e] = f.x[e];
    d.sx[e] = f.sx[e];
    d.y[e] = f.y[e];
    d.sy[e] = f.sy[e];
    d.z[e] = f.z[e];
    d.sz[e] = f.sz[e];
}

__global__ void pria_convert_kernel(PriaArrays<float> f, PriaArrays<double> d) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < f.ND) pria_convert_one(pria_view(f), pria_view(d), e);
}

__global__ void pria_convert_jobs_kernel(const PriaJob<float>* from, const PriaJob<double>* to) {
    const PriaView<float> f = pria_view(from[blockIdx.x].a);
    const PriaView<double> d = pria_view(to[blockIdx.x].a);
    for (int e = threadIdx.x; e < f.ND; e += blockDim.x) pria_convert_one(f, d, e);
}

int pria_convert_state(const PriaArrays<float>& from, const PriaArrays<double>& to,
                       hipStream_t stream) {
    hipLaunchKernelGGL(pria_convert_kernel, dim3((from.ND + 255) / 256), dim3(256), 0, stream, from,
                       to);
    return hipGetLastError() == hipSuccess ? 0 : kPriaErrorDevice;
}

int pria_convert_jobs(const PriaJob<float>* from, const PriaJob<double>* to, int count,
                      hipStream_t stream) {
    if (count <= 0) return 0;
    hipLaunchKernelGGL(pria_convert_jobs_kernel, dim3(count), dim3(1024), 0, stream, from, to);
    return hipGetLastError() == hipSuccess ? 0 : kPriaErrorDevice;
}

template <class Real>
int pria_launch_wg(const PriaJob<Real>* jobs, int count, int max_T, hipStream_t stream) {
    if (count <= 0) return 0;
    if (max_T > kPriaWgMaxT) {
        record_message("the one-workgroup PRIA tier takes T' <= " + std::to_string(kPriaWgMaxT) +
                       ", not " + std::to_string(max_T) + "; use the device tier");
        return kPriaErrorArgument;
    }
    size_t lds = pria_x_lds_bytes<Real>(max_T);
    if (lds > 65536) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(pria_wg_solve<Real>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)) != hipSuccess)
            return kPriaErrorDevice;
    }
    hipLaunchKernelGGL(pria_wg_solve<Real>, dim3(count), dim3(1024), lds, stream, jobs);
    return hipGetLastError() == hipSuccess ? 0 : kPriaErrorDevice;
}

template int pria_launch_wg<float>(const PriaJob<float>*, int, int, hipStream_t);
template int pria_launch_wg<double>(const PriaJob<double>*, int, int, hipStream_t);

}  // namespace reform
