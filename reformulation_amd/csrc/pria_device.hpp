// PRIA device code: the fused passes of one PDHG step on the JRP LP and the
// check-pass sums.  Shared by the one-workgroup whole-solve kernel and the
// device-wide kernels.
//
// Layout in HBM.  T' is T padded to even, N items.  Every per-entry array
// holds the upper triangle s <= t in "slice" order: slice s (one order period)
// is the N (T' - s) entries {(i, t) : t >= s}, item-major, so
//     idx(s, i, t) = off(s) + i (T' - s) + (t - s),
//     off(s)       = N (s T' - s (s - 1) / 2),
// and ND = N T' (T' + 1) / 2 entries in all.  In this order each projection
// segment of y (row (i, s), target K_i) and z (slice s, target K0) is
// contiguous; the x segments (column (i, t) over s <= t, target 1) are read
// in tiles of consecutive t, one coalesced row of the tile per s.
//
// Arrays: h (scaled holding cost), and for v in {x, y, z}: vs (the projected
// start), vc (the
// Halpern-combined point the next step starts from), v0 (anchor), v and sv
// (the last point and its step, written only when a check or restart
// follows), plus r = sigma (2 x2 - x1), which the x pass hands to the y/z pass.
//
// Per step:  x pass   x2 = P_X(x1 - tau (h + y1 + z1)), r, xc <- blend
//            yz pass  y2 = P_Y(y1 + r), z2 = P_Z(z1 + r), yc, zc <- blend
// Reference: _raw_step jrp/pria-jax/pria/solver.py:409-425, _reflect_step
// :428-447, bounds :347-402, canonical norm halpern_operator.py:5-19.
// Simplex projections use Michelot's method with FP64 accumulation: theta is
// (sum of active - target) / |active|, repeated until the active set stops
// shrinking; the same threshold the sort-based double_simplex_projection
// (jrp/pria-jax/pria/projection.py:71-130) computes.  A target <= 0 gives 0.
// Each segment keeps its last threshold (thx per column, thy per row, thz
// per slice; NaN = none) and the next step starts from the active set
// {v > last}: if that set lies inside the solution's support its theta is a
// lower bound of the answer, so one more pass lands on a superset of the
// support and the usual monotone shrinking follows; from a set that already
// contains the support the usual argument applies.  The answer is theta of
// the final active set either way, so the start changes the number of passes,
// not the result; between PDHG steps the set rarely changes and two passes
// usually suffice (against about seven from the full set on x columns).
#pragma once

#include <type_traits>

#include "common.hpp"
#include "halpern.hpp"

namespace reform {

template <class Real>
struct PriaArrays {
    int N, T;          // items, padded periods
    int ND;
    const Real* h;
    const Real* k;     // [N] scaled K_i
    Real k0;           // scaled K0
    const Real *xs, *ys, *zs;   // projected start
    Real *xc, *x0, *x, *sx;
    Real *yc, *y0, *y, *sy;
    Real *zc, *z0, *z, *sz;
    Real* r;
    double *thx, *thy, *thz;    // last thresholds: [N T'] columns, [N T'] rows, [T'] slices
};

// The device view of PriaArrays: the same arrays through global-memory
// pointers (see gptr); every device function works on a view.
template <class Real>
struct PriaView {
    int N, T;
    int ND;
    gptr<const Real> h;
    gptr<const Real> k;
    Real k0;
    gptr<const Real> xs, ys, zs;
    gptr<Real> xc, x0, x, sx;
    gptr<Real> yc, y0, y, sy;
    gptr<Real> zc, z0, z, sz;
    gptr<Real> r;
    gptr<double> thx, thy, thz;
};

template <class Real>
__device__ __forceinline__ PriaView<Real> pria_view(const PriaArrays<Real>& a) {
    PriaView<Real> v;
    v.N = a.N;
    v.T = a.T;
    v.ND = a.ND;
    v.h = to_global(a.h);
    v.k = to_global(a.k);
    v.k0 = a.k0;
    v.xs = to_global(a.xs);
    v.ys = to_global(a.ys);
    v.zs = to_global(a.zs);
    v.xc = to_global(a.xc);
    v.x0 = to_global(a.x0);
    v.x = to_global(a.x);
    v.sx = to_global(a.sx);
    v.yc = to_global(a.yc);
    v.y0 = to_global(a.y0);
    v.y = to_global(a.y);
    v.sy = to_global(a.sy);
    v.zc = to_global(a.zc);
    v.z0 = to_global(a.z0);
    v.z = to_global(a.z);
    v.sz = to_global(a.sz);
    v.r = to_global(a.r);
    v.thx = to_global(a.thx);
    v.thy = to_global(a.thy);
    v.thz = to_global(a.thz);
    return v;
}

__host__ __device__ inline int pria_off(int N, int T, int s) {
    return N * (s * T - (s * (s - 1)) / 2);
}

template <class Real>
struct PriaStep {
    Real lwn, rwn;     // weights of the next step (for the vc prediction)
    Real tau, sigma;
    bool reflect;      // start from vc (Halpern) or from sv (vanilla)
    bool store;        // write v, sv
};

template <class Real>
__host__ __device__ inline PriaStep<Real> pria_step_scalars(int64_t restart_iter, double omega,
                                                            double eta, bool reflect, bool store) {
    PriaStep<Real> s;
    double lwn, rwn;
    halpern_weights(reflect ? restart_iter + 1 : 0, &lwn, &rwn);
    s.lwn = Real(lwn);
    s.rwn = Real(rwn);
    Real re = Real(eta), ro = Real(omega);
    s.tau = re / ro;
    s.sigma = re * ro;
    s.reflect = reflect;
    s.store = store || !reflect;
    return s;
}

// The thresholds of a fresh solve: none (NaN), so the first projection of
// each segment starts from its full set; entries k, k + stride, ...
template <class Real>
__device__ inline void pria_reset_thresholds(const PriaView<Real>& a, int k, int stride) {
    const double none = __builtin_nan("");
    for (int e = k; e < a.N * a.T; e += stride) {
        a.thx[e] = none;
        a.thy[e] = none;
    }
    for (int e = k; e < a.T; e += stride) a.thz[e] = none;
}

// Rows of an x tile a thread loads as one batch (per G rows of the tile).
constexpr int kPriaXBatch = 8;

// --- x pass: one tile of TW consecutive columns t0.. of item i ------------
//
// lds: (T' x (TW + 1)) pre-projection values + TW thresholds.  With HOLD > 0
// a thread keeps x1 of its first HOLD batches of U rows in registers for the
// write phase instead of reading it again (the same values, so the same bits).
template <class Real, int TW, int HOLD = 0>
__device__ void pria_x_tile(const PriaView<Real>& a, const PriaStep<Real>& st, int i, int t0,
                            Real* lds) {
    const int T = a.T, N = a.N;
    const int BS = blockDim.x, tid = threadIdx.x;
    const int G = BS / TW;            // threads per column
    const int stride = TW + 1;
    Real* theta = lds + T * stride;
    const int tcount = min(TW, T - t0);
    const int smax = t0 + tcount - 1;
    const gptr<const Real> srcx = st.reflect ? a.xc : a.sx;
    const gptr<const Real> srcy = st.reflect ? a.yc : a.sy;
    const gptr<const Real> srcz = st.reflect ? a.zc : a.sz;

    // load: lanes over columns, G rows of s at a time (coalesced per s); U
    // rows' loads are issued before any is used, so their latencies overlap
    constexpr int U = kPriaXBatch;
    const int c = tid % TW, row = tid / TW, t = t0 + c;
    const bool colok = c < tcount;
    Real held[HOLD > 0 ? HOLD * U : 1];
    // (the loads are unconditional -- entries outside the tile read entry 0
    // and are dropped -- so the compiler issues the batch without branches)
    auto load_batch = [&](int s0, auto keep, Real* kept) {
        int g[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int s = s0 + u * G;
            g[u] = colok && s <= t ? pria_off(N, T, s) + i * (T - s) + (t - s) : 0;
        }
        Real vx[U], vh[U], vy[U], vz[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            vx[u] = srcx[g[u]];
            vh[u] = a.h[g[u]];
            vz[u] = srcz[g[u]];
            vy[u] = srcy[g[u]];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int s = s0 + u * G;
            if constexpr (decltype(keep)::value) kept[u] = vx[u];
            if (colok && s <= t) lds[s * stride + c] = vx[u] - st.tau * ((vh[u] + vz[u]) + vy[u]);
        }
    };
#pragma unroll
    for (int j = 0; j < HOLD; ++j)
        if (row + j * G * U <= smax) load_batch(row + j * G * U, std::true_type(), held + j * U);
    for (int s0 = row + HOLD * G * U; s0 <= smax; s0 += G * U)
        load_batch(s0, std::false_type(), nullptr);
    __syncthreads();
    // project: G lanes per column, spread across the wave (column = wave-local
    // index, q = the lane's share), reduced by xor over the q bits
    {
        const int cpw = kWave / G;                        // columns per wave
        const int lane = tid & (kWave - 1), wave = tid / kWave;
        const int c = wave * cpw + (lane % cpw), q = lane / cpw;
        const int t = t0 + c;
        const int len = c < tcount ? t + 1 : 0;
        const double start = len > 0 ? a.thx[i * T + t] : 0.0;
        double S = 0.0, C = 0.0, Sw = 0.0, Cw = 0.0;
        for (int s = q; s < len; s += G) {
            const double v = double(lds[s * stride + c]);
            S += v;
            C += 1.0;
            if (v > start) {
                Sw += v;
                Cw += 1.0;
            }
        }
        for (int off = cpw; off < kWave; off <<= 1) {
            S += __shfl_xor(S, off, kWave);
            C += __shfl_xor(C, off, kWave);
            Sw += __shfl_xor(Sw, off, kWave);
            Cw += __shfl_xor(Cw, off, kWave);
        }
        bool first = Cw > 0.0;  // a start (NaN compares false: none)
        if (first) {
            S = Sw;
            C = Cw;
        }
        double th = len > 0 ? (S - 1.0) / C : 0.0;
        bool done = len == 0;
        while (__any(!done)) {
            double S2 = 0.0, C2 = 0.0;
            if (!done) {
                for (int s = q; s < len; s += G) {
                    double v = double(lds[s * stride + c]);
                    if (v > th) {
                        S2 += v;
                        C2 += 1.0;
                    }
                }
            }
            for (int off = cpw; off < kWave; off <<= 1) {
                S2 += __shfl_xor(S2, off, kWave);
                C2 += __shfl_xor(C2, off, kWave);
            }
            if (!done) {
                if (C2 == C || C2 == 0.0 || (C2 > C && !first)) {
                    done = true;
                } else {
                    C = C2;
                    th = (S2 - 1.0) / C2;
                }
                first = false;
            }
        }
        if (q == 0 && c < TW) theta[c] = Real(th);
        if (q == 0 && len > 0) a.thx[i * T + t] = th;
    }
    __syncthreads();
    // write: x2, r, the next xc (and x, sx when storing); loads batched as above
    {
        const Real th = theta[c < TW ? c : 0];
        auto write_batch = [&](int s0, auto keep, const Real* kept) {
            int g[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int s = s0 + u * G;
                g[u] = colok && s <= t ? pria_off(N, T, s) + i * (T - s) + (t - s) : 0;
            }
            Real vx[U], v0[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if constexpr (decltype(keep)::value) vx[u] = kept[u];
                else vx[u] = srcx[g[u]];
            }
            if (st.reflect) {
#pragma unroll
                for (int u = 0; u < U; ++u) v0[u] = a.x0[g[u]];
            } else {
#pragma unroll
                for (int u = 0; u < U; ++u) v0[u] = vx[u];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int s = s0 + u * G;
                if (colok && s <= t) {
                    const int gu = g[u];
                    const Real x1 = vx[u];
                    const Real d = lds[s * stride + c] - th;
                    const Real x2 = d > Real(0) ? d : Real(0);
                    const Real refl = Real(2) * x2 - x1;
                    a.r[gu] = st.sigma * refl;
                    a.xc[gu] = st.lwn * refl + st.rwn * v0[u];
                    if (st.store) {
                        a.x[gu] = x1;
                        a.sx[gu] = x2;
                    }
                }
            }
        };
#pragma unroll
        for (int j = 0; j < HOLD; ++j)
            if (row + j * G * U <= smax)
                write_batch(row + j * G * U, std::true_type(), held + j * U);
        for (int s0 = row + HOLD * G * U; s0 <= smax; s0 += G * U)
            write_batch(s0, std::false_type(), nullptr);
    }
    __syncthreads();
}

// Michelot thresholds of the row segments a wavefront holds, one row per
// group of `width` lanes (16, 32 or 64), VR values per lane (e = lr + width k
// for the group lane lr), len valid per row; rows whose `live` is false only
// join the reductions.  The first active set is {u > start} when that is not
// empty, else all; every group stops when its set stops shrinking.
template <class Real, int VR>
__device__ __forceinline__ double group_michelot(const Real (&u)[VR], int len, int width, int lr,
                                                 bool live, double target, double start) {
    double S = 0.0, C = 0.0, Sw = 0.0, Cw = 0.0;
#pragma unroll
    for (int k = 0; k < VR; ++k) {
        const int e = lr + width * k;
        if (live && e < len) {
            const double v = double(u[k]);
            S += v;
            C += 1.0;
            if (v > start) {
                Sw += v;
                Cw += 1.0;
            }
        }
    }
    S = group_sum_dpp(S, width);
    C = group_sum_dpp(C, width);
    Sw = group_sum_dpp(Sw, width);
    Cw = group_sum_dpp(Cw, width);
    bool first = Cw > 0.0;
    if (first) {
        S = Sw;
        C = Cw;
    }
    double th = live ? (S - target) / C : 0.0;
    bool done = !live;
    for (int round = 0; round <= len && __any(!done); ++round) {
        double S2 = 0.0, C2 = 0.0;
#pragma unroll
        for (int k = 0; k < VR; ++k) {
            const int e = lr + width * k;
            const double v = double(u[k]);
            if (!done && e < len && v > th) {
                S2 += v;
                C2 += 1.0;
            }
        }
        S2 = group_sum_dpp(S2, width);
        C2 = group_sum_dpp(C2, width);
        if (!done) {
            if (C2 == C || C2 == 0.0 || (C2 > C && !first)) {
                done = true;
            } else {
                C = C2;
                th = (S2 - target) / C2;
            }
            first = false;
        }
    }
    return th;
}

// A y row longer than a lane group's registers hold (len > 64 VR, only with
// T' > 64 VR): one wavefront, values u = y1 + r read from memory on every
// Michelot pass (a row is at most T' values, resident in L2).  Lane lr sums
// its entries lr + 64 k in order of k and the wavefront's DPP sum follows,
// as in group_michelot; every tier runs rows of this length here, so they
// give the same bits.
template <class Real>
__device__ __forceinline__ void pria_y_row_memory(const PriaView<Real>& a,
                                                  const PriaStep<Real>& st,
                                                  gptr<const Real> srcy, int rb, int len,
                                                  int lane, bool row, int i, int s) {
    const double target = row ? double(a.k[i]) : 0.0;
    const bool live = row && target > 0.0;
    const double start = live ? a.thy[i * a.T + s] : 0.0;
    auto u_at = [&](int e) { return double(Real(srcy[rb + e] + a.r[rb + e])); };
    double S = 0.0, C = 0.0, Sw = 0.0, Cw = 0.0;
    if (live)
        for (int e = lane; e < len; e += kWave) {
            const double v = u_at(e);
            S += v;
            C += 1.0;
            if (v > start) {
                Sw += v;
                Cw += 1.0;
            }
        }
    S = wave_sum(S);
    C = wave_sum(C);
    Sw = wave_sum(Sw);
    Cw = wave_sum(Cw);
    bool first = Cw > 0.0;
    if (first) {
        S = Sw;
        C = Cw;
    }
    double th = live ? (S - target) / C : 0.0;
    bool done = !live;
    for (int round = 0; round <= len && !done; ++round) {   // wave-uniform: one row
        double S2 = 0.0, C2 = 0.0;
        for (int e = lane; e < len; e += kWave) {
            const double v = u_at(e);
            if (v > th) {
                S2 += v;
                C2 += 1.0;
            }
        }
        S2 = wave_sum(S2);
        C2 = wave_sum(C2);
        if (C2 == C || C2 == 0.0 || (C2 > C && !first)) {
            done = true;
        } else {
            C = C2;
            th = (S2 - target) / C2;
        }
        first = false;
    }
    if (live && lane == 0) a.thy[i * a.T + s] = th;
    if (!row) return;
    for (int e = lane; e < len; e += kWave) {
        const int g = rb + e;
        const Real v1 = srcy[g];
        const Real u = v1 + a.r[g];
        const Real an = st.reflect ? a.y0[g] : v1;
        const Real d = u - Real(th);
        const Real y2 = live ? (d > Real(0) ? d : Real(0)) : Real(0);
        a.yc[g] = st.lwn * (Real(2) * y2 - v1) + st.rwn * an;
        if (st.store) {
            a.y[g] = v1;
            a.sy[g] = y2;
        }
    }
}

// --- y/z pass: one slice s -------------------------------------------------
//
// part: 1 the y rows [i_lo, i_hi), 2 the z slice, 3 both; a block running
// only rows needs no LDS and no barrier.
// y rows: all N rows of the slice have len = T' - s entries; each row goes
// to a group of 16, 32 or 64 lanes (the smallest holding len in VR values
// per lane), so a wavefront projects 4, 2 or 1 rows at once; rows longer
// than 64 VR take pria_y_row_memory, one per wavefront.
// Widths follow kPriaRow16 / kPriaRow32, not VR.
// z slice: the whole workgroup, values in registers when the slice fits
// (N (T' - s) <= blockDim * VZ), else re-read from memory each round; its
// reductions take one barrier each.  `red`: 8 16 doubles of LDS.  `zcache`
// (LDS for `zcap` values, or null): on the memory path the slice's first
// zcap sums z + r are kept there for the Michelot rounds, which then re-read
// only the rest from memory (each thread reads back its own entries; the
// values and their order are the same either way).
// LDS doubles pria_yz_slice's `red` must hold.
constexpr int kPriaYZRed = 8 * 16;
// Dynamic LDS of a z-slice workgroup caching a long slice's z + r for its
// Michelot rounds; with `red`, three 512-thread workgroups still fit a CU.
constexpr int kPriaZCacheBytes = 48 * 1024;

// Row group widths by row length, the same for every register shape: a row
// of at most 80 entries takes 16 lanes, at most 160 takes 32, longer rows a
// whole wavefront.  VR (values per lane) only has to hold the row, so every
// shape with VR >= ceil(len / width) gives the same sums bit for bit; so does
// VZ (the register path and the memory path add a thread's values in the
// same order).
constexpr int kPriaRow16 = 80, kPriaRow32 = 160;

template <class Real, int VR, int VZ>
__device__ void pria_yz_slice(const PriaView<Real>& a, const PriaStep<Real>& st, int s,
                              double* red, int part = 3, int i_lo = 0, int i_hi = 1 << 30,
                              Real* zcache = nullptr, int zcap = 0) {
    const int T = a.T, N = a.N;
    const int BS = blockDim.x, tid = threadIdx.x;
    const int lane = tid & (kWave - 1), wave = tid / kWave, waves = BS / kWave;
    const int len = T - s;
    const int base = pria_off(N, T, s);
    const gptr<const Real> srcy = st.reflect ? a.yc : a.sy;
    const gptr<const Real> srcz = st.reflect ? a.zc : a.sz;

    // ---- y: rows [i_lo, i_hi) of the slice, 64 / width rows per wavefront
    // at a time ----
    const int width = len <= kPriaRow16 ? 16 : (len <= kPriaRow32 ? 32 : 64);
    const int rpw = kWave / width, lr = lane % width, sub = lane / width;
    const int rows_end = (part & 1) ? min(N, i_hi) : 0;
    const bool long_rows = len > kWave * VR;   // block-uniform: every row of the slice
    // a block running the whole slice parks the r values the rows read (the
    // slice's entries e < rcap) in zcache for the z part, which then reads
    // them from LDS instead of memory (with the compact register shapes, 64
    // bytes of slice values per thread: the wider ones would spill)
    constexpr bool park = VZ * int(sizeof(Real)) <= 64;
    const int rcap =
        (park && part == 3 && zcache != nullptr && !long_rows) ? min(zcap, N * len) : 0;
    for (int i0 = i_lo + wave * rpw; i0 < rows_end; i0 += waves * rpw) {
        const int i = i0 + sub;
        const bool row = i < rows_end;
        const int rb = base + (row ? i : 0) * len;
        if (long_rows) {   // width 64: one row per wavefront
            pria_y_row_memory<Real>(a, st, srcy, rb, len, lane, row, i, s);
            continue;
        }
        // unconditional loads (entries past the row read the row's first
        // entry and are dropped), issued as one batch
        int g[VR];
#pragma unroll
        for (int k = 0; k < VR; ++k) {
            const int e = lr + width * k;
            g[k] = rb + (row && e < len ? e : 0);
        }
        Real y1[VR], u[VR], an[VR];
#pragma unroll
        for (int k = 0; k < VR; ++k) {
            y1[k] = srcy[g[k]];
            u[k] = a.r[g[k]];
        }
        if (st.reflect) {
#pragma unroll
            for (int k = 0; k < VR; ++k) an[k] = a.y0[g[k]];
        } else {
#pragma unroll
            for (int k = 0; k < VR; ++k) an[k] = y1[k];
        }
        if (park && rcap > 0) {
#pragma unroll
            for (int k = 0; k < VR; ++k) {
                const int e = lr + width * k, es = (rb - base) + e;
                if (row && e < len && es < rcap) zcache[es] = u[k];
            }
        }
#pragma unroll
        for (int k = 0; k < VR; ++k) {
            const bool in = row && lr + width * k < len;
            u[k] = in ? y1[k] + u[k] : Real(0);
            y1[k] = in ? y1[k] : Real(0);
            an[k] = in ? an[k] : Real(0);
        }
        const double target = row ? double(a.k[i]) : 0.0;
        const bool live = row && target > 0.0;
        const double th = group_michelot<Real, VR>(u, len, width, lr, live, target,
                                                   live ? a.thy[i * T + s] : 0.0);
        if (live && lr == 0) a.thy[i * T + s] = th;
#pragma unroll
        for (int k = 0; k < VR; ++k) {
            const int e = lr + width * k;
            if (row && e < len) {
                const int g = rb + e;
                const Real v1 = y1[k];
                const Real d = u[k] - Real(th);
                const Real y2 = live ? (d > Real(0) ? d : Real(0)) : Real(0);
                a.yc[g] = st.lwn * (Real(2) * y2 - v1) + st.rwn * an[k];
                if (st.store) {
                    a.y[g] = v1;
                    a.sy[g] = y2;
                }
            }
        }
    }

    // ---- z: the whole slice ----
    if (!(part & 2)) return;
    if (park && rcap > 0) __syncthreads();   // the rows' r values are in zcache
    const int L = N * len;
    const double target = double(a.k0);
    double th = 0.0;
    int parity = 0;   // block_sums_alt's LDS half
    int ucap = 0;     // the memory path kept u = z1 + r of entries e < ucap in zcache
    // register path: u = z1 + r of entry tid + BS k in u[k]; with `kept`
    // (the slice fits the LDS cache) z1 of entry e also in zcache[e], and the
    // register path writes the slice itself, reading neither z1 nor r again
    bool kept = false;
    if (target > 0.0) {
        const double start = a.thz[s];
        bool first;
        if (L <= BS * VZ) {
            kept = zcache != nullptr && L <= zcap;
            Real u[VZ];
            double v4[4] = {0.0, 0.0, 0.0, 0.0};
            // unconditional loads up to the last k any thread needs (a
            // block-uniform bound); entries past the slice read its first
            // entry and are dropped.  r from zcache when the rows parked the
            // whole slice's (each thread then overwrites its own entries with
            // z1 below)
            const int kn = (L + BS - 1) / BS;
            Real rv[VZ];
            if (park && L <= rcap) {
#pragma unroll
                for (int k = 0; k < VZ; ++k) {
                    const int e = tid + BS * k;
                    const int el = e < L ? e : 0;
                    u[k] = Real(0);
                    rv[k] = Real(0);
                    if (k < kn) {
                        u[k] = srcz[base + el];
                        rv[k] = zcache[el];
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < VZ; ++k) {
                    const int e = tid + BS * k;
                    const int g = base + (e < L ? e : 0);
                    u[k] = Real(0);
                    rv[k] = Real(0);
                    if (k < kn) {
                        u[k] = srcz[g];
                        rv[k] = a.r[g];
                    }
                }
            }
            if (kept) {
#pragma unroll
                for (int k = 0; k < VZ; ++k) {
                    const int e = tid + BS * k;
                    if (e < L) zcache[e] = u[k];
                }
            }
#pragma unroll
            for (int k = 0; k < VZ; ++k) u[k] += rv[k];
#pragma unroll
            for (int k = 0; k < VZ; ++k) {
                const int e = tid + BS * k;
                if (e < L) {
                    const double v = double(u[k]);
                    v4[0] += v;
                    v4[1] += 1.0;
                    if (v > start) {
                        v4[2] += v;
                        v4[3] += 1.0;
                    }
                }
            }
            block_sums_alt<4>(v4, red, parity);
            first = v4[3] > 0.0;
            double C = first ? v4[3] : v4[1];
            th = ((first ? v4[2] : v4[0]) - target) / C;
            for (int round = 0; round <= L; ++round) {
                double w2[2] = {0.0, 0.0};
                // u as the rounds' input, so its FP64 widening is redone each
                // round rather than held (2 VZ registers) beside u for the write
                if (sizeof(Real) == 4) {
#pragma unroll
                    for (int k = 0; k < VZ; ++k) asm volatile("" : "+v"(u[k]));
                }
#pragma unroll
                for (int k = 0; k < VZ; ++k) {
                    const int e = tid + BS * k;
                    const double v = double(u[k]);
                    if (e < L && v > th) {
                        w2[0] += v;
                        w2[1] += 1.0;
                    }
                }
                block_sums_alt<2>(w2, red, parity);
                if (w2[1] == C || w2[1] == 0.0 || (w2[1] > C && !first)) break;
                C = w2[1];
                th = (w2[0] - target) / w2[1];
                first = false;
            }
            if (kept) {
                if (tid == 0) a.thz[s] = th;
                // the write from the registers and the cache; only z0 is read
                // (when reflecting).  d is u - thr with u = z1 + r rounded as
                // the reloading write below rounds it: the same bits.
                const Real thr = Real(th);
                constexpr int U = 4;
#pragma unroll
                for (int k0 = 0; k0 < VZ; k0 += U) {
                    if (tid + BS * k0 >= L) break;
                    Real an[U];
#pragma unroll
                    for (int q = 0; q < U; ++q) {
                        const int e = tid + BS * (k0 + q);
                        an[q] = st.reflect ? a.z0[base + (e < L ? e : 0)] : Real(0);
                    }
#pragma unroll
                    for (int q = 0; q < U; ++q) {
                        const int e = tid + BS * (k0 + q);
                        if (e < L) {
                            const int g = base + e;
                            const Real v1 = zcache[e];
                            const Real d = u[k0 + q] - thr;
                            const Real z2 = d > Real(0) ? d : Real(0);
                            a.zc[g] = st.lwn * (Real(2) * z2 - v1) +
                                      st.rwn * (st.reflect ? an[q] : v1);
                            if (st.store) {
                                a.z[g] = v1;
                                a.sz[g] = z2;
                            }
                        }
                    }
                }
            }
        } else {
            double v4[4] = {0.0, 0.0, 0.0, 0.0};
            if (!zcache) zcap = 0;
            if (park) ucap = zcap;
            for (int e = tid; e < L; e += BS) {
                const Real u =
                    srcz[base + e] + (park && e < rcap ? zcache[e] : Real(a.r[base + e]));
                if (e < zcap) zcache[e] = u;
                const double v = double(u);
                v4[0] += v;
                v4[1] += 1.0;
                if (v > start) {
                    v4[2] += v;
                    v4[3] += 1.0;
                }
            }
            block_sums_alt<4>(v4, red, parity);
            first = v4[3] > 0.0;
            double C = first ? v4[3] : v4[1];
            th = ((first ? v4[2] : v4[0]) - target) / C;
            for (int round = 0; round <= L; ++round) {
                double w2[2] = {0.0, 0.0};
                for (int e = tid; e < L; e += BS) {
                    const double v = double(e < zcap ? zcache[e]
                                                     : Real(srcz[base + e] + a.r[base + e]));
                    if (v > th) {
                        w2[0] += v;
                        w2[1] += 1.0;
                    }
                }
                block_sums_alt<2>(w2, red, parity);
                if (w2[1] == C || w2[1] == 0.0 || (w2[1] > C && !first)) break;
                C = w2[1];
                th = (w2[0] - target) / w2[1];
                first = false;
            }
        }
        if (tid == 0) a.thz[s] = th;
    }
    if (kept) {   // written by the register path
        __syncthreads();
        return;
    }
    const Real thr = Real(th);
    // write: loads of U entries issued before any is used; with the compact
    // shapes u = z1 + r from the cache where the memory path kept it (the
    // same bits), and r not read at all without a projection (z2 = 0)
    constexpr int U = 4;
    for (int e0 = tid; e0 < L; e0 += BS * U) {
        int g[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int e = e0 + BS * q;
            g[q] = base + (e < L ? e : 0);
        }
        Real v1[U], rr[U], an[U];
#pragma unroll
        for (int q = 0; q < U; ++q) {
            v1[q] = srcz[g[q]];
            rr[q] = Real(0);
        }
        if (!park || (target > 0.0 && e0 + BS * (U - 1) >= ucap)) {
#pragma unroll
            for (int q = 0; q < U; ++q) rr[q] = a.r[g[q]];
        }
        if (st.reflect) {
#pragma unroll
            for (int q = 0; q < U; ++q) an[q] = a.z0[g[q]];
        } else {
#pragma unroll
            for (int q = 0; q < U; ++q) an[q] = v1[q];
        }
#pragma unroll
        for (int q = 0; q < U; ++q) {
            const int e = e0 + BS * q;
            if (e < L) {
                const int g = base + e;
                const Real d = (park && e < ucap ? zcache[e] : Real(v1[q] + rr[q])) - thr;
                const Real z2 = target > 0.0 ? (d > Real(0) ? d : Real(0)) : Real(0);
                a.zc[g] = st.lwn * (Real(2) * z2 - v1[q]) + st.rwn * an[q];
                if (st.store) {
                    a.z[g] = v1[q];
                    a.sz[g] = z2;
                }
            }
        }
    }
    __syncthreads();
}

// Tile and register shapes.  x tiles of 128-byte rows: 32 FP32 or 16 FP64
// columns, so a 512-thread tile of T' = 450 rows takes 60 KB of LDS and two
// fit a CU.  VR: row values per lane (longer rows from memory); VZ: slice values per
// thread held in registers.  The compact shape (T' <= 320) keeps the
// batched y/z kernel within 80 VGPRs, three 512-thread blocks per CU.
template <class Real> struct PriaShape;
template <> struct PriaShape<float> { static constexpr int TW = 32, VR = 10, VZ = 32; };
template <> struct PriaShape<double> { static constexpr int TW = 16, VR = 10, VZ = 32; };
template <class Real> struct PriaCompactShape;
template <> struct PriaCompactShape<float> { static constexpr int VR = 5, VZ = 16; };
template <> struct PriaCompactShape<double> { static constexpr int VR = 5, VZ = 8; };
constexpr int kPriaCompactT = 64 * 5;   // largest T' the compact shape holds

template <class Real>
__host__ __device__ inline size_t pria_x_lds_bytes(int T) {
    return sizeof(Real) * (size_t(T) * (PriaShape<Real>::TW + 1) + PriaShape<Real>::TW);
}

// --- check sums ------------------------------------------------------------

enum PriaSum {
    kPPNorm = 0,   // sum (x - sx)^2
    kPDNorm = 1,   // sum (y - sy)^2 + (z - sz)^2
    kPCross = 2,   // sum (x - sx)((y - sy) + (z - sz))
    kPPDrift = 3,  // sum (x - x0)^2
    kPDDrift = 4,  // sum (y - y0)^2 + (z - z0)^2
    kPUpper = 5,   // sum_i K_i max_row bx + K0 max_slice bx + <h, bx>
    kPLower = 6,   // sum over columns of min (h + by + bz)
    kPriaSums = 7
};

// Slice part of the check (everything but the column minima), by one
// workgroup; returns the six sums in acc[0..5] of thread 0 (all threads hold
// the totals after block_sums).  Bounds at (bx): upper bound terms.
template <class Real>
__device__ void pria_slice_sums(const PriaView<Real>& a, int s, gptr<const Real> bx, double* red,
                                double (&acc)[kPriaSums]) {
    const int T = a.T, N = a.N;
    const int BS = blockDim.x, tid = threadIdx.x;
    const int lane = tid & (kWave - 1), wave = tid / kWave, waves = BS / kWave;
    const int len = T - s, base = pria_off(N, T, s), L = N * len;
    double v[6] = {0, 0, 0, 0, 0, 0};
    double smax = -INFINITY;
    for (int e = tid; e < L; e += BS) {
        int g = base + e;
        double dx = double(a.x[g]) - double(a.sx[g]);
        double dy = double(a.y[g]) - double(a.sy[g]);
        double dz = double(a.z[g]) - double(a.sz[g]);
        v[0] += dx * dx;
        v[1] += dy * dy + dz * dz;
        v[2] += dx * dy + dx * dz;
        double rx = double(a.x[g]) - double(a.x0[g]);
        double ry = double(a.y[g]) - double(a.y0[g]);
        double rz = double(a.z[g]) - double(a.z0[g]);
        v[3] += rx * rx;
        v[4] += ry * ry + rz * rz;
        double b = bx[g];
        v[5] += double(a.h[g]) * b;
        smax = b > smax ? b : smax;
    }
    // row maxima: one wave per row, K_i max_t bx
    double rows = 0.0;
    for (int i = wave; i < N; i += waves) {
        double m = -INFINITY;
        for (int e = lane; e < len; e += kWave) {
            double b = bx[base + i * len + e];
            m = b > m ? b : m;
        }
        m = wave_max(m);
        if (lane == 0) rows += double(a.k[i]) * m;
    }
    v[5] += rows;
    double gmax = block_max1(smax, red);
    block_sums<6>(v, red);
    for (int q = 0; q < 5; ++q) acc[q] += v[q];
    acc[kPUpper] += v[5] + double(a.k0) * gmax;
}

// Column minima of (h + by + bz) over s <= t for the tile of item i starting
// at t0 (TW columns); adds the tile's sum of minima to acc[kPLower].
template <class Real, int TW>
__device__ void pria_column_sums(const PriaView<Real>& a, int i, int t0, gptr<const Real> by,
                                 gptr<const Real> bz, double* red, double (&acc)[kPriaSums]) {
    const int T = a.T, N = a.N;
    const int BS = blockDim.x, tid = threadIdx.x;
    const int G = BS / TW;
    const int c = tid % TW, row = tid / TW, t = t0 + c;
    const int tcount = min(TW, T - t0);
    double m = INFINITY;
    if (c < tcount) {
        for (int s = row; s <= t; s += G) {
            int g = pria_off(N, T, s) + i * (T - s) + (t - s);
            double v = double(a.h[g]) + double(by[g]) + double(bz[g]);
            m = v < m ? v : m;
        }
    }
    // min over the G rows of each column: LDS reduce by column
    __shared__ double colmin[1024];
    colmin[tid] = m;
    __syncthreads();
    double total = 0.0;
    if (tid < TW && tid < tcount) {
        double mm = colmin[tid];
        for (int q = 1; q < G; ++q) {
            double o = colmin[q * TW + tid];
            mm = o < mm ? o : mm;
        }
        total = mm;
    }
    __syncthreads();
    total = block_sum1(total, red);
    acc[kPLower] += total;
}

__host__ __device__ inline CheckSums pria_check_sums(const double* s, double scale) {
    CheckSums c;
    c.pnorm = s[kPPNorm];
    c.dnorm = s[kPDNorm];
    c.cross = s[kPCross];
    c.pdrift = s[kPPDrift];
    c.ddrift = s[kPDDrift];
    c.upper = s[kPUpper] * scale;
    c.lower = s[kPLower] * scale;
    return c;
}

}  // namespace reform
