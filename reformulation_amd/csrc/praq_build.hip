// The PRAQ instance built on the device from a COO instance in device memory
// (praq_solve_device): validation, the objective scale, the half-edge CSR
// (a stable radix sort of the half-edges by node), degrees, hub lists,
// scaled data and the cold start -- the same arrays, bit for bit, as the
// host build (praq_build_host) writes, so the solve that follows is the
// host entry's.  Only a few scalars cross PCIe.
#include "praq_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace reform {

namespace {

constexpr int kBuildBlock = 256;
// scan: 256 threads x 8 items per tile
constexpr int kScanItems = 8;
constexpr int kScanTile = kBuildBlock * kScanItems;
// radix sort: 8-bit digits, 256 threads x 8 items per tile
constexpr int kSortItems = 8;
constexpr int kSortTile = kBuildBlock * kSortItems;
constexpr int kRadixBins = 256;

int blocks_for(int64_t count, int per) { return int((count + per - 1) / per); }

// ---- validation, narrowing and degree counts ------------------------------

__global__ void __launch_bounds__(kBuildBlock)
praq_dev_edges(const int64_t* rows, const int64_t* cols, int m, int64_t n, int* e0, int* e1,
               int* count, int* bad) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int64_t i = rows[e], j = cols[e];
    if (i < 0 || i >= n || j < 0 || j >= n || i == j) {
        atomicOr(bad, 1);
        return;
    }
    e0[e] = int(i);
    e1[e] = int(j);
    atomicAdd(&count[i], 1);
    atomicAdd(&count[j], 1);
}

// ---- the objective scale: praq_sumsq_blocked's order ----------------------

// One thread per chunk of kPraqScaleChunk terms, summed in order; which = 0
// the diagonal's d_i d_i, 1 the weights' (4 v) v.
__global__ void __launch_bounds__(kBuildBlock)
praq_dev_sumsq_chunks(const double* v, int64_t len, int which, double* partial) {
#pragma clang fp contract(off)
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t c0 = c * kPraqScaleChunk;
    if (c0 >= len) return;
    const int64_t c1 = c0 + kPraqScaleChunk < len ? c0 + kPraqScaleChunk : len;
    double part = 0.0;
    for (int64_t k = c0; k < c1; ++k) part += which ? 4.0 * v[k] * v[k] : v[k] * v[k];
    partial[c] = part;
}

// One thread: the chunk sums in order, diagonal then weights (the host
// takes the square root, as the host build does).
__global__ void praq_dev_squares(const double* pd, int nd, const double* pw, int nw,
                                 double* squares) {
#pragma clang fp contract(off)
    double d = 0.0, w = 0.0;
    for (int c = 0; c < nd; ++c) d += pd[c];
    for (int c = 0; c < nw; ++c) w += pw[c];
    *squares = d + w;
}

// ---- exclusive scan of int32, in place -------------------------------------

// Each tile of kScanTile values: an exclusive scan in place (each thread's
// kScanItems consecutive values in order, the thread sums scanned across
// the block), the tile's total into sums[tile].
__global__ void __launch_bounds__(kBuildBlock) scan_tiles(int* data, int64_t count, int* sums) {
    __shared__ int part[kBuildBlock];
    const int tid = threadIdx.x;
    const int64_t at = int64_t(blockIdx.x) * kScanTile + int64_t(tid) * kScanItems;
    int v[kScanItems];
    int mine = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = at + j < count ? data[at + j] : 0;
        mine += v[j];
    }
    part[tid] = mine;
    __syncthreads();
    // inclusive Hillis-Steele scan of the thread sums
    for (int off = 1; off < kBuildBlock; off <<= 1) {
        const int add = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += add;
        __syncthreads();
    }
    int run = part[tid] - mine;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        if (at + j < count) data[at + j] = run;
        run += v[j];
    }
    if (tid == kBuildBlock - 1) sums[blockIdx.x] = part[tid];
}

__global__ void __launch_bounds__(kBuildBlock) scan_add(int* data, int64_t count, const int* sums) {
    const int64_t at = int64_t(blockIdx.x) * kScanTile;
    const int add = sums[blockIdx.x];
    for (int j = threadIdx.x; j < kScanTile; j += kBuildBlock)
        if (at + j < count) data[at + j] += add;
}

// Scratch ints an exclusive scan of `count` values needs for its tile sums,
// all levels.
size_t scan_scratch(int64_t count) {
    size_t total = 0;
    int64_t level = count;
    while (level > 1) {
        level = blocks_for(level, kScanTile);
        total += size_t(level);
    }
    return total + 1;
}

int exclusive_scan(int* data, int64_t count, int* scratch, hipStream_t s) {
    if (count <= 0) return 0;
    const int nb = blocks_for(count, kScanTile);
    int* sums = scratch;
    hipLaunchKernelGGL(scan_tiles, dim3(nb), dim3(kBuildBlock), 0, s, data, count, sums);
    if (nb > 1) {
        if (exclusive_scan(sums, nb, scratch + nb, s)) return kErrorDevice;
        hipLaunchKernelGGL(scan_add, dim3(nb), dim3(kBuildBlock), 0, s, data, count, sums);
    }
    REFORM_HIP_CHECK(hipGetLastError());
    return 0;
}

// ---- stable LSD radix sort of (key, value) int pairs by key ---------------

// hist[d * tiles + b]: how many keys of tile b have digit d
__global__ void __launch_bounds__(kBuildBlock)
radix_count(const int* keys, int64_t count, int shift, int tiles, int* hist) {
    __shared__ int h[kRadixBins];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t t0 = int64_t(blockIdx.x) * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
        const int64_t k = t0 + r * kBuildBlock + threadIdx.x;
        if (k < count) atomicAdd(&h[(keys[k] >> shift) & (kRadixBins - 1)], 1);
    }
    __syncthreads();
    hist[size_t(threadIdx.x) * tiles + blockIdx.x] = h[threadIdx.x];
}

// Tile b's keys to their places: offset[d * tiles + b] (exclusive scan of
// hist) plus the key's rank among the tile's earlier keys of its digit.  A
// tile is taken in kSortItems rounds of kBuildBlock keys (tile order: round,
// then thread); within a round a key's rank is its rank among its
// wavefront's lanes of the same digit (lanes matched by eight ballots)
// after the earlier wavefronts' counts of that digit.
__global__ void __launch_bounds__(kBuildBlock)
radix_place(const int* keys, const int* vals, int64_t count, int shift, int tiles,
            const int* offset, int* keys_out, int* vals_out) {
    constexpr int kWaves = kBuildBlock / kWave;
    __shared__ int base[kRadixBins];
    __shared__ int wcount[kWaves][kRadixBins];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    base[tid] = offset[size_t(tid) * tiles + blockIdx.x];
    const unsigned long long below = lane ? (~0ull >> (kWave - lane)) : 0ull;
    const int64_t t0 = int64_t(blockIdx.x) * kSortTile;
    for (int r = 0; r < kSortItems; ++r) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) wcount[w][tid] = 0;
        __syncthreads();
        const int64_t k = t0 + r * kBuildBlock + tid;
        const bool valid = k < count;
        const int key = valid ? keys[k] : 0;
        const int d = (key >> shift) & (kRadixBins - 1);
        unsigned long long same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool set = (d >> b) & 1;
            const unsigned long long ones = __ballot(set);
            same &= set ? ones : ~ones;
        }
        const int rank = __popcll(same & below);
        if (valid && rank == 0) wcount[wave][d] = __popcll(same);
        __syncthreads();
        if (valid) {
            int pos = base[d] + rank;
            for (int w = 0; w < wave; ++w) pos += wcount[w][d];
            keys_out[pos] = key;
            vals_out[pos] = vals[k];
        }
        __syncthreads();
        int round_total = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) round_total += wcount[w][tid];
        base[tid] += round_total;
        __syncthreads();
    }
}

// ---- the half-edges, degrees, hubs, data and start -------------------------

// element q = 2e + slot of the edge-ordered half-edge sequence: key the node,
// value the half-edge (slot m + e for slot 1)
__global__ void __launch_bounds__(kBuildBlock)
praq_dev_halves(const int* e0, const int* e1, int m, int* keys, int* vals) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    keys[2 * size_t(e)] = e0[e];
    vals[2 * size_t(e)] = e;
    keys[2 * size_t(e) + 1] = e1[e];
    vals[2 * size_t(e) + 1] = m + e;
}

// per node: degree (floored at 1, or n - 1 in the dense mode), the start
// 0.5, hub flag and chunk count (into n + 1 arrays whose last entry is 0)
__global__ void __launch_bounds__(kBuildBlock)
praq_dev_nodes(const int* ptr, const double* diag, int n, int dense, int heavy_degree,
               const double* scale, double* sdiag, double* deg, double* xs, int* hub,
               int* chunks) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        hub[n] = 0;
        chunks[n] = 0;
        return;
    }
    const int len = ptr[i + 1] - ptr[i];
    const double incident = dense ? double(n - 1) : double(len);
    deg[i] = incident > 1.0 ? incident : 1.0;
    sdiag[i] = diag[i] / *scale;
    xs[i] = 0.5;
    const bool h = len > heavy_degree;
    hub[i] = h;
    chunks[i] = h ? (len + kPraqChunk - 1) / kPraqChunk : 0;
}

// hub lists: heavy[h] = node, hchunk[h] = its first chunk, chunk_node[c] = h
__global__ void __launch_bounds__(kBuildBlock)
praq_dev_hubs(const int* ptr, int n, int heavy_degree, const int* hpos, const int* cpos,
              int* heavy, int* hchunk, int* chunk_node) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int len = ptr[i + 1] - ptr[i];
    if (len <= heavy_degree) return;
    const int h = hpos[i], c0 = cpos[i];
    heavy[h] = i;
    hchunk[h] = c0;
    for (int c = 0; c < (len + kPraqChunk - 1) / kPraqChunk; ++c) chunk_node[c0 + c] = h;
}

// per edge: w = 2 W / scale (signed) and the cold-start dual, -W on both
// slots projected onto the edge's set (project_edge_host's operations)
__global__ void __launch_bounds__(kBuildBlock)
praq_dev_edge_data(const double* vals, int m, const double* scale, double* w, double* ys) {
#pragma clang fp contract(off)
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const double we = 2.0 * vals[e] / *scale;
    w[e] = we;
    const double aw = fabs(we);
    double v0 = -0.5 * we, v1 = v0;
    if (we > 0.0) {
        double mu = 0.5 * (v0 + v1);
        mu = mu < -aw ? -aw : (mu > 0.0 ? 0.0 : mu);
        v0 = v1 = mu;
    } else {
        const double line = v0 + 0.5 * (aw - v0 - v1);
        v0 = line < 0.0 ? 0.0 : (line > aw ? aw : line);
        v1 = aw - v0;
    }
    ys[e] = v0;
    ys[size_t(m) + e] = v1;
}

template <class Real>
__global__ void __launch_bounds__(kBuildBlock) narrow(const double* src, size_t count, Real* dst) {
    for (size_t q = size_t(blockIdx.x) * blockDim.x + threadIdx.x; q < count;
         q += size_t(gridDim.x) * blockDim.x)
        dst[q] = Real(src[q]);
}

}  // namespace

PraqDeviceBuild::~PraqDeviceBuild() {
    if (arena) (void)hipFreeAsync(arena, stream);
}

int praq_build_device(const PraqInstance& in, int dense_degree, int lanes, hipStream_t s,
                      PraqDeviceBuild& b, PraqHostInstance& shape) {
    const int64_t limit = std::numeric_limits<int>::max();
    if (in.n <= 0 || in.n > limit || in.m < 0 || in.m > limit / 2) return kErrorInstance;
    if (!in.diag || (in.m > 0 && (!in.rows || !in.cols || !in.vals))) return kErrorInstance;
    const int n = int(in.n), m = int(in.m);
    const int64_t halves = 2 * int64_t(m);
    const int tiles = blocks_for(halves, kSortTile);
    const int chunks_d = blocks_for(n, kPraqScaleChunk), chunks_w = blocks_for(m, kPraqScaleChunk);
    // one stream-ordered arena: the image arrays, then work space
    const size_t nn = size_t(n), mm = size_t(m), h2 = size_t(halves);
    size_t at = 0;
    auto carve = [&](size_t bytes) {
        const size_t off = at;
        at += align256(bytes ? bytes : 1);
        return off;
    };
    const size_t o_e0 = carve(4 * mm), o_e1 = carve(4 * mm), o_ptr = carve(4 * (nn + 1)),
                 o_half = carve(4 * h2), o_w = carve(8 * mm), o_diag = carve(8 * nn),
                 o_deg = carve(8 * nn), o_xs = carve(8 * nn), o_ys = carve(8 * h2),
                 o_hub = carve(4 * (nn + 1)), o_chunks = carve(4 * (nn + 1)),
                 o_keys = carve(4 * h2), o_keys2 = carve(4 * h2), o_vals2 = carve(4 * h2),
                 o_hist = carve(4 * size_t(kRadixBins) * tiles),
                 o_scan = carve(4 * std::max({scan_scratch(nn + 1),
                                              scan_scratch(int64_t(kRadixBins) * tiles)})),
                 o_part = carve(8 * size_t(chunks_d + chunks_w)), o_scalars = carve(64);
    b.stream = s;
    REFORM_HIP_CHECK(device_malloc_async(&b.arena, at, s));
    char* base = static_cast<char*>(b.arena);
    auto I = [&](size_t off) { return reinterpret_cast<int*>(base + off); };
    auto D = [&](size_t off) { return reinterpret_cast<double*>(base + off); };
    b.e0 = I(o_e0);
    b.e1 = I(o_e1);
    b.ptr = I(o_ptr);
    b.half = I(o_half);
    b.w = D(o_w);
    b.diag = D(o_diag);
    b.deg = D(o_deg);
    b.xs = D(o_xs);
    b.ys = D(o_ys);
    int* count = I(o_ptr);          // degrees, scanned into ptr in place
    int* bad = I(o_scalars);
    double* scale = D(o_scalars + 8);
    int* totals = I(o_scalars + 16);   // nh, nc
    const dim3 blk(kBuildBlock);
    REFORM_HIP_CHECK(hipMemsetAsync(count, 0, 4 * (nn + 1), s));
    REFORM_HIP_CHECK(hipMemsetAsync(bad, 0, 4, s));
    if (m > 0)
        hipLaunchKernelGGL(praq_dev_edges, dim3(blocks_for(m, kBuildBlock)), blk, 0, s, in.rows,
                           in.cols, m, in.n, b.e0, b.e1, count, bad);
    double* pd = D(o_part);
    double* pw = pd + chunks_d;
    hipLaunchKernelGGL(praq_dev_sumsq_chunks, dim3(blocks_for(chunks_d, kBuildBlock)), blk, 0, s,
                       in.diag, int64_t(n), 0, pd);
    if (m > 0)
        hipLaunchKernelGGL(praq_dev_sumsq_chunks, dim3(blocks_for(chunks_w, kBuildBlock)), blk, 0,
                           s, in.vals, int64_t(m), 1, pw);
    hipLaunchKernelGGL(praq_dev_squares, dim3(1), dim3(1), 0, s, pd, chunks_d, pw,
                       m > 0 ? chunks_w : 0, scale);
    REFORM_HIP_CHECK(hipGetLastError());
    int bad_host = 0;
    double squares = 0.0;
    REFORM_HIP_CHECK(hipMemcpyAsync(&bad_host, bad, 4, hipMemcpyDeviceToHost, s));
    REFORM_HIP_CHECK(hipMemcpyAsync(&squares, scale, 8, hipMemcpyDeviceToHost, s));
    REFORM_HIP_CHECK(hipStreamSynchronize(s));
    if (bad_host) return kErrorInstance;
    const double scale_host = squares > 0.0 ? std::sqrt(squares) : 1.0;
    REFORM_HIP_CHECK(hipMemcpyAsync(scale, &scale_host, 8, hipMemcpyHostToDevice, s));
    // ptr: the degrees scanned
    if (exclusive_scan(count, int64_t(n) + 1, I(o_scan), s)) return kErrorDevice;
    // half: the edge-ordered half-edges sorted stably by node
    if (m > 0) {
        int *k_in = I(o_keys), *v_in = b.half, *k_out = I(o_keys2), *v_out = I(o_vals2);
        hipLaunchKernelGGL(praq_dev_halves, dim3(blocks_for(m, kBuildBlock)), blk, 0, s, b.e0,
                           b.e1, m, k_in, v_in);
        int bits = 0;
        while (bits < 31 && (int64_t(1) << bits) < int64_t(n)) ++bits;
        const int passes = (bits + 7) / 8;
        for (int p = 0; p < passes; ++p) {
            int* hist = I(o_hist);
            hipLaunchKernelGGL(radix_count, dim3(tiles), blk, 0, s, k_in, halves, 8 * p, tiles,
                               hist);
            if (exclusive_scan(hist, int64_t(kRadixBins) * tiles, I(o_scan), s))
                return kErrorDevice;
            hipLaunchKernelGGL(radix_place, dim3(tiles), blk, 0, s, k_in, v_in, halves, 8 * p,
                               tiles, hist, k_out, v_out);
            std::swap(k_in, k_out);
            std::swap(v_in, v_out);
        }
        if (v_in != b.half)
            REFORM_HIP_CHECK(hipMemcpyAsync(b.half, v_in, 4 * h2, hipMemcpyDeviceToDevice, s));
    }
    // degrees, scaled diagonal, start, hubs
    const int heavy_degree = praq_heavy_degree(lanes);
    int* hub = I(o_hub);
    int* chunks = I(o_chunks);
    hipLaunchKernelGGL(praq_dev_nodes, dim3(blocks_for(int64_t(n) + 1, kBuildBlock)), blk, 0, s,
                       b.ptr, in.diag, n, dense_degree, heavy_degree, scale, b.diag, b.deg, b.xs,
                       hub, chunks);
    if (exclusive_scan(hub, int64_t(n) + 1, I(o_scan), s) ||
        exclusive_scan(chunks, int64_t(n) + 1, I(o_scan), s))
        return kErrorDevice;
    REFORM_HIP_CHECK(hipMemcpyAsync(totals, hub + n, 4, hipMemcpyDeviceToDevice, s));
    REFORM_HIP_CHECK(hipMemcpyAsync(totals + 1, chunks + n, 4, hipMemcpyDeviceToDevice, s));
    int counts[2] = {0, 0};
    REFORM_HIP_CHECK(hipMemcpyAsync(counts, totals, 8, hipMemcpyDeviceToHost, s));
    if (m > 0)
        hipLaunchKernelGGL(praq_dev_edge_data, dim3(blocks_for(m, kBuildBlock)), blk, 0, s,
                           in.vals, m, scale, b.w, b.ys);
    REFORM_HIP_CHECK(hipGetLastError());
    REFORM_HIP_CHECK(hipStreamSynchronize(s));
    b.nh = counts[0];
    b.nc = counts[1];
    // the hub lists, in a second arena now that their sizes are known
    const size_t hub_bytes = align256(4 * size_t(b.nh)) + align256(4 * (size_t(b.nh) + 1)) +
                             align256(4 * size_t(b.nc) + 4);
    REFORM_HIP_CHECK(device_malloc_async(&b.hub_arena, hub_bytes, s));
    char* hb = static_cast<char*>(b.hub_arena);
    b.heavy = reinterpret_cast<int*>(hb);
    b.hchunk = reinterpret_cast<int*>(hb + align256(4 * size_t(b.nh)));
    b.chunk_node = reinterpret_cast<int*>(hb + align256(4 * size_t(b.nh)) +
                                          align256(4 * (size_t(b.nh) + 1)));
    hipLaunchKernelGGL(praq_dev_hubs, dim3(blocks_for(n, kBuildBlock)), blk, 0, s, b.ptr, n,
                       heavy_degree, hub, chunks, b.heavy, b.hchunk, b.chunk_node);
    REFORM_HIP_CHECK(hipMemcpyAsync(b.hchunk + b.nh, chunks + n, 4, hipMemcpyDeviceToDevice, s));
    REFORM_HIP_CHECK(hipGetLastError());
    // the shape the layout and the binding read
    shape = PraqHostInstance();
    shape.n = n;
    shape.m = m;
    shape.scale = scale_host;
    shape.lanes = lanes;
    shape.heavy_degree = heavy_degree;
    shape.heavy.assign(size_t(b.nh), 0);
    shape.hchunk.assign(size_t(b.nh) + 1, 0);
    shape.chunk_node.assign(size_t(b.nc), 0);
    b.n = n;
    b.m = m;
    return 0;
}

template <class Real>
int praq_fill_image_device(const PraqDeviceBuild& b, const PraqLayout& L, char* image,
                           hipStream_t s) {
    auto put_int = [&](int k, const int* src, size_t count) -> int {
        if (count)
            REFORM_HIP_CHECK(hipMemcpyAsync(image + L.off[k], src, 4 * count,
                                            hipMemcpyDeviceToDevice, s));
        return 0;
    };
    auto put_real = [&](int k, const double* src, size_t count) -> int {
        if (!count) return 0;
        Real* dst = reinterpret_cast<Real*>(image + L.off[k]);
        if constexpr (sizeof(Real) == 8) {
            REFORM_HIP_CHECK(hipMemcpyAsync(dst, src, 8 * count, hipMemcpyDeviceToDevice, s));
        } else {
            const int grid = int(std::min<size_t>((count + kBuildBlock - 1) / kBuildBlock, 4096));
            hipLaunchKernelGGL(narrow<Real>, dim3(grid), dim3(kBuildBlock), 0, s, src, count, dst);
            REFORM_HIP_CHECK(hipGetLastError());
        }
        return 0;
    };
    const size_t n = size_t(b.n), m = size_t(b.m);
    if (put_int(PraqLayout::kE0, b.e0, m) || put_int(PraqLayout::kE1, b.e1, m) ||
        put_int(PraqLayout::kPtr, b.ptr, n + 1) || put_int(PraqLayout::kHalf, b.half, 2 * m) ||
        put_int(PraqLayout::kHeavy, b.heavy, size_t(b.nh)) ||
        put_int(PraqLayout::kHChunk, b.hchunk, size_t(b.nh) + 1) ||
        put_int(PraqLayout::kChunkNode, b.chunk_node, size_t(b.nc)) ||
        put_real(PraqLayout::kW, b.w, m) || put_real(PraqLayout::kDiag, b.diag, n) ||
        put_real(PraqLayout::kDeg, b.deg, n) || put_real(PraqLayout::kXS, b.xs, n) ||
        put_real(PraqLayout::kYS, b.ys, 2 * m))
        return kErrorDevice;
    return 0;
}

template int praq_fill_image_device<float>(const PraqDeviceBuild&, const PraqLayout&, char*,
                                           hipStream_t);
template int praq_fill_image_device<double>(const PraqDeviceBuild&, const PraqLayout&, char*,
                                            hipStream_t);

}  // namespace reform
