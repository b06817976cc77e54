// Shared device helpers for the gfx950 PDHG kernels: wave64 shuffles, fixed-
// order (deterministic) FP64 block reductions, and the Halpern step scalars.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>

namespace reform {

constexpr int kWave = 64;   // CDNA wavefront width

// A pointer into global memory (address space 1).  Array pointers that a
// kernel reads out of memory (a job record, a shared copy of one) are generic
// to the compiler, which then emits FLAT accesses: they also count against
// the LDS counter and, since they may alias the record and LDS, the compiler
// reloads the record after every store and waits for each access before the
// next.  Device code works on views whose pointers carry this type, so every
// access is a global_load/global_store with loads kept in flight.
template <class T>
using gptr = __attribute__((address_space(1))) T*;

template <class T>
__device__ __forceinline__ gptr<T> to_global(T* p) {
    return (gptr<T>)p;
}


template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, kWave);
        v = o > v ? o : v;
    }
    return v;
}

template <class T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, kWave);
        v = o < v ? o : v;
    }
    return v;
}

// Cross-lane move by DPP within rows of 16 lanes (gfx9 controls), 64-bit.
template <int Ctrl>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    return __hiloint2double(__builtin_amdgcn_update_dpp(hi, hi, Ctrl, 0xf, 0xf, false),
                            __builtin_amdgcn_update_dpp(lo, lo, Ctrl, 0xf, 0xf, false));
}

// Sum over groups of `width` consecutive lanes (4, 8, 16, 32 or 64; uniform
// over the wave): xor 1 and xor 2 by quad permutes, then the other quad of
// the half-row and the other half of the row by mirrors -- each a pairwise
// exchange, so both partners add the same two numbers and every lane of a
// group ends with the same bits -- then xor 16 and xor 32 by shuffles.
__device__ __forceinline__ double group_sum_dpp(double v, int width) {
    v += dpp_move<0xb1>(v);                      // quad_perm [1, 0, 3, 2]
    v += dpp_move<0x4e>(v);                      // quad_perm [2, 3, 0, 1]
    if (width > 4) v += dpp_move<0x141>(v);      // row_half_mirror: quad q <-> q ^ 1
    if (width > 8) v += dpp_move<0x140>(v);      // row_mirror: half h <-> h ^ 1
    if (width > 16) v += __shfl_xor(v, 16, kWave);
    if (width > 32) v += __shfl_xor(v, 32, kWave);
    return v;
}

// Sum over a wave64: every lane ends with the same bits, and the order of
// additions depends only on the lane layout.
__device__ __forceinline__ double wave_sum(double v) { return group_sum_dpp(v, kWave); }

// Sum of `width` consecutive lanes (a power of two <= 64), xor butterfly.
template <class T>
__device__ __forceinline__ T group_sum(T v, int width) {
    for (int off = width >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// K FP64 sums over the whole workgroup.  `scratch` is LDS of at least
// K * (blockDim.x / 64) doubles.  Every thread returns with the totals in
// v[]; the order of additions depends only on the block shape.
template <int K>
__device__ __forceinline__ void block_sums(double (&v)[K], double* scratch) {
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int waves = blockDim.x / kWave;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        v[k] = wave_sum(v[k]);
        if (lane == 0) scratch[k * 16 + wave] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double t = 0.0;
        for (int w = 0; w < waves; ++w) t += scratch[k * 16 + w];
        v[k] = t;
    }
    __syncthreads();
}

// block_sums (K <= 4) with one barrier instead of two: successive calls
// alternate between the two halves of `scratch` (2 x 64 doubles, whatever
// each call's K), so a call's writes cannot meet the previous call's reads
// (every wave has passed the barrier in between).  `parity` is the caller's
// count of calls, the same on every thread; the sums equal block_sums' bit
// for bit.
template <int K>
__device__ __forceinline__ void block_sums_alt(double (&v)[K], double* scratch, int& parity) {
    static_assert(K <= 4, "a half holds 4 x 16 doubles");
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int waves = blockDim.x / kWave;
    double* half = scratch + (parity & 1) * 4 * 16;
    ++parity;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        v[k] = wave_sum(v[k]);
        if (lane == 0) half[k * 16 + wave] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double t = 0.0;
        for (int w = 0; w < waves; ++w) t += half[k * 16 + w];
        v[k] = t;
    }
}

// One FP64 value: sum / max / min over the workgroup (scratch: 16 doubles).
__device__ __forceinline__ double block_sum1(double v, double* scratch) {
    double a[1] = {v};
    block_sums<1>(a, scratch);
    return a[0];
}

template <class T>
__device__ __forceinline__ T block_max1(T v, T* scratch) {
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    v = wave_max(v);
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    T t = scratch[0];
    for (int w = 1; w < (int)(blockDim.x / kWave); ++w) t = scratch[w] > t ? scratch[w] : t;
    __syncthreads();
    return t;
}

template <class T>
__device__ __forceinline__ T block_min1(T v, T* scratch) {
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    v = wave_min(v);
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    T t = scratch[0];
    for (int w = 1; w < (int)(blockDim.x / kWave); ++w) t = scratch[w] < t ? scratch[w] : t;
    __syncthreads();
    return t;
}

template <class Real>
__device__ __forceinline__ Real clampr(Real v, Real lo, Real hi) {
    return v < lo ? lo : (v > hi ? hi : v);
}

// Halpern weights of a reflected step with restart counter r:
// k = r + 1, left = k / (k + 1), right = 1 / (k + 1)
// (qubo/praq-jax/praq/solver.py:276-292, jrp/pria-jax/pria/solver.py:428-447).
__host__ __device__ inline void halpern_weights(int64_t r, double* lw, double* rw) {
    double k = double(r + 1);
    *lw = k / (k + 1.0);
    *rw = 1.0 / (k + 1.0);
}

// Every device allocation of the library goes through device_malloc or
// device_malloc_async, so that a test can make one of them fail as out of
// memory (include/reform_debug.h); device_alloc_fault() counts the
// allocation and says whether it is the one to fail.
bool device_alloc_fault();

template <class T>
inline hipError_t device_malloc(T** p, size_t bytes) {
    if (device_alloc_fault()) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc(reinterpret_cast<void**>(p), bytes);
}

template <class T>
inline hipError_t device_malloc_async(T** p, size_t bytes, hipStream_t stream) {
    if (device_alloc_fault()) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMallocAsync(reinterpret_cast<void**>(p), bytes, stream);
}

// The last failed HIP call of the library (call text, place, HIP error),
// readable through reform_last_error().
inline std::mutex& last_error_mutex() {
    static std::mutex m;
    return m;
}
inline std::string& last_error_text() {
    static std::string s;
    return s;
}
inline void record_error(const char* call, hipError_t err, const char* file, int line) {
    std::lock_guard<std::mutex> lock(last_error_mutex());
    last_error_text() = std::string(call) + " at " + file + ":" + std::to_string(line) + ": " +
                        hipGetErrorString(err);
}
// A failure that is no HIP call's (an argument an entry point rejects with
// an explanation), readable through reform_last_error().
inline void record_message(const std::string& text) {
    std::lock_guard<std::mutex> lock(last_error_mutex());
    last_error_text() = text;
}

}  // namespace reform

#define REFORM_HIP_CHECK(call)                                                \
    do {                                                                      \
        hipError_t reform_err_ = (call);                                      \
        if (reform_err_ != hipSuccess) {                                      \
            ::reform::record_error(#call, reform_err_, __FILE__, __LINE__);   \
            return kErrorDevice;                                              \
        }                                                                     \
    } while (0)
