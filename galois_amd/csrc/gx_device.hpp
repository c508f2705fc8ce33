// Device-side building blocks for gfx950 (CDNA4, wave64): lane/wave helpers,
// wave-aggregated queue appends, wave and workgroup reductions/scans, the
// counter-based hash used by the generators, and fill kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gx {

constexpr int kWave = 64;

__device__ __forceinline__ unsigned lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of `mask` in lanes below the calling lane.
__device__ __forceinline__ unsigned lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}

// Append `item` to queue[*count] for every lane with `pred`; one atomic per
// wavefront.  Returns false (and appends nothing for this lane) when the slot
// would reach `cap`; the caller raises an overflow flag.
template <class T>
__device__ __forceinline__ bool wave_append(bool pred, T item, T* queue, unsigned* count,
                                            unsigned cap, unsigned* overflow) {
  uint64_t mask = __ballot(pred);
  if (mask == 0) return true;
  unsigned total = __popcll(mask);
  unsigned leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
  unsigned base = 0;
  if (lane_id() == leader) base = atomicAdd(count, total);
  base = __shfl(base, leader);
  if (pred) {
    unsigned slot = base + lanes_below(mask);
    if (slot < cap) {
      queue[slot] = item;
    } else {
      atomicOr(overflow, 1u);
      return false;
    }
  }
  return true;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off);
    v = o > v ? o : v;
  }
  return v;
}

template <class T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off);
    v = o < v ? o : v;
  }
  return v;
}

// Inclusive scan across the wavefront.
template <class T>
__device__ __forceinline__ T wave_inclusive_scan(T v) {
  unsigned lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    T o = __shfl_up(v, off);
    if (lane >= static_cast<unsigned>(off)) v += o;
  }
  return v;
}

// Workgroup exclusive scan of one value per thread (BLOCK a multiple of 64).
// `lds` must hold BLOCK/64 + 1 elements.  Returns the exclusive prefix and
// writes the block total to *total (all threads).
template <int BLOCK, class T>
__device__ __forceinline__ T block_exclusive_scan(T v, T* lds, T* total) {
  constexpr int NW = BLOCK / kWave;
  unsigned lane = lane_id();
  unsigned wid = threadIdx.x / kWave;
  T inc = wave_inclusive_scan(v);
  if (lane == kWave - 1) lds[wid] = inc;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
      T t = lds[i];
      lds[i] = run;
      run += t;
    }
    lds[NW] = run;
  }
  __syncthreads();
  T res = lds[wid] + inc - v;
  *total = lds[NW];
  __syncthreads();
  return res;
}

// Error-free transformation of a + b (Knuth TwoSum): s + e == a + b exactly.
__device__ __forceinline__ void two_sum(float a, float b, float& s, float& e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Compensated (Kahan-Babuska / Neumaier) accumulator in FP32.
struct KahanF32 {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) {
    float t = s + x;
    if (fabsf(s) >= fabsf(x))
      c += (s - t) + x;
    else
      c += (x - t) + s;
    s = t;
  }
  // Merge another accumulator.
  __device__ __forceinline__ void merge(float os, float oc) {
    add(os);
    c += oc;
  }
  __device__ __forceinline__ float value() const { return s + c; }
};

// ---------------------------------------------------------- hashing (PRNG)
// splitmix64 finaliser; the generators draw every random number as
// mix64(seed ^ f(counter)) so results depend only on (seed, counter).
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__host__ __device__ __forceinline__ uint64_t hash2(uint64_t seed, uint64_t a, uint64_t b) {
  return mix64(mix64(seed ^ mix64(a)) ^ (b * 0xD1B54A32D192ED03ull));
}

// ---------------------------------------------------------------- fills
template <class T>
__global__ void fill_kernel(T* p, T v, int64_t n) {
  int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

}  // namespace gx
