// Scans and radix sorts of the graph-build path (transpose, degree order,
// generator cleanup, propagation-blocking layout), written for gfx950.
//
// Scan (uint32, sum or max, exclusive or inclusive): three launches.  Each
// workgroup of kScanBlock lanes owns a tile of kScanBlock * kScanItems
// consecutive items: it reduces its tile, the tile totals are scanned (by the
// same routine, recursively, until one tile holds them all), and every tile
// scans itself on top of its offset.  Lane t of a tile holds items
// t * kScanItems .. + kScanItems - 1, so the per-lane pass is sequential and
// the cross-lane pass is one wavefront scan (shuffles) plus a scan of the
// wave totals in LDS.
//
// Radix sort (uint64 keys over bits [0, end_bit), optional uint32 values;
// least significant digit first, 8-bit digits, stable): per digit a histogram
// launch writes every tile's 256 digit counts digit-major (counts[digit *
// tiles + tile]), whose exclusive scan is the first output slot of every
// (digit, tile) pair; a scatter launch then ranks the tile's keys stably
// (within a wavefront from 8 ballots of the digit bits, across wavefronts and
// strips of 256 keys from LDS counters), stages them digit-sorted in LDS and
// writes them out in that order, so consecutive lanes store consecutive
// slots of one digit's run.  Passes alternate between the output and a
// scratch buffer so that the last one writes the output.
#include <algorithm>

#include "gx_device.hpp"
#include "gx_internal.hpp"

namespace gx {
namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int64_t kScanTile = int64_t(kScanBlock) * kScanItems;

struct SumOp {
  __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; }
};
struct MaxOp {
  __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};

// Exclusive scan of v over the workgroup (kScanBlock lanes) starting at
// `init`; the workgroup's combined value (without init) in *total.
template <class Op>
__device__ __forceinline__ uint32_t block_scan_excl(uint32_t v, uint32_t init, uint32_t* lds, uint32_t* total,
                                                    Op op) {
  const unsigned lane = lane_id(), wave = threadIdx.x / kWave;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t o = __shfl_up(incl, off);
    if (lane >= static_cast<unsigned>(off)) incl = op(o, incl);
  }
  const uint32_t prev = __shfl_up(incl, 1);
  if (lane == kWave - 1) lds[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    // wave totals -> exclusive prefixes (lds[w]) and the total (lds[waves])
    uint32_t run = lds[0];
    lds[0] = init;
    for (int w = 1; w < kScanBlock / kWave; ++w) {
      const uint32_t t = lds[w];
      lds[w] = op(init, run);
      run = op(run, t);
    }
    lds[kScanBlock / kWave] = run;
  }
  __syncthreads();
  const uint32_t r = lane == 0 ? lds[wave] : op(lds[wave], prev);
  *total = lds[kScanBlock / kWave];
  __syncthreads();
  return r;
}

template <class Op>
__global__ void __launch_bounds__(kScanBlock) scan_reduce_kernel(const uint32_t* __restrict__ in, int64_t count,
                                                                 uint32_t* __restrict__ totals, Op op) {
  __shared__ uint32_t lds[kScanBlock / kWave + 1];
  const int64_t base = blockIdx.x * kScanTile + static_cast<int64_t>(threadIdx.x) * kScanItems;
  uint32_t acc = 0;
  bool any = false;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k)
    if (base + k < count) {
      acc = any ? op(acc, in[base + k]) : in[base + k];
      any = true;
    }
  // lanes without items contribute the identity of sum and of max over
  // unsigned values (0)
  uint32_t total;
  (void)block_scan_excl(any ? acc : 0u, 0u, lds, &total, op);
  if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

// out = scan of in over the tile, starting at offset[tile] (null: 0).  in and
// out may alias: every item is read before the tile writes.
template <class Op, bool EXCL>
__global__ void __launch_bounds__(kScanBlock) scan_tile_kernel(const uint32_t* in, int64_t count,
                                                               const uint32_t* __restrict__ offset, uint32_t* out,
                                                               Op op) {
  __shared__ uint32_t lds[kScanBlock / kWave + 1];
  const int64_t base = blockIdx.x * kScanTile + static_cast<int64_t>(threadIdx.x) * kScanItems;
  uint32_t v[kScanItems];
  uint32_t acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = base + k < count ? in[base + k] : 0u;
    acc = k == 0 ? v[k] : op(acc, v[k]);
  }
  uint32_t total;
  uint32_t r = block_scan_excl(acc, offset ? offset[blockIdx.x] : 0u, lds, &total, op);
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (EXCL) {
      if (base + k < count) out[base + k] = r;
      r = op(r, v[k]);
    } else {
      r = op(r, v[k]);
      if (base + k < count) out[base + k] = r;
    }
  }
}

// Scan of `count` items of in into out (may alias), exclusive or inclusive,
// starting at 0.  Tile totals are scanned by the same routine.
template <class Op>
void scan_u32(const uint32_t* in, uint32_t* out, int64_t count, bool exclusive, Op op, hipStream_t s) {
  if (count <= 0) return;
  const int64_t tiles = (count + kScanTile - 1) / kScanTile;
  GX_REQUIRE(tiles < (int64_t(1) << 31), "scan too large");
  if (tiles == 1) {
    if (exclusive)
      scan_tile_kernel<Op, true><<<1, kScanBlock, 0, s>>>(in, count, nullptr, out, op);
    else
      scan_tile_kernel<Op, false><<<1, kScanBlock, 0, s>>>(in, count, nullptr, out, op);
    GX_LAUNCHED("scan_tile_kernel");
    return;
  }
  DevBuf tot;
  tot.ensure(tiles * 4);
  scan_reduce_kernel<Op><<<static_cast<unsigned>(tiles), kScanBlock, 0, s>>>(in, count, tot.as<uint32_t>(), op);
  GX_LAUNCHED("scan_reduce_kernel");
  scan_u32(tot.as<uint32_t>(), tot.as<uint32_t>(), tiles, true, op, s);
  if (exclusive)
    scan_tile_kernel<Op, true><<<static_cast<unsigned>(tiles), kScanBlock, 0, s>>>(in, count, tot.as<uint32_t>(),
                                                                                   out, op);
  else
    scan_tile_kernel<Op, false><<<static_cast<unsigned>(tiles), kScanBlock, 0, s>>>(in, count, tot.as<uint32_t>(),
                                                                                    out, op);
  GX_LAUNCHED("scan_tile_kernel");
  // `tot` is freed when this returns: wait for the launches that read it
  GX_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------ radix sort
constexpr int kSortBlock = 256;
constexpr int kSortItems = 16;
constexpr int64_t kSortTile = int64_t(kSortBlock) * kSortItems;
constexpr int kDigits = 256;
constexpr int kSortWaves = kSortBlock / kWave;

__device__ __forceinline__ uint32_t digit_of(uint64_t key, int shift, uint32_t mask) {
  return static_cast<uint32_t>(key >> shift) & mask;
}

__global__ void __launch_bounds__(kSortBlock) radix_hist_kernel(const uint64_t* __restrict__ keys, int64_t count,
                                                                int shift, uint32_t mask,
                                                                uint32_t* __restrict__ counts, int64_t tiles) {
  __shared__ uint32_t h[kDigits];
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = blockIdx.x * kSortTile;
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const int64_t i = base + k * kSortBlock + threadIdx.x;
    if (i < count) atomicAdd(&h[digit_of(keys[i], shift, mask)], 1u);
  }
  __syncthreads();
  counts[static_cast<int64_t>(threadIdx.x) * tiles + blockIdx.x] = h[threadIdx.x];
}

// One tile: stable ranks by digit, staged digit-sorted in LDS, written out at
// offsets[digit * tiles + tile] + position within the tile's run of the digit.
template <bool VALS>
__global__ void __launch_bounds__(kSortBlock) radix_scatter_kernel(const uint64_t* __restrict__ kin,
                                                                   const uint32_t* __restrict__ vin, int64_t count,
                                                                   int shift, uint32_t mask,
                                                                   const uint32_t* __restrict__ offsets,
                                                                   int64_t tiles, uint64_t* __restrict__ kout,
                                                                   uint32_t* __restrict__ vout) {
  __shared__ uint64_t skeys[kSortTile];
  __shared__ uint32_t svals[VALS ? kSortTile : 1];
  __shared__ uint32_t cnt[kDigits];                // keys of each digit in earlier strips
  __shared__ uint32_t wcnt[kSortWaves][kDigits];   // keys of each digit per wavefront in this strip
  __shared__ uint32_t dstart[kDigits], gbase[kDigits];
  __shared__ uint32_t scan_lds[kScanBlock / kWave + 1];
  const unsigned tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const int64_t base = blockIdx.x * kSortTile;
  uint64_t key[kSortItems];
  uint32_t val[kSortItems], rank[kSortItems], dig[kSortItems];
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const int64_t i = base + k * kSortBlock + tid;
    key[k] = i < count ? kin[i] : 0ull;
    if constexpr (VALS) val[k] = i < count ? vin[i] : 0u;
  }
  cnt[tid] = 0;
  const uint64_t below = (uint64_t(1) << lane) - 1;
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const bool valid = base + k * kSortBlock + tid < count;
    const uint32_t d = digit_of(key[k], shift, mask);
    dig[k] = d;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    uint64_t match = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      match &= bit ? bal : ~bal;
    }
    const uint32_t r = static_cast<uint32_t>(__popcll(match & below));
    if (valid && r == 0) wcnt[wave][d] = static_cast<uint32_t>(__popcll(match));
    __syncthreads();
    if (valid) {
      uint32_t before = cnt[d];
      for (unsigned w = 0; w < wave; ++w) before += wcnt[w][d];
      rank[k] = before + r;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < kSortWaves; ++w) add += wcnt[w][tid];
    cnt[tid] += add;
    __syncthreads();
  }
  // cnt holds the tile's digit histogram: runs start at its exclusive scan.
  uint32_t total;
  const uint32_t ds = block_scan_excl(cnt[tid], 0u, scan_lds, &total, SumOp{});
  dstart[tid] = ds;
  gbase[tid] = offsets[static_cast<int64_t>(tid) * tiles + blockIdx.x];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    if (base + k * kSortBlock + tid < count) {
      const uint32_t pos = dstart[dig[k]] + rank[k];
      skeys[pos] = key[k];
      if constexpr (VALS) svals[pos] = val[k];
    }
  }
  __syncthreads();
  const int64_t here = count - base < kSortTile ? count - base : kSortTile;
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    const int64_t i = k * kSortBlock + tid;
    if (i < here) {
      const uint64_t kk = skeys[i];
      const uint32_t d = digit_of(kk, shift, mask);
      const int64_t pos = static_cast<int64_t>(gbase[d]) + (i - dstart[d]);
      kout[pos] = kk;
      if constexpr (VALS) vout[pos] = svals[i];
    }
  }
}

}  // namespace

void scan_u32_sum(const uint32_t* in, uint32_t* out, int64_t count, hipStream_t s) {
  scan_u32(in, out, count, true, SumOp{}, s);
}

void inclusive_max_scan_u32(const uint32_t* in, uint32_t* out, int64_t count, hipStream_t s) {
  scan_u32(in, out, count, false, MaxOp{}, s);
}

void radix_sort_u64(const uint64_t* kin, uint64_t* kout, const uint32_t* vin, uint32_t* vout, int64_t count,
                    int end_bit, hipStream_t s) {
  GX_REQUIRE(count >= 0 && count < (int64_t(1) << 32), "sort: too many keys");
  GX_REQUIRE(end_bit >= 0 && end_bit <= 64, "sort: bad bit range");
  GX_REQUIRE(kin != kout && (!vin || vin != vout), "sort: input and output must differ");
  const bool vals = vin != nullptr;
  GX_REQUIRE(!vals || vout != nullptr, "sort: values without an output");
  if (count == 0) return;
  const int passes = (end_bit + 7) / 8;
  if (passes == 0) {
    GX_HIP(hipMemcpyAsync(kout, kin, count * 8, hipMemcpyDeviceToDevice, s));
    if (vals) GX_HIP(hipMemcpyAsync(vout, vin, count * 4, hipMemcpyDeviceToDevice, s));
    GX_HIP(hipStreamSynchronize(s));
    return;
  }
  const int64_t tiles = (count + kSortTile - 1) / kSortTile;
  GX_REQUIRE(tiles * kDigits < (int64_t(1) << 32), "sort: too many tiles");
  DevBuf tk, tv, counts;
  tk.ensure(count * 8);
  if (vals) tv.ensure(count * 4);
  counts.ensure(tiles * kDigits * 4);
  const uint64_t* sk = kin;
  const uint32_t* sv = vin;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    const int bits = std::min(8, end_bit - shift);
    const uint32_t mask = (1u << bits) - 1u;
    const bool to_out = (passes - p) % 2 == 1;  // the last pass writes the output
    uint64_t* dk = to_out ? kout : tk.as<uint64_t>();
    uint32_t* dv = vals ? (to_out ? vout : tv.as<uint32_t>()) : nullptr;
    radix_hist_kernel<<<static_cast<unsigned>(tiles), kSortBlock, 0, s>>>(sk, count, shift, mask,
                                                                          counts.as<uint32_t>(), tiles);
    GX_LAUNCHED("radix_hist_kernel");
    scan_u32(counts.as<uint32_t>(), counts.as<uint32_t>(), tiles * kDigits, true, SumOp{}, s);
    if (vals)
      radix_scatter_kernel<true><<<static_cast<unsigned>(tiles), kSortBlock, 0, s>>>(
          sk, sv, count, shift, mask, counts.as<uint32_t>(), tiles, dk, dv);
    else
      radix_scatter_kernel<false><<<static_cast<unsigned>(tiles), kSortBlock, 0, s>>>(
          sk, nullptr, count, shift, mask, counts.as<uint32_t>(), tiles, dk, nullptr);
    GX_LAUNCHED("radix_scatter_kernel");
    sk = dk;
    sv = dv;
  }
  GX_HIP(hipStreamSynchronize(s));  // scratch buffers are freed on return
}

}  // namespace gx
