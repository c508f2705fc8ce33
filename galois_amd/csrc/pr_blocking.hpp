// Propagation-blocked PageRank sweep (gfx950).
//
// The pull sweep of rowsweep.hpp gathers y[u] for every in-edge: ~1e9
// 4-byte random reads per RMAT-26 iteration, each of which moves a whole
// cache line, so the sweep is bound by wasted line traffic (about ten times
// the algorithmic bytes).  Propagation blocking (Beamer, Asanovic, Patterson,
// "Reducing PageRank Communication via Propagation Blocking", IPDPS 2017)
// replaces the random reads with two streaming phases over a static edge
// layout built once per graph:
//
//   layout  edges sorted by (source chunk, destination bin, source,
//           destination), kSpan vertices per chunk and per bin.  A run is
//           the edges of one (chunk, bin) pair; a group is the edges of one
//           source inside a run, which all carry the same contribution.  Per
//           edge: the in-bin destination (dst16, bit 15 set on the first
//           edge of a group; stored bin-major, so that a bin's runs from
//           every chunk form one stretch, ordered by the length class of
//           their first job); per group, in run-major order:
//           the in-chunk source (gsrc16) and the contribution slot
//           vals[group].
//   phase 1 (gather_kernel) one workgroup per piece of a chunk: the chunk's
//           kSpan contributions y = x / outdeg are computed from x and the
//           out-degrees and staged in LDS (FP32: encoded as fixed-point
//           words, see encode_fixed), and vals[q] = y[gsrc16[q]] is streamed
//           out in layout order (2-byte read, 4-byte write per group).
//   phase 2 (accum_kernel) one workgroup per task (a bin, or a slice of a
//           heavy bin): the bin's runs are read from every chunk, the group
//           of each edge is counted from the group flags with wavefront
//           ballots, and vals[group] is added into an LDS accumulator per
//           destination (2 bytes per edge and 4 per group read), then the
//           bin's rows are finalised (x' only: the vertices without
//           out-edges, whose mass is redistributed, are the tail of the
//           degree order, so no out-degree is read there).
//
// On RMAT graphs a hub source has many edges in one bin, so there are far
// fewer groups than edges (about 0.33 groups per edge at RMAT-26): per edge
// the sweep moves 2 + 10 * groups / edges bytes instead of 12.
//
// FP32 row sums are exact: every contribution (0 <= y <= 1 for the
// normalised iteration) is rounded to a 64-bit fixed-point number with 2^-62
// resolution and added with integer atomics, so the sum does not depend on
// the order of the additions and the result is deterministic.  The FP64
// variant streams FP64 contributions and adds them with FP64 LDS atomics
// (2^-62 fixed point would be coarser than FP64 for small contributions);
// its sums depend on the addition order in the last bit.  Slices of a heavy bin add
// their LDS sums into a per-bin global accumulator that multi_kernel
// finalises (and clears for the next iteration).
//
// Phase-2 work is described by jobs (start, length, group base) of at most
// kJob edges of one run, in three size classes, so that runs of every length
// keep the lanes busy: long jobs (65..512 edges) take a wavefront each, mid
// jobs (9..64) eight lanes each, short jobs (1..8) one lane each.
#pragma once

#include <cstdint>

#include "gx_device.hpp"
#include "gx_internal.hpp"
#include "rowsweep.hpp"

namespace gx {
namespace pb {

constexpr int kShift = 14;           // log2 of the vertices per chunk and per bin
constexpr int kSpan = 1 << kShift;   // 16K: 64 KB of FP32 y, 128 KB of 64-bit sums
constexpr int kBlock = 1024;
constexpr int kWaves = kBlock / kWave;
constexpr int kJob = 512;            // entries per long job: 8 per lane
constexpr int kMid = 64;             // mid jobs: 8 lanes x 8 entries
constexpr int kShort = 8;            // short jobs: 1 lane x 8 entries
constexpr int64_t kPieceEntries = int64_t(1) << 18;  // phase-1 piece length (groups)
constexpr int64_t kMaxKeys = int64_t(1) << 28;       // chunk x bin table limit
constexpr uint32_t kSingle = 0xffffffffu;            // task finalises its bin itself
constexpr uint32_t kFirst = 1u << 15;                // dst16 flag: first edge of a group

struct Piece {
  uint32_t chunk, p0, p1, pad;
};
// Edges [start, start + len) of one run, positions in the bin-major dst16.
// The group of its i-th edge is gbase + (number of group flags among its
// edges 0 .. i), modulo 2^32: gbase is the number of groups that begin
// before the job's first edge in run-major order, minus one.
struct Job {
  uint32_t start, len, gbase;
};
struct Task {
  uint32_t bin, slot;  // slot: global accumulator of a sliced bin, or kSingle
  uint32_t job0[3], job1[3];  // [begin, end) in the long / mid / short job lists
};
struct Multi {
  uint32_t bin, slot;
};

// Whether the layout can be built for this graph (nonempty, table size).
inline bool supported(int64_t n, int64_t m) {
  const int64_t nb = (n + kSpan - 1) / kSpan;
  return n > 0 && m > 0 && m < (int64_t(1) << 31) && nb * nb <= kMaxKeys;
}

// y value -> fixed point with 2^-62 resolution, rounded to nearest (ties
// up).  Valid for 0 <= v < 2; the normalised iteration keeps every x (and so
// every y and every row sum) within [0, 1].
__host__ __device__ __forceinline__ unsigned long long to_fixed(float v) {
  const uint32_t b = __builtin_bit_cast(uint32_t, v);
  const int e = static_cast<int>(b >> 23);  // biased exponent (v >= 0)
  if (e == 0) return 0;                     // zero / subnormal: below 2^-63
  const unsigned long long mant = (b & 0x7fffffu) | 0x800000u;
  const int sh = e - 88;                    // v * 2^62 = mant * 2^(e - 127 - 23 + 62)
  if (sh >= 0) return mant << sh;
  if (sh <= -25) return 0;
  const int r = -sh;
  return (mant + (1ull << (r - 1))) >> r;
}

// to_fixed(v) in 32 bits, as m << s with m < 2^24 in bits 0..23 and s <= 39
// in bits 24..29: the mantissa with its shift when v >= 2^-39 (exact), else
// the rounded value itself (s = 0).  Phase 1 encodes once per group, so that
// phase 2 decodes every edge's contribution with one 64-bit shift instead of
// the branches of to_fixed.
__host__ __device__ __forceinline__ uint32_t encode_fixed(float v) {
  const uint32_t b = __builtin_bit_cast(uint32_t, v);
  const int e = static_cast<int>(b >> 23);
  if (e == 0) return 0;
  const uint32_t mant = (b & 0x7fffffu) | 0x800000u;
  const int sh = e - 88;
  if (sh >= 0) return mant | (static_cast<uint32_t>(sh) << 24);
  if (sh <= -25) return 0;
  const int r = -sh;
  return (mant + (1u << (r - 1))) >> r;  // < 2^24
}
__host__ __device__ __forceinline__ unsigned long long decode_fixed(uint32_t w) {
  return static_cast<unsigned long long>(w & 0xffffffu) << (w >> 24);
}

// Exact fixed-point sum -> unevaluated FP32 pair hi + lo (about 48 bits).
__device__ __forceinline__ void fixed_to_float2(unsigned long long s, float& hi, float& lo) {
  const float h = __ull2float_rn(s);
  const long long r = static_cast<long long>(s - __float2ull_rn(h));  // exact, may be negative
  hi = h * 0x1p-62f;
  lo = __ll2float_rn(r) * 0x1p-62f;
}

// Per precision: the accumulator type T, the type V of the per-group
// contributions phase 1 writes, and the conversions.  FP32 contributions are
// written as encode_fixed words and summed exactly as 64-bit fixed point
// (integer atomics, order independent); FP64 contributions are written and
// summed as FP64 (atomics, order dependent in the last bit).
template <class Real>
struct PbAcc;
template <>
struct PbAcc<float> {
  using T = unsigned long long;
  using V = uint32_t;
  static __device__ __forceinline__ V encode(float v) { return encode_fixed(v); }
  static __device__ __forceinline__ T conv(V w) { return decode_fixed(w); }
};
template <>
struct PbAcc<double> {
  using T = double;
  using V = double;
  static __device__ __forceinline__ V encode(double v) { return v; }
  static __device__ __forceinline__ T conv(V v) { return v; }
};

// Phase 1: vals[q] = encode(y[chunk base + gsrc16[q]]) over one piece of
// groups, with y = x / outdeg computed (and encoded) once per vertex as the
// chunk is staged: the finaliser then writes x' only.
template <class Real>
__global__ void __launch_bounds__(kBlock) gather_kernel(const Piece* __restrict__ pieces,
                                                        const uint16_t* __restrict__ src16,
                                                        const Real* __restrict__ x,
                                                        const uint32_t* __restrict__ outdeg, int64_t n,
                                                        typename PbAcc<Real>::V* __restrict__ vals,
                                                        const unsigned* stop) {
  using V = typename PbAcc<Real>::V;
  if (*stop) return;
  __shared__ V ys[kSpan];
  const Piece P = pieces[blockIdx.x];
  const int64_t base = static_cast<int64_t>(P.chunk) << kShift;
  const int cnt = static_cast<int>(n - base < kSpan ? n - base : kSpan);
  for (int i = threadIdx.x; i < cnt; i += kBlock) {
    const uint32_t d = rs::ld_stream(&outdeg[base + i]);
    const Real xv = rs::ld_stream(&x[base + i]);
    ys[i] = PbAcc<Real>::encode(d ? xv / static_cast<Real>(d) : Real(0));
  }
  __syncthreads();
  // Body in 4-entry groups: one 8-byte index load and one 16-byte store per
  // lane, contiguous across the wavefront; the unaligned head and tail of the
  // piece entry by entry.
  const int64_t a4 = (static_cast<int64_t>(P.p0) + 3) & ~int64_t(3);
  const int64_t b4 = static_cast<int64_t>(P.p1) & ~int64_t(3);
  if (a4 >= b4) {
    for (int64_t q = static_cast<int64_t>(P.p0) + threadIdx.x; q < P.p1; q += kBlock)
      __builtin_nontemporal_store(ys[src16[q]], &vals[q]);
    return;
  }
  if (threadIdx.x < a4 - P.p0) {
    const int64_t q = P.p0 + threadIdx.x;
    vals[q] = ys[src16[q]];
  }
  if (threadIdx.x < P.p1 - b4) {
    const int64_t q = b4 + threadIdx.x;
    vals[q] = ys[src16[q]];
  }
  typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
  typedef V r4 __attribute__((ext_vector_type(4)));
  const u32x2* s4 = reinterpret_cast<const u32x2*>(src16 + a4);
  r4* v4 = reinterpret_cast<r4*>(vals + a4);
  const int64_t groups = (b4 - a4) >> 2;
  constexpr int kU = 4;
  for (int64_t gi = threadIdx.x; gi < groups; gi += kU * kBlock) {
    u32x2 w[kU];
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int64_t q = gi + k * kBlock;
      if (q < groups) w[k] = __builtin_nontemporal_load(&s4[q]);
    }
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const int64_t q = gi + k * kBlock;
      if (q < groups) {
        r4 o;
        o.x = ys[w[k].x & 0xffffu];
        o.y = ys[w[k].x >> 16];
        o.z = ys[w[k].y & 0xffffu];
        o.w = ys[w[k].y >> 16];
        __builtin_nontemporal_store(o, &v4[q]);
      }
    }
  }
}

// Adds the edges of one job class into the LDS sums.  Lanes are grouped
// LPJ to a job; lane `sub` of a group takes edges sub + LPJ*k, k < 8, so for
// each k the LPJ lanes of a job hold LPJ consecutive edges and a ballot of
// the group flags, masked to those lanes, counts the groups they open.  A
// wavefront works in rounds of kU job groups, software-pipelined: while a
// round loads its contributions and adds them, the next round's destination
// ids and the job entries of the round after that are already in flight, so
// each round waits for one load latency instead of three (job entry ->
// destination ids -> contribution).
template <int LPJ, class Real>
__device__ __forceinline__ void accum_jobs(const Job* __restrict__ jobs, uint32_t j0, uint32_t j1,
                                           const typename PbAcc<Real>::V* __restrict__ vals,
                                           const uint16_t* __restrict__ dst16, typename PbAcc<Real>::T* acc) {
  using V = typename PbAcc<Real>::V;
  constexpr int kGroups = kWave / LPJ;
  constexpr int kU = sizeof(Real) == 4 ? 2 : 1;
  constexpr uint32_t kStride = kWaves * kGroups * kU;
  const unsigned lane = lane_id(), wave = threadIdx.x / kWave;
  const unsigned grp = lane / LPJ, sub = lane % LPJ;
  const uint64_t mine = LPJ == kWave ? ~0ull : ((uint64_t(1) << LPJ) - 1) << (grp * LPJ);
  const uint64_t upto = mine & ((uint64_t(2) << lane) - 1);  // my job's lanes up to mine (lane 63: all)
  auto load_jobs = [&](uint32_t jb, Job (&J)[kU]) {
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      const uint32_t j = jb + u * kGroups + grp;
      J[u] = Job{0, 0, 0};
      if (j < j1) J[u] = jobs[j];
    }
  };
  auto load_dst = [&](const Job (&J)[kU], uint32_t (&d)[kU][8]) {
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint32_t i = sub + LPJ * k;
        d[u][k] = i < J[u].len ? rs::ld_stream(&dst16[J[u].start + i]) : 0u;
      }
  };
  uint32_t jb = j0 + wave * kGroups * kU;
  if (jb >= j1) return;  // wavefront-uniform
  Job J[kU], Jn[kU];
  uint32_t d[kU][8];
  load_jobs(jb, J);
  load_dst(J, d);
  load_jobs(jb + kStride, Jn);
  for (;;) {
    V v[kU][8];
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      uint32_t carry = J[u].gbase;
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint32_t own = (d[u][k] & kFirst) >> 15;
        const uint64_t f = __ballot(own != 0);
        uint32_t q;
        if constexpr (LPJ == kWave) {  // flags of the lanes below, plus my own
          q = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(f >> 32),
                                        __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(f), carry + own));
        } else {
          q = carry + static_cast<uint32_t>(__popcll(f & upto));
        }
        carry += static_cast<uint32_t>(__popcll(f & mine));
        v[u][k] = V(0);
        if (sub + LPJ * k < J[u].len) v[u][k] = vals[q];
      }
    }
    const bool more = jb + kStride < j1;  // wavefront-uniform
    uint32_t dn[kU][8];
    Job Jnn[kU];
    if (more) {
      load_dst(Jn, dn);
      load_jobs(jb + 2 * kStride, Jnn);
    }
#pragma unroll
    for (int u = 0; u < kU; ++u)
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (sub + LPJ * k < J[u].len) atomicAdd(&acc[d[u][k] & (kSpan - 1)], PbAcc<Real>::conv(v[u][k]));
    if (!more) break;
    jb += kStride;
#pragma unroll
    for (int u = 0; u < kU; ++u) {
      J[u] = Jn[u];
      Jn[u] = Jnn[u];
#pragma unroll
      for (int k = 0; k < 8; ++k) d[u][k] = dn[u][k];
    }
  }
}

// Phase 2: one task.  Fin::operator()(row, accumulated sum, resid, dang)
// finalises a row.
template <class Real, class Fin>
__global__ void __launch_bounds__(kBlock) accum_kernel(const Task* __restrict__ tasks, const Job* __restrict__ jl,
                                                       const Job* __restrict__ jm, const Job* __restrict__ js,
                                                       const typename PbAcc<Real>::V* __restrict__ vals,
                                                       const uint16_t* __restrict__ dst16,
                                                       typename PbAcc<Real>::T* __restrict__ gacc, int64_t n, Fin fin,
                                                       double* __restrict__ part_resid,
                                                       double* __restrict__ part_dang, const unsigned* stop) {
  using Acc = typename PbAcc<Real>::T;
  if (*stop) return;
  __shared__ Acc acc[kSpan];
  __shared__ double red[kWaves];
  const Task T = tasks[blockIdx.x];
  for (int i = threadIdx.x; i < kSpan; i += kBlock) acc[i] = Acc(0);
  __syncthreads();
  accum_jobs<kWave, Real>(jl, T.job0[0], T.job1[0], vals, dst16, acc);
  accum_jobs<8, Real>(jm, T.job0[1], T.job1[1], vals, dst16, acc);
  accum_jobs<1, Real>(js, T.job0[2], T.job1[2], vals, dst16, acc);
  __syncthreads();
  const int64_t r0 = static_cast<int64_t>(T.bin) << kShift;
  const int rows = static_cast<int>(n - r0 < kSpan ? n - r0 : kSpan);
  double resid = 0, dang = 0;
  if (T.slot == kSingle) {
    fin.load();
    for (int i = threadIdx.x; i < rows; i += kBlock) fin(r0 + i, acc[i], resid, dang);
  } else {
    Acc* g = gacc + static_cast<size_t>(T.slot) * kSpan;
    for (int i = threadIdx.x; i < rows; i += kBlock)
      if (acc[i] != Acc(0)) atomicAdd(&g[i], acc[i]);
  }
  resid = rs::block_sum<kBlock>(resid, red);
  dang = rs::block_sum<kBlock>(dang, red);
  if (threadIdx.x == 0) {
    part_resid[blockIdx.x] = resid;
    part_dang[blockIdx.x] = dang;
  }
}

// Finalises the sliced bins from their global accumulators and clears them.
template <class Real, class Fin>
__global__ void __launch_bounds__(kBlock) multi_kernel(const Multi* __restrict__ multi,
                                                       typename PbAcc<Real>::T* __restrict__ gacc, int64_t n,
                                                       Fin fin, double* __restrict__ part_resid,
                                                       double* __restrict__ part_dang, const unsigned* stop) {
  using Acc = typename PbAcc<Real>::T;
  if (*stop) return;
  __shared__ double red[kWaves];
  const Multi M = multi[blockIdx.x];
  const int64_t r0 = static_cast<int64_t>(M.bin) << kShift;
  const int rows = static_cast<int>(n - r0 < kSpan ? n - r0 : kSpan);
  Acc* g = gacc + static_cast<size_t>(M.slot) * kSpan;
  fin.load();
  double resid = 0, dang = 0;
  for (int i = threadIdx.x; i < rows; i += kBlock) {
    const Acc sv = g[i];
    g[i] = Acc(0);
    fin(r0 + i, sv, resid, dang);
  }
  resid = rs::block_sum<kBlock>(resid, red);
  dang = rs::block_sum<kBlock>(dang, red);
  if (threadIdx.x == 0) {
    part_resid[blockIdx.x] = resid;
    part_dang[blockIdx.x] = dang;
  }
}

// Builds g->degree_order.pb from the CSR and the degree order (which must
// exist).  Requires supported(n, m).
void build_layout(gx_graph* g);

}  // namespace pb
}  // namespace gx
