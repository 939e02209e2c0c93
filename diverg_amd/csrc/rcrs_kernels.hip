// gfx950 kernels of the range-CRS Boolean engine.
//
// Range SpGEMM (Gustavson, row-wise): row i of C = A.B is the union of the
// B rows named by A(i)'s ranges, emitted as maximal runs
// (range_sparse.hpp:1184-1264 count, :1340-1441 write).  Here each row is
// owned by a group of G lanes inside one wave64 (G = 64 for long B slices,
// fewer for short ones), with a private bit window in LDS:
//   1. stream the row's B slices -- for each A range [s, e] the flat,
//      contiguous entries B.entries[B.rowmap[s] .. B.rowmap[e+1]) -- as
//      coalesced 8-byte (start, end) loads and reduce the column extent
//      [lo, hi];
//   2. clear the window words covering [lo, hi + 1] and OR each range in
//      (ds_or_b64 on the two boundary words, plain all-ones stores between);
//   3. gather: per word, run starts map01(x, c) = x & ~(x << 1 | c) and
//      run-end successors map10(x, c) = (x << 1 | c) & ~x with c the top bit
//      of the previous word (hbitvector.hpp:577-597); a group prefix sum of
//      their popcounts places each bound, and the k-th bound of the row is
//      written as its bit position minus (k & 1).
// Bit hi + 1 always lies inside the window, so every run end is produced by
// map10 and no last-word fix-up (range_sparse.hpp:1435-1440) is needed.
// Windows wider than the LDS budget are processed in chunks, carrying the
// top bit across chunk boundaries.
//
// The counting pass keeps the first `stage` entries of every row, so rows
// whose output fits are never recomputed: after the row-pointer scan they
// are copied into place, and only the listed long rows run the scatter a
// second time in direct-write mode.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "rcrs_kernels.hpp"

namespace dv {
namespace {

constexpr int BLOCK = 256;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int G>
__device__ __forceinline__ uint32_t grp_min(uint32_t v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o, G));
  return v;
}

template <int G>
__device__ __forceinline__ uint32_t grp_max(uint32_t v) {
#pragma unroll
  for (int o = G / 2; o > 0; o >>= 1) v = max(v, (uint32_t)__shfl_xor((int)v, o, G));
  return v;
}

template <int G>
__device__ __forceinline__ uint32_t grp_scan_incl(uint32_t v, uint32_t lane) {
#pragma unroll
  for (int d = 1; d < G; d <<= 1) {
    uint32_t t = (uint32_t)__shfl_up((int)v, d, G);
    if (lane >= (uint32_t)d) v += t;
  }
  return v;
}

__device__ __forceinline__ void lds_or(uint64_t *p, uint64_t m) {
  __hip_atomic_fetch_or(p, m, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// OR bits [s, f] (window-relative, s <= f) into the window words.
__device__ __forceinline__ void window_set(uint64_t *w, uint32_t s, uint32_t f) {
  const uint32_t ws = s >> 6, wf = f >> 6;
  const uint64_t ms = ~0ull << (s & 63);
  const uint64_t mf = ~0ull >> (63 - (f & 63));
  if (ws == wf) {
    lds_or(&w[ws], ms & mf);
    return;
  }
  lds_or(&w[ws], ms);
  // All-ones is absorbing for OR, so plain stores may race with other
  // lanes' ORs into the same word.
  for (uint32_t k = ws + 1; k < wf; ++k) w[k] = ~0ull;
  lds_or(&w[wf], mf);
}

__device__ __forceinline__ uint64_t run_bounds(uint64_t x, uint64_t c) {
  const uint64_t sh = (x << 1) | c;
  return (x & ~sh) | (sh & ~x);  // map01 | map10
}

// Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
// share of the logical tiles so neighbouring rows (which read overlapping B
// slices) meet in the same L2.  gridDim.x is a multiple of 8.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb) {
  return (b & 7u) * (nb >> 3) + (b >> 3);
}

template <typename S>
struct Mat {
  const S *__restrict__ rm;
  const uint32_t *__restrict__ en;
};

// One row of C = A.B by a group of G lanes.  DIRECT: write entries at
// c_en[c_base ...]; otherwise count, stage and list long rows.
template <typename S, int G, bool DIRECT>
__device__ __forceinline__ void spgemm_row(uint64_t row, Mat<S> A, Mat<S> B,
                                           uint64_t *lds, uint32_t W,
                                           uint32_t lane, uint32_t KS,
                                           uint32_t *cnt, uint32_t *stage,
                                           uint32_t *ovf_list,
                                           unsigned int *ovf_count,
                                           uint32_t *c_en, uint64_t c_base,
                                           unsigned long long &nnz_acc,
                                           unsigned long long &slice_acc) {
  const uint64_t a0 = A.rm[row], a1 = A.rm[row + 1];

  // 1. column extent of the row's B slices
  uint32_t lo = 0xFFFFFFFFu, hi = 0u;
  for (uint64_t ai = a0; ai < a1; ai += 2) {
    const uint32_t s = A.en[ai], e = A.en[ai + 1];
    const uint64_t p0 = B.rm[s], p1 = B.rm[(uint64_t)e + 1];
    if (!DIRECT && lane == 0) slice_acc += (p1 - p0) >> 1;
    for (uint64_t p = p0 + 2 * lane; p < p1; p += 2 * G) {
      const uint2 r = *reinterpret_cast<const uint2 *>(B.en + p);
      lo = min(lo, r.x);
      hi = max(hi, r.y);
    }
  }
  lo = grp_min<G>(lo);
  hi = grp_max<G>(hi);
  if (lo > hi) {
    if (!DIRECT && lane == 0) cnt[row] = 0;
    return;
  }

  const uint64_t w0 = lo >> 6;
  const uint64_t nwords = (((uint64_t)hi + 1) >> 6) - w0 + 1;
  uint32_t count = 0;
  uint64_t carry = 0;
  for (uint64_t c0 = 0; c0 < nwords; c0 += W) {
    const uint32_t cw = (uint32_t)min((uint64_t)W, nwords - c0);
    for (uint32_t j = lane; j < cw; j += G) lds[j] = 0;
    wave_sync();

    // 2. scatter the slice ranges clipped to this chunk
    const uint64_t cb = (w0 + c0) << 6;
    const uint64_t cl = cb + ((uint64_t)cw << 6) - 1;
    for (uint64_t ai = a0; ai < a1; ai += 2) {
      const uint32_t s = A.en[ai], e = A.en[ai + 1];
      const uint64_t p1 = B.rm[(uint64_t)e + 1];
      for (uint64_t p = (uint64_t)B.rm[s] + 2 * lane; p < p1; p += 2 * G) {
        const uint2 r = *reinterpret_cast<const uint2 *>(B.en + p);
        const uint64_t rs = max((uint64_t)r.x, cb), rf = min((uint64_t)r.y, cl);
        if (rs <= rf) window_set(lds, (uint32_t)(rs - cb), (uint32_t)(rf - cb));
      }
    }
    wave_sync();

    // 3. gather: lane owns words [j0, j1)
    const uint32_t q = (cw + G - 1) / G;
    const uint32_t j0 = min(lane * q, cw), j1 = min(j0 + q, cw);
    const uint64_t cy0 = (j0 == 0) ? carry : (lds[j0 - 1] >> 63);
    uint32_t nb = 0;
    {
      uint64_t cy = cy0;
      for (uint32_t j = j0; j < j1; ++j) {
        const uint64_t x = lds[j];
        nb += __popcll(run_bounds(x, cy));
        if (!DIRECT) nnz_acc += __popcll(x);
        cy = x >> 63;
      }
    }
    const uint32_t incl = grp_scan_incl<G>(nb, lane);
    const uint32_t tot = (uint32_t)__shfl((int)incl, G - 1, G);
    uint32_t idx = count + incl - nb;
    {
      uint64_t cy = cy0;
      for (uint32_t j = j0; j < j1; ++j) {
        const uint64_t x = lds[j];
        uint64_t bnd = run_bounds(x, cy);
        cy = x >> 63;
        while (bnd) {
          const uint32_t pos = (uint32_t)__builtin_ctzll(bnd);
          bnd &= bnd - 1;
          const uint32_t val = (uint32_t)(cb + ((uint64_t)j << 6) + pos - (idx & 1u));
          if (DIRECT)
            c_en[c_base + idx] = val;
          else if (idx < KS)
            stage[row * KS + idx] = val;
          ++idx;
        }
      }
    }
    count += tot;
    carry = lds[cw - 1] >> 63;
    wave_sync();
  }
  if (!DIRECT && lane == 0) {
    cnt[row] = count;
    if (count > KS) {
      const unsigned int slot = atomicAdd(ovf_count, 1u);
      ovf_list[slot] = (uint32_t)row;
    }
  }
}

template <int G>
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
  if constexpr (G == 64) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
  } else {
    return v;
  }
}

// Block sum of v added to *dst by one lane.  Callable repeatedly (the
// trailing barrier protects the reduction buffer).
__device__ void block_add_u64(unsigned long long v, unsigned long long *dst) {
  __shared__ unsigned long long red[BLOCK / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < BLOCK / 64; ++w) t += red[w];
    if (t) atomicAdd(dst, t);
  }
  __syncthreads();
}

template <typename S, int G>
__global__ __launch_bounds__(BLOCK) void k_spgemm_count(
    Mat<S> A, Mat<S> B, uint64_t nrows, uint32_t W, uint32_t KS, uint32_t RPG,
    uint32_t *__restrict__ cnt, uint32_t *__restrict__ stage,
    uint32_t *__restrict__ ovf_list, OpCounters *__restrict__ ctr) {
  extern __shared__ uint64_t lds_all[];
  constexpr uint32_t GPB = BLOCK / G;
  const uint32_t grp = threadIdx.x / G, lane = threadIdx.x % G;
  uint64_t *lds = lds_all + (size_t)grp * W;
  unsigned long long nnz_acc = 0, slice_acc = 0;
  const uint64_t groups = (uint64_t)gridDim.x * GPB;
  const uint64_t gid = (uint64_t)xcd_tile(blockIdx.x, gridDim.x) * GPB + grp;
  for (uint64_t r0 = gid * RPG; r0 < nrows; r0 += groups * RPG) {
    const uint64_t r1 = min(r0 + RPG, nrows);
    for (uint64_t row = uniform_u64<G>(r0); row < r1; ++row)
      spgemm_row<S, G, false>(row, A, B, lds, W, lane, KS, cnt, stage, ovf_list,
                              &ctr->overflow, nullptr, 0, nnz_acc, slice_acc);
  }
  block_add_u64(nnz_acc, &ctr->nnz);
  block_add_u64(slice_acc, &ctr->slice);
}

template <typename S, int G>
__global__ __launch_bounds__(BLOCK) void k_spgemm_direct(
    Mat<S> A, Mat<S> B, uint32_t W, const uint32_t *__restrict__ rows,
    uint32_t nlisted, const S *__restrict__ c_rm, uint32_t *__restrict__ c_en) {
  extern __shared__ uint64_t lds_all[];
  constexpr uint32_t GPB = BLOCK / G;
  const uint32_t grp = threadIdx.x / G, lane = threadIdx.x % G;
  uint64_t *lds = lds_all + (size_t)grp * W;
  unsigned long long unused = 0, unused2 = 0;
  const uint32_t groups = gridDim.x * GPB;
  for (uint32_t i = blockIdx.x * GPB + grp; i < nlisted; i += groups) {
    const uint64_t row = uniform_u64<G>(rows[i]);
    spgemm_row<S, G, true>(row, A, B, lds, W, lane, 0, nullptr, nullptr, nullptr,
                           nullptr, c_en, (uint64_t)c_rm[row], unused, unused2);
  }
}

template <typename S>
__global__ __launch_bounds__(BLOCK) void k_spgemm_emit_staged(
    uint64_t nrows, const S *__restrict__ c_rm, const uint32_t *__restrict__ stage,
    uint32_t KS, uint32_t *__restrict__ c_en) {
  for (uint64_t row = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; row < nrows;
       row += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t b = c_rm[row], e = c_rm[row + 1];
    const uint64_t c = e - b;
    if (c <= KS)
      for (uint64_t k = 0; k < c; ++k) c_en[b + k] = stage[row * KS + k];
  }
}

// ---------------------------------------------------------------------------
// Row-pointer scan: tiles of BLOCK * ITEMS counts.
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = BLOCK * SCAN_ITEMS;

__device__ unsigned long long block_scan_incl(unsigned long long v,
                                              unsigned long long &total) {
  __shared__ unsigned long long wsum[BLOCK / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  if (lane == 63) wsum[wid] = v;
  __syncthreads();
  unsigned long long off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < BLOCK / 64; ++w) {
    if (w < wid) off += wsum[w];
    tot += wsum[w];
  }
  __syncthreads();
  total = tot;
  return v + off;
}

__global__ __launch_bounds__(BLOCK) void k_scan_tiles(const uint32_t *__restrict__ cnt,
                                                      uint64_t n,
                                                      unsigned long long *__restrict__ tile_sum) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  unsigned long long s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k)
    if (base + k < n) s += cnt[base + k];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  __shared__ unsigned long long red[BLOCK / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < BLOCK / 64; ++w) t += red[w];
    tile_sum[blockIdx.x] = t;
  }
}

__global__ __launch_bounds__(BLOCK) void k_scan_spine(unsigned long long *__restrict__ tile_sum,
                                                      uint64_t ntiles,
                                                      OpCounters *__restrict__ ctr) {
  unsigned long long carry = 0;
  for (uint64_t base = 0; base < ntiles; base += BLOCK) {
    const uint64_t i = base + threadIdx.x;
    const unsigned long long v = i < ntiles ? tile_sum[i] : 0ull;
    unsigned long long tot;
    const unsigned long long incl = block_scan_incl(v, tot);
    if (i < ntiles) tile_sum[i] = carry + incl - v;
    carry += tot;
  }
  if (threadIdx.x == 0) ctr->total = carry;
}

template <typename S>
__global__ __launch_bounds__(BLOCK) void k_scan_write(const uint32_t *__restrict__ cnt,
                                                      uint64_t n,
                                                      const unsigned long long *__restrict__ tile_off,
                                                      S *__restrict__ rm) {
  const uint64_t base = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  unsigned long long s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = base + k < n ? cnt[base + k] : 0u;
    s += v[k];
  }
  unsigned long long tot;
  unsigned long long run = block_scan_incl(s, tot) - s + tile_off[blockIdx.x];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    run += v[k];
    if (base + k < n) rm[base + k + 1] = (S)run;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) rm[0] = 0;
}

// ---------------------------------------------------------------------------
// Range SpAdd, one lane per row (range_sparse.hpp:145-224 and :273-356).
// The merge takes the side with the smaller start (A on ties), folds a range
// into the pending one while rs <= hi + 1, and once either side is used up
// folds only until the first range that does not touch; that range and all
// later ones are copied as they are.
template <typename S, bool WRITE>
__device__ __forceinline__ uint64_t spadd_row(const uint32_t *__restrict__ ae, uint64_t ai,
                                              uint64_t aend, const uint32_t *__restrict__ be,
                                              uint64_t bi, uint64_t bend,
                                              uint32_t *__restrict__ out) {
  uint64_t c = 0;
  auto put = [&](uint32_t x) {
    if (WRITE) out[c] = x;
    ++c;
  };
  if (ai < aend && bi < bend) {
    uint32_t lo, hi;
    auto take_min = [&](uint32_t &rs, uint32_t &re) {
      if (ae[ai] <= be[bi]) {
        rs = ae[ai];
        re = ae[ai + 1];
        ai += 2;
      } else {
        rs = be[bi];
        re = be[bi + 1];
        bi += 2;
      }
    };
    take_min(lo, hi);
    while (ai < aend && bi < bend) {
      uint32_t rs, re;
      take_min(rs, re);
      if (rs <= hi + 1u) {
        lo = min(lo, rs);
        hi = max(hi, re);
      } else {
        put(lo);
        put(hi);
        lo = rs;
        hi = re;
      }
    }
    while (ai < aend || bi < bend) {
      uint32_t rs, re;
      if (ai == aend) {
        rs = be[bi];
        re = be[bi + 1];
        bi += 2;
      } else {
        rs = ae[ai];
        re = ae[ai + 1];
        ai += 2;
      }
      if (rs <= hi + 1u) {
        lo = min(lo, rs);
        hi = max(hi, re);
        continue;
      }
      put(lo);
      put(hi);
      lo = rs;
      hi = re;
      break;
    }
    put(lo);
    put(hi);
  }
  for (; ai < aend; ++ai) put(ae[ai]);
  for (; bi < bend; ++bi) put(be[bi]);
  return c;
}

template <typename S>
__global__ __launch_bounds__(BLOCK) void k_spadd_count(Mat<S> A, Mat<S> B, uint64_t n,
                                                       uint32_t *__restrict__ cnt) {
  for (uint64_t row = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; row < n;
       row += (uint64_t)gridDim.x * BLOCK)
    cnt[row] = (uint32_t)spadd_row<S, false>(A.en, A.rm[row], A.rm[row + 1], B.en,
                                             B.rm[row], B.rm[row + 1], nullptr);
}

template <typename S>
__global__ __launch_bounds__(BLOCK) void k_spadd_write(Mat<S> A, Mat<S> B, uint64_t n,
                                                       const S *__restrict__ c_rm,
                                                       uint32_t *__restrict__ c_en) {
  for (uint64_t row = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; row < n;
       row += (uint64_t)gridDim.x * BLOCK)
    spadd_row<S, true>(A.en, A.rm[row], A.rm[row + 1], B.en, B.rm[row], B.rm[row + 1],
                       c_en + c_rm[row]);
}

// nnz = sum of (e - s + 1) over all ranges (crs_matrix.hpp:175-189).
__global__ __launch_bounds__(BLOCK) void k_range_nnz(const uint2 *__restrict__ r, uint64_t nr,
                                                     OpCounters *__restrict__ ctr) {
  unsigned long long s = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < nr;
       i += (uint64_t)gridDim.x * BLOCK) {
    const uint2 v = r[i];
    s += (unsigned long long)v.y - v.x + 1;
  }
  block_add_u64(s, &ctr->nnz);
}

template <typename S>
__global__ __launch_bounds__(BLOCK) void k_identity(uint64_t n, S *__restrict__ rm,
                                                    uint32_t *__restrict__ en) {
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * BLOCK) {
    reinterpret_cast<uint2 *>(en)[i] = make_uint2((uint32_t)i, (uint32_t)i);
    rm[i + 1] = (S)(2 * (i + 1));
    if (i == 0) rm[0] = 0;
  }
}

// Lower bound of j in the row's flat entries; an odd position steps back to
// the range start (crs_matrix.hpp:149-169).
template <typename S>
__global__ __launch_bounds__(BLOCK) void k_query(Mat<S> M, uint64_t nrows, uint64_t ncols,
                                                 const uint32_t *__restrict__ qi,
                                                 const uint32_t *__restrict__ qj,
                                                 uint8_t *__restrict__ out, uint64_t n,
                                                 uint32_t *__restrict__ bad) {
  for (uint64_t q = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; q < n;
       q += (uint64_t)gridDim.x * BLOCK) {
    const uint32_t i = qi[q], j = qj[q];
    if (i >= nrows || j >= ncols) {
      out[q] = 2;
      *bad = 1u;
      continue;
    }
    const uint64_t b = M.rm[i], e = M.rm[(uint64_t)i + 1];
    uint64_t first = b, len = e - b;
    while (len > 0) {
      const uint64_t half = len >> 1;
      if (M.en[first + half] < j) {
        first += half + 1;
        len -= half + 1;
      } else {
        len = half;
      }
    }
    if ((first - b) & 1) --first;
    out[q] = (uint8_t)(first != e && M.en[first] <= j);
  }
}

template <typename S>
__global__ __launch_bounds__(BLOCK) void k_slice_lengths(Mat<S> A, Mat<S> B, uint64_t n,
                                                         uint32_t *__restrict__ len) {
  for (uint64_t row = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; row < n;
       row += (uint64_t)gridDim.x * BLOCK) {
    uint64_t t = 0;
    for (uint64_t ai = A.rm[row]; ai < A.rm[row + 1]; ai += 2)
      t += (uint64_t)B.rm[(uint64_t)A.en[ai + 1] + 1] - B.rm[A.en[ai]];
    t >>= 1;
    len[row] = t > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)t;
  }
}

// Input validation for device-resident matrices handed to the library.
template <typename S>
__global__ __launch_bounds__(BLOCK) void k_validate(Mat<S> M, uint64_t nrows, uint64_t ncols,
                                                    uint64_t n_entries,
                                                    uint32_t *__restrict__ flag) {
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  const uint64_t t0 = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (t0 == 0 && (M.rm[0] != 0 || (uint64_t)M.rm[nrows] != n_entries)) *flag = 1u;
  for (uint64_t row = t0; row < nrows; row += stride) {
    const uint64_t b = M.rm[row], e = M.rm[row + 1];
    if (e < b || ((e - b) & 1) || e > n_entries) *flag = 1u;
  }
  const uint2 *r = reinterpret_cast<const uint2 *>(M.en);
  for (uint64_t k = t0; k < n_entries / 2; k += stride) {
    const uint2 v = r[k];
    if (v.x > v.y || v.y >= ncols) *flag = 1u;
  }
}

__global__ __launch_bounds__(BLOCK) void k_fill_u32(uint32_t *__restrict__ p, uint64_t n,
                                                    uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * BLOCK)
    p[i] = v;
}

inline uint32_t grid_for(uint64_t n, uint64_t per_block = BLOCK, uint32_t cap = 65536) {
  uint64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  return (uint32_t)(g > cap ? cap : g);
}

template <typename S>
Mat<S> mat(DView v) {
  return Mat<S>{static_cast<const S *>(v.rowmap), v.entries};
}

template <typename S, int G>
void launch_count(hipStream_t s, DView a, DView b, const SpgemmShape &sh, uint32_t *cnt,
                  uint32_t *stage, uint32_t *ovf, OpCounters *ctr) {
  const size_t lds = (size_t)(BLOCK / G) * sh.lds_words * sizeof(uint64_t);
  hipLaunchKernelGGL((k_spgemm_count<S, G>), dim3(sh.grid), dim3(BLOCK), lds, s, mat<S>(a),
                     mat<S>(b), a.nrows, sh.lds_words, sh.stage, sh.rows_per_group, cnt,
                     stage, ovf, ctr);
}

template <typename S, int G>
void launch_direct(hipStream_t s, DView a, DView b, const SpgemmShape &sh,
                   const uint32_t *rows, uint32_t nl, const void *c_rm, uint32_t *c_en) {
  const size_t lds = (size_t)(BLOCK / G) * sh.lds_words * sizeof(uint64_t);
  const uint32_t grid = grid_for(nl, BLOCK / G, 1u << 20);
  hipLaunchKernelGGL((k_spgemm_direct<S, G>), dim3(grid), dim3(BLOCK), lds, s, mat<S>(a),
                     mat<S>(b), sh.lds_words, rows, nl, static_cast<const S *>(c_rm), c_en);
}

template <typename S>
void count_by_lanes(hipStream_t s, DView a, DView b, const SpgemmShape &sh, uint32_t *cnt,
                    uint32_t *stage, uint32_t *ovf, OpCounters *ctr) {
  switch (sh.lanes) {
    case 4: launch_count<S, 4>(s, a, b, sh, cnt, stage, ovf, ctr); break;
    case 8: launch_count<S, 8>(s, a, b, sh, cnt, stage, ovf, ctr); break;
    case 16: launch_count<S, 16>(s, a, b, sh, cnt, stage, ovf, ctr); break;
    case 32: launch_count<S, 32>(s, a, b, sh, cnt, stage, ovf, ctr); break;
    default: launch_count<S, 64>(s, a, b, sh, cnt, stage, ovf, ctr); break;
  }
}

template <typename S>
void direct_by_lanes(hipStream_t s, DView a, DView b, const SpgemmShape &sh,
                     const uint32_t *rows, uint32_t nl, const void *c_rm, uint32_t *c_en) {
  switch (sh.lanes) {
    case 4: launch_direct<S, 4>(s, a, b, sh, rows, nl, c_rm, c_en); break;
    case 8: launch_direct<S, 8>(s, a, b, sh, rows, nl, c_rm, c_en); break;
    case 16: launch_direct<S, 16>(s, a, b, sh, rows, nl, c_rm, c_en); break;
    case 32: launch_direct<S, 32>(s, a, b, sh, rows, nl, c_rm, c_en); break;
    default: launch_direct<S, 64>(s, a, b, sh, rows, nl, c_rm, c_en); break;
  }
}

}  // namespace

void fill_u32(hipStream_t s, uint32_t *p, uint64_t n, uint32_t v) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(BLOCK), 0, s, p, n, v);
}

void zero_counters(hipStream_t s, OpCounters *c) {
  fill_u32(s, reinterpret_cast<uint32_t *>(c), sizeof(OpCounters) / 4, 0u);
}

void identity(hipStream_t s, uint32_t sb, uint64_t n, void *rowmap, uint32_t *entries) {
  if (n == 0) {
    // rowmap[0] = 0 for an empty matrix
    fill_u32(s, static_cast<uint32_t *>(rowmap), sb / 4, 0u);
    return;
  }
  if (sb == 4)
    hipLaunchKernelGGL(k_identity<uint32_t>, dim3(grid_for(n)), dim3(BLOCK), 0, s, n,
                       static_cast<uint32_t *>(rowmap), entries);
  else
    hipLaunchKernelGGL(k_identity<uint64_t>, dim3(grid_for(n)), dim3(BLOCK), 0, s, n,
                       static_cast<uint64_t *>(rowmap), entries);
}

uint64_t scan_tmp_words(uint64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

void scan_counts(hipStream_t s, uint32_t sb, const uint32_t *cnt, uint64_t n, void *rowmap,
                 unsigned long long *tmp, OpCounters *ctr) {
  if (n == 0) {
    fill_u32(s, static_cast<uint32_t *>(rowmap), sb / 4, 0u);
    return;
  }
  const uint64_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  hipLaunchKernelGGL(k_scan_tiles, dim3((uint32_t)ntiles), dim3(BLOCK), 0, s, cnt, n, tmp);
  hipLaunchKernelGGL(k_scan_spine, dim3(1), dim3(BLOCK), 0, s, tmp, ntiles, ctr);
  if (sb == 4)
    hipLaunchKernelGGL(k_scan_write<uint32_t>, dim3((uint32_t)ntiles), dim3(BLOCK), 0, s, cnt,
                       n, tmp, static_cast<uint32_t *>(rowmap));
  else
    hipLaunchKernelGGL(k_scan_write<uint64_t>, dim3((uint32_t)ntiles), dim3(BLOCK), 0, s, cnt,
                       n, tmp, static_cast<uint64_t *>(rowmap));
}

void spadd_count(hipStream_t s, uint32_t sb, DView a, DView b, uint32_t *cnt) {
  if (!a.nrows) return;
  if (sb == 4)
    hipLaunchKernelGGL(k_spadd_count<uint32_t>, dim3(grid_for(a.nrows)), dim3(BLOCK), 0, s,
                       mat<uint32_t>(a), mat<uint32_t>(b), a.nrows, cnt);
  else
    hipLaunchKernelGGL(k_spadd_count<uint64_t>, dim3(grid_for(a.nrows)), dim3(BLOCK), 0, s,
                       mat<uint64_t>(a), mat<uint64_t>(b), a.nrows, cnt);
}

void spadd_write(hipStream_t s, uint32_t sb, DView a, DView b, const void *c_rowmap,
                 uint32_t *c_entries) {
  if (!a.nrows) return;
  if (sb == 4)
    hipLaunchKernelGGL(k_spadd_write<uint32_t>, dim3(grid_for(a.nrows)), dim3(BLOCK), 0, s,
                       mat<uint32_t>(a), mat<uint32_t>(b), a.nrows,
                       static_cast<const uint32_t *>(c_rowmap), c_entries);
  else
    hipLaunchKernelGGL(k_spadd_write<uint64_t>, dim3(grid_for(a.nrows)), dim3(BLOCK), 0, s,
                       mat<uint64_t>(a), mat<uint64_t>(b), a.nrows,
                       static_cast<const uint64_t *>(c_rowmap), c_entries);
}

void range_nnz(hipStream_t s, const uint32_t *entries, uint64_t n_entries, OpCounters *ctr) {
  if (n_entries < 2) return;
  const uint64_t nr = n_entries / 2;
  hipLaunchKernelGGL(k_range_nnz, dim3(grid_for(nr, BLOCK * 8, 8192)), dim3(BLOCK), 0, s,
                     reinterpret_cast<const uint2 *>(entries), nr, ctr);
}

void spgemm_count(hipStream_t s, uint32_t sb, DView a, DView b, const SpgemmShape &shape,
                  uint32_t *cnt, uint32_t *stage_buf, uint32_t *ovf_list, OpCounters *ctr) {
  if (!a.nrows) return;
  if (sb == 4)
    count_by_lanes<uint32_t>(s, a, b, shape, cnt, stage_buf, ovf_list, ctr);
  else
    count_by_lanes<uint64_t>(s, a, b, shape, cnt, stage_buf, ovf_list, ctr);
}

void spgemm_emit_staged(hipStream_t s, uint32_t sb, uint64_t nrows, const void *c_rowmap,
                        const uint32_t *stage_buf, uint32_t stage, uint32_t *c_entries) {
  if (!nrows) return;
  if (sb == 4)
    hipLaunchKernelGGL(k_spgemm_emit_staged<uint32_t>, dim3(grid_for(nrows)), dim3(BLOCK), 0,
                       s, nrows, static_cast<const uint32_t *>(c_rowmap), stage_buf, stage,
                       c_entries);
  else
    hipLaunchKernelGGL(k_spgemm_emit_staged<uint64_t>, dim3(grid_for(nrows)), dim3(BLOCK), 0,
                       s, nrows, static_cast<const uint64_t *>(c_rowmap), stage_buf, stage,
                       c_entries);
}

void spgemm_direct(hipStream_t s, uint32_t sb, DView a, DView b, const SpgemmShape &shape,
                   const uint32_t *rows, uint32_t nrows_listed, const void *c_rowmap,
                   uint32_t *c_entries) {
  if (!nrows_listed) return;
  if (sb == 4)
    direct_by_lanes<uint32_t>(s, a, b, shape, rows, nrows_listed, c_rowmap, c_entries);
  else
    direct_by_lanes<uint64_t>(s, a, b, shape, rows, nrows_listed, c_rowmap, c_entries);
}

void query(hipStream_t s, uint32_t sb, DView m, const uint32_t *qi, const uint32_t *qj,
           uint8_t *out, uint64_t n, uint32_t *bad) {
  if (!n) return;
  if (sb == 4)
    hipLaunchKernelGGL(k_query<uint32_t>, dim3(grid_for(n)), dim3(BLOCK), 0, s,
                       mat<uint32_t>(m), m.nrows, m.ncols, qi, qj, out, n, bad);
  else
    hipLaunchKernelGGL(k_query<uint64_t>, dim3(grid_for(n)), dim3(BLOCK), 0, s,
                       mat<uint64_t>(m), m.nrows, m.ncols, qi, qj, out, n, bad);
}

void validate_matrix(hipStream_t s, uint32_t sb, DView m, uint64_t n_entries, uint32_t *flag) {
  const uint64_t work = std::max<uint64_t>(m.nrows, n_entries / 2);
  if (sb == 4)
    hipLaunchKernelGGL(k_validate<uint32_t>, dim3(grid_for(work)), dim3(BLOCK), 0, s,
                       mat<uint32_t>(m), m.nrows, m.ncols, n_entries, flag);
  else
    hipLaunchKernelGGL(k_validate<uint64_t>, dim3(grid_for(work)), dim3(BLOCK), 0, s,
                       mat<uint64_t>(m), m.nrows, m.ncols, n_entries, flag);
}

void slice_lengths(hipStream_t s, uint32_t sb, DView a, DView b, uint32_t *len) {
  if (!a.nrows) return;
  if (sb == 4)
    hipLaunchKernelGGL(k_slice_lengths<uint32_t>, dim3(grid_for(a.nrows)), dim3(BLOCK), 0, s,
                       mat<uint32_t>(a), mat<uint32_t>(b), a.nrows, len);
  else
    hipLaunchKernelGGL(k_slice_lengths<uint64_t>, dim3(grid_for(a.nrows)), dim3(BLOCK), 0, s,
                       mat<uint64_t>(a), mat<uint64_t>(b), a.nrows, len);
}

}  // namespace dv
