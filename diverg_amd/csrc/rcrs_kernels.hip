// gfx950 kernels of the range-CRS Boolean engine.
//
// Range SpGEMM (Gustavson, row-wise): row i of C = A.B is the union of the
// B rows named by A(i)'s ranges, emitted as maximal runs
// (range_sparse.hpp:1184-1264 count, :1340-1441 write).  For each A range
// [s, e] the B rows s..e form one flat, contiguous slice of B's entries,
// B.entries[B.rowmap[s] .. B.rowmap[e+1]), read as 8-byte (start, end) pairs.
//
// Rows are split by work into three tiers, all producing the same canonical
// runs:
//   lane tier   one lane per row: the union is folded range by range into
//               at most 4 disjoint intervals held in registers, then sorted
//               and swept (a range-merge compaction).  Rows of banded
//               matrices read neighbouring slices, so lane-per-row loads
//               coalesce.  Covers rows whose union has at most 4 runs.
//   16-lane and 64-lane groups (rows the lane tier hands over through
//               wave-aggregated lists): the group reduces the extent
//               [lo, hi], clears a private LDS bit window over [lo, hi + 1],
//               ORs every range in (ds_or_b64 on boundary words, all-ones
//               stores between), then gathers.  Slices up to 4 ranges per
//               lane stay in registers between the two passes.
// Gather, in every tier: per 64-bit word x with c the top bit of the previous
// word, run starts are map01(x, c) = x & ~(x << 1 | c) and run-end successors
// map10(x, c) = (x << 1 | c) & ~x (hbitvector.hpp:577-597); the k-th bound of
// a row is written as its bit position minus (k & 1).  Bit hi + 1 always lies
// inside the window, so every run end comes from map10 and no last-word
// fix-up (range_sparse.hpp:1435-1440) is needed.  Windows wider than the LDS
// budget are processed in chunks, carrying the top bit across chunks.
//
// The counting pass keeps the first `stage` entries of each row (range k of
// row r as one uint2 at stage[k * nrows + r], so lane-tier stores and the
// later copy coalesce across rows); after the row-pointer scan they are
// copied into place and only rows with longer output are recomputed, in
// direct-write mode.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>

#include "rcrs_kernels.hpp"

// Tests on bools are written with & and | where both operands are cheap:
// the short-circuit forms let the compiler branch, and a load it sinks into
// such a branch costs a round trip of its own.
#pragma clang diagnostic ignored "-Wbitwise-instead-of-logical"

namespace dv {
namespace {

constexpr int BLOCK = 256;
constexpr int CACHE = 4;             // group tiers: cached ranges per lane

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

__device__ __forceinline__ uint2 ld_range(const uint32_t *__restrict__ en, uint64_t p) {
  return *reinterpret_cast<const uint2 *>(en + p);
}

// Use a loaded value here, so that the compiler waits for its load at this
// point.  Without it the wait can land at a join or a loop back edge, where
// the memory counter (loads and stores complete in order) also makes the wave
// wait for every load and store issued after it: prefetches meant for a later
// row, or the current row's stores.
template <typename T>
__device__ __forceinline__ void need(T x) {
  asm volatile("" ::"v"(x));
}

// Append `value` to list[*counter ...] for the flagged lanes of the wave with
// one atomic per wave.  Must be reached by every lane of the wave.
__device__ __forceinline__ void wave_append(bool flag, uint32_t value, uint32_t *list,
                                            unsigned int *counter) {
  const unsigned long long m = __ballot(flag);
  if (m == 0) return;
  const uint32_t lane = __lane_id();
  const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
  unsigned int base = 0;
  if (lane == leader) base = atomicAdd(counter, (unsigned int)__popcll(m));
  base = (unsigned int)__shfl((int)base, (int)leader);
  if (flag) list[base + __popcll(m & ((1ull << lane) - 1))] = value;
}

// Sums over a wave / a block (every thread of the block must call it; the
// trailing barrier protects the reduction buffer for the next call).
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  return v;
}

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v) {
  __shared__ unsigned long long red[BLOCK / 64];
  v = wave_sum_u64(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  unsigned long long t = 0;
#pragma unroll
  for (int w = 0; w < BLOCK / 64; ++w) t += red[w];
  __syncthreads();
  return t;
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

// ---------------------------------------------------------------------------
// Lane tier: one lane per row.  The row's union is kept as at most
// LANE_SLOTS disjoint intervals in registers.  Each slice range is folded
// into the first interval it touches (overlap or adjacency), else into a free
// slot; when no slot is free the intervals are compacted (sorted by start and
// swept, merging touching neighbours) and the insert retried.  The final
// compaction leaves the maximal runs in ascending order.  A row whose union
// needs more than LANE_SLOTS runs, or whose slice exceeds lane_max ranges,
// is handed to the group tiers.
constexpr int LANE_SLOTS = 4;
constexpr int LANE_BATCH = 8;  // slice loads issued together per lane

struct Runs {
  uint32_t lo[LANE_SLOTS], hi[LANE_SLOTS];  // empty slot: lo = 0xFFFFFFFF, hi = 0
};

__device__ __forceinline__ void cswap(Runs &r, int i, int j) {
  const bool sw = r.lo[j] < r.lo[i];
  const uint32_t l = sw ? r.lo[j] : r.lo[i], h = sw ? r.hi[j] : r.hi[i];
  r.lo[j] = sw ? r.lo[i] : r.lo[j];
  r.hi[j] = sw ? r.hi[i] : r.hi[j];
  r.lo[i] = l;
  r.hi[i] = h;
}

__device__ __forceinline__ void sort4(Runs &r) {
  cswap(r, 0, 1);
  cswap(r, 2, 3);
  cswap(r, 0, 2);
  cswap(r, 1, 3);
  cswap(r, 1, 2);
}

// Sort slots by start (empty slots, start 0xFFFFFFFF, go last), fold each
// slot that touches its predecessor forward into it (the predecessor is
// emptied, so slot j always holds the running merged interval), and sort
// again to move the emptied slots behind the live ones.
__device__ __forceinline__ void compact(Runs &r) {
  sort4(r);
#pragma unroll
  for (int j = 1; j < LANE_SLOTS; ++j) {
    const bool live = r.lo[j - 1] <= r.hi[j - 1] && r.lo[j] <= r.hi[j];
    if (live && r.lo[j] <= r.hi[j - 1] + 1u) {
      r.lo[j] = r.lo[j - 1];
      r.hi[j] = max(r.hi[j - 1], r.hi[j]);
      r.lo[j - 1] = 0xFFFFFFFFu;
      r.hi[j - 1] = 0u;
    }
  }
  sort4(r);
}

// Fold [s, e] into the runs; false when no slot can take it.
__device__ __forceinline__ bool insert(Runs &r, uint32_t s, uint32_t e) {
  bool done = false;
#pragma unroll
  for (int t = 0; t < LANE_SLOTS; ++t) {
    const bool live = r.lo[t] <= r.hi[t];
    if (!done && live && s <= r.hi[t] + 1u && e + 1u >= r.lo[t]) {
      r.lo[t] = min(r.lo[t], s);
      r.hi[t] = max(r.hi[t], e);
      done = true;
    }
  }
#pragma unroll
  for (int t = 0; t < LANE_SLOTS; ++t) {
    if (!done && r.lo[t] > r.hi[t]) {
      r.lo[t] = s;
      r.hi[t] = e;
      done = true;
    }
  }
  return done;
}

// Fold one range into the running interval [cl, ch] / the slots.
__device__ __forceinline__ void fold(uint32_t vs, uint32_t ve, uint32_t &cl, uint32_t &ch,
                                     Runs &r, bool &ok) {
  if (vs > ve) return;
  if (vs <= ch + 1u && ve + 1u >= cl) {
    cl = min(cl, vs);
    ch = max(ch, ve);
  } else {
    if (cl <= ch && ok && !insert(r, cl, ch)) {
      compact(r);
      ok = insert(r, cl, ch);
    }
    cl = vs;
    ch = ve;
  }
}

// Fold a batch of ranges (empty ones are (1, 0)) in order.  The common case,
// every range touching the running interval as it grows (a banded slice), is
// done without branches; a lane with any other range redoes the batch from
// the same state through fold(), so the result is fold()'s either way.
template <int N>
__device__ __forceinline__ void fold_batch(const uint2 (&v)[N], uint32_t &cl, uint32_t &ch,
                                           Runs &r, bool &ok) {
  uint32_t ncl = cl, nch = ch;
  bool all = true;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const bool valid = v[k].x <= v[k].y;
    // an empty running interval takes the range as it is
    const bool touch = ncl > nch || (v[k].x <= nch + 1u && v[k].y + 1u >= ncl);
    all = all && (!valid || touch);
    ncl = valid && touch ? min(ncl, v[k].x) : ncl;
    nch = valid && touch ? max(nch, v[k].y) : nch;
  }
  if (all) {
    cl = ncl;
    ch = nch;
  } else {
#pragma unroll
    for (int k = 0; k < N; ++k) fold(v[k].x, v[k].y, cl, ch, r, ok);
  }
}

// Fold the flat run of ranges en[p0 .. p1) with LANE_BATCH loads in flight.
// Offsets are of the rowmap's size type (32-bit arithmetic for 32-bit rowmaps).
// FAST selects fold_batch's branch-free common case; without it each range
// goes through fold() (fewer registers).
template <typename S, bool FAST = true>
__device__ __forceinline__ void fold_segment(const uint32_t *__restrict__ en, S p0, S p1,
                                             uint32_t &cl, uint32_t &ch, Runs &r, bool &ok,
                                             uint32_t &gathered) {
  gathered += (uint32_t)((p1 - p0) >> 1);
  for (S p = p0; ok && p < p1;) {
    const S rem = p1 - p;  // no offset past p1 is formed (it could wrap)
    uint2 v[LANE_BATCH];
#pragma unroll
    for (int k = 0; k < LANE_BATCH; ++k)
      v[k] = (S)(2 * k) < rem ? ld_range(en, p + 2 * k) : make_uint2(1u, 0u);
    if (FAST) {
      fold_batch(v, cl, ch, r, ok);
    } else {
#pragma unroll
      for (int k = 0; k < LANE_BATCH; ++k) fold(v[k].x, v[k].y, cl, ch, r, ok);
    }
    p = rem > (S)(2 * LANE_BATCH) ? p + 2 * LANE_BATCH : p1;
  }
}

// Interval summaries of the right operand, per block k of W rows (rows
// Wk .. Wk+W-1, W a power of two from 32 to 256, chosen per product from how
// many rows its ranges span): pre[r] is the union of rows Wk .. r and suf[r]
// the union of rows r .. Wk+W-1, each as its one or two pieces when it has at
// most two (ivl_store), else "unknown".  The whole block's union is
// pre[Wk + W - 1].
// Partial unions in the scan are kept exactly as up to two disjoint
// intervals, so a row with a gap (a bubble's alternate locus reaches past the
// reference locus) does not spoil the block when neighbouring rows fill the
// gap; a partial union of three or more pieces becomes "unknown" (the lane
// tier then reads those rows' ranges).  A summary is never claimed for a
// union that is not one interval.
constexpr uint32_t SUMM_ROWS_MIN = 32, SUMM_ROWS_MAX = 256;
constexpr uint32_t SUMM_EMPTY = 0xFFFFFFFFu;

// A partial union: up to two pieces in ascending order, an unused piece is
// (EMPTY, 0); the empty union has lo0 = EMPTY, an unknown one is (1, 0).
struct Ivl2 {
  uint32_t lo0, hi0, lo1, hi1;
};

__device__ __forceinline__ bool ivl_empty(const Ivl2 &v) { return v.lo0 == SUMM_EMPTY; }
__device__ __forceinline__ bool ivl_unknown(const Ivl2 &v) {
  return v.lo0 != SUMM_EMPTY && v.lo0 > v.hi0;
}

__device__ __forceinline__ void cswap_iv(uint32_t &la, uint32_t &ha, uint32_t &lb, uint32_t &hb) {
  const bool sw = lb < la;
  const uint32_t l = sw ? lb : la, h = sw ? hb : ha;
  lb = sw ? la : lb;
  hb = sw ? ha : hb;
  la = l;
  ha = h;
}

// The union of two partial unions.  The merge of the general case is
// computed with selects (a wave's lanes take different cases in it, and a
// branch per case would run them one after another).
__device__ __forceinline__ Ivl2 ivl_join(const Ivl2 &a, const Ivl2 &b) {
  if (ivl_empty(a)) return b;
  if (ivl_empty(b)) return a;
  const Ivl2 u{1u, 0u, SUMM_EMPTY, 0u};
  if (ivl_unknown(a) | ivl_unknown(b)) return u;
  // the common case: two single intervals that touch
  if ((a.lo1 == SUMM_EMPTY) & (b.lo1 == SUMM_EMPTY) & (b.lo0 <= a.hi0 + 1u) &
      (b.hi0 + 1u >= a.lo0))
    return Ivl2{min(a.lo0, b.lo0), max(a.hi0, b.hi0), SUMM_EMPTY, 0u};
  uint32_t L0 = a.lo0, H0 = a.hi0, L1 = a.lo1, H1 = a.hi1, L2 = b.lo0, H2 = b.hi0, L3 = b.lo1,
           H3 = b.hi1;
  // each side's pieces are already ascending (an unused piece starts at
  // EMPTY): merging the two sorted pairs takes three compare-swaps
  cswap_iv(L0, H0, L2, H2);
  cswap_iv(L1, H1, L3, H3);
  cswap_iv(L1, H1, L2, H2);
  // sweep the pieces in order, merging each into the last output piece it
  // touches; a third output piece makes the union unknown
  uint32_t l0 = SUMM_EMPTY, h0 = 0u, l1 = SUMM_EMPTY, h1 = 0u;
  uint32_t k = 0;  // output pieces so far
  bool over = false;
  auto put = [&](uint32_t L, uint32_t H) {
    const bool valid = L <= H;  // an unused piece is skipped
    const bool t0 = L <= h0 + 1u, t1 = L <= h1 + 1u;
    const bool first = valid & (k == 0);
    const bool grow0 = valid & (k == 1) & t0, second = valid & (k == 1) & !t0;
    const bool grow1 = valid & (k == 2) & t1;
    over = over | (valid & (k == 2) & !t1);
    l0 = first ? L : l0;
    h0 = first ? H : grow0 ? max(h0, H) : h0;
    l1 = second ? L : l1;
    h1 = second ? H : grow1 ? max(h1, H) : h1;
    k = first ? 1u : second ? 2u : k;
  };
  put(L0, H0);
  put(L1, H1);
  put(L2, H2);
  put(L3, H3);
  return over ? u : Ivl2{l0, h0, l1, h1};
}

// A stored summary is 8 bytes: (lo, hi) for one interval, (EMPTY, 0) for
// the empty union, (1, 0) for an unknown one and (2, 0) for a union of two
// pieces, which are then stored in a 16-byte side entry of the same row
// (written only for such rows).
constexpr uint32_t SUMM_UNKNOWN = 1u, SUMM_TWO = 2u;

__device__ __forceinline__ void ivl_store(const Ivl2 &v, uint2 *__restrict__ one,
                                          uint4 *__restrict__ two) {
  if (ivl_empty(v)) {
    *one = make_uint2(SUMM_EMPTY, 0u);
  } else if (ivl_unknown(v)) {
    *one = make_uint2(SUMM_UNKNOWN, 0u);
  } else if (v.lo1 == SUMM_EMPTY) {
    *one = make_uint2(v.lo0, v.hi0);
  } else {
    *one = make_uint2(SUMM_TWO, 0u);
    *two = make_uint4(v.lo0, v.hi0, v.lo1, v.hi1);
  }
}

// Shuffles within aligned groups of w lanes (w a power of two, at most 64).
constexpr int SUMM_ROWS_PER_LANE = 4;

// The value of lane L - D (UP) or L + D in a group of W lanes.  Groups of at
// most 16 lanes stay within a 16-lane DPP row, so a row shift does it (one
// VALU move; a lane whose source lies outside the row keeps its own value);
// wider groups go through the LDS crossbar.  Callers use the value only where
// the source lane is in the lane's own group.
template <int W, int D, bool UP>
__device__ __forceinline__ uint32_t lane_shift(uint32_t x) {
  if constexpr (W <= 16)
    return (uint32_t)__builtin_amdgcn_update_dpp((int)x, (int)x, (UP ? 0x110 : 0x100) + D, 0xF,
                                                 0xF, false);
  else
    return UP ? (uint32_t)__shfl_up((int)x, D, W) : (uint32_t)__shfl_down((int)x, D, W);
}

template <int W, int D, bool UP>
__device__ __forceinline__ Ivl2 ivl_shift(const Ivl2 &v) {
  return Ivl2{lane_shift<W, D, UP>(v.lo0), lane_shift<W, D, UP>(v.hi0),
              lane_shift<W, D, UP>(v.lo1), lane_shift<W, D, UP>(v.hi1)};
}

// Single intervals [lo, hi]: one step of the scans of the fast path, which
// holds while every join is of two touching intervals (ok).
__device__ __forceinline__ bool iv_touch(uint32_t alo, uint32_t ahi, uint32_t blo, uint32_t bhi) {
  return (blo <= ahi + 1u) & (bhi + 1u >= alo);
}

template <int W, int D>
__device__ __forceinline__ void iv_scan_step(uint32_t &flo, uint32_t &fhi, uint32_t &glo,
                                             uint32_t &ghi, bool &ok, int pos) {
  if constexpr (D < W) {
    const uint32_t olo = lane_shift<W, D, true>(flo), ohi = lane_shift<W, D, true>(fhi);
    const uint32_t ulo = lane_shift<W, D, false>(glo), uhi = lane_shift<W, D, false>(ghi);
    // (selects and bitwise tests: no branch)
    const bool up = pos >= D, down = pos + D < W;
    ok = ok & (!up | iv_touch(olo, ohi, flo, fhi)) & (!down | iv_touch(glo, ghi, ulo, uhi));
    flo = up ? min(flo, olo) : flo;
    fhi = up ? max(fhi, ohi) : fhi;
    glo = down ? min(glo, ulo) : glo;
    ghi = down ? max(ghi, uhi) : ghi;
  }
}

// One step of the inclusive scans of the lane totals: prefix f from the
// lane D before, suffix g from the lane D after.
template <int W, int D>
__device__ __forceinline__ void ivl_scan_step(Ivl2 &f, Ivl2 &g, int pos) {
  if constexpr (D < W) {
    const Ivl2 o = ivl_shift<W, D, true>(f);
    if (pos >= D) f = ivl_join(o, f);
    const Ivl2 u = ivl_shift<W, D, false>(g);
    if (pos + D < W) g = ivl_join(g, u);
  }
}

// Each lane takes K (SUMM_ROWS_PER_LANE unless given) consecutive rows: their unions are
// joined serially into the lane's own prefixes and suffixes, the lane totals
// are scanned across the block's LANES lanes (both directions), and each
// row's summary is the join of the lanes before (after) it with its own
// prefix (suffix).  Joins of summaries in a different order may give
// "unknown" where another order would not, never a different interval.
// Summaries of one lane's K rows r0 .. r0+K-1: x[k] is row
// k's first range ((EMPTY, 0) for an empty row), `single` says whether every
// row is exactly one range, and more(k, v) joins row k's further ranges
// into v (only called on the general path).
template <int LANES, int K = SUMM_ROWS_PER_LANE, typename More>
__device__ __forceinline__ void summary_lane(const uint2 (&x)[K], bool single,
                                             More more, uint64_t r0, uint64_t nrows, int pos,
                                             uint2 *__restrict__ pre, uint2 *__restrict__ suf,
                                             uint4 *__restrict__ pre2, uint4 *__restrict__ suf2) {
  const Ivl2 none{SUMM_EMPTY, 0u, SUMM_EMPTY, 0u};
  // Fast path: every row one range and every join below one of two
  // touching intervals, so each partial union is a single interval and
  // the two-piece joins would give the same.  A wave with any other case
  // takes the general path.
  {
    bool ok = single;
    uint32_t plo[K], phi[K], slo[K], shi[K];
    plo[0] = x[0].x;
    phi[0] = x[0].y;
#pragma unroll
    for (int k = 1; k < K; ++k) {
      ok = ok & iv_touch(plo[k - 1], phi[k - 1], x[k].x, x[k].y);
      plo[k] = min(plo[k - 1], x[k].x);
      phi[k] = max(phi[k - 1], x[k].y);
    }
    slo[K - 1] = x[K - 1].x;
    shi[K - 1] = x[K - 1].y;
#pragma unroll
    for (int k = K - 2; k >= 0; --k) {
      ok = ok & iv_touch(x[k].x, x[k].y, slo[k + 1], shi[k + 1]);
      slo[k] = min(slo[k + 1], x[k].x);
      shi[k] = max(shi[k + 1], x[k].y);
    }
    uint32_t flo = plo[K - 1], fhi = phi[K - 1], glo = flo, ghi = fhi;
    iv_scan_step<LANES, 1>(flo, fhi, glo, ghi, ok, pos);
    iv_scan_step<LANES, 2>(flo, fhi, glo, ghi, ok, pos);
    iv_scan_step<LANES, 4>(flo, fhi, glo, ghi, ok, pos);
    iv_scan_step<LANES, 8>(flo, fhi, glo, ghi, ok, pos);
    iv_scan_step<LANES, 16>(flo, fhi, glo, ghi, ok, pos);
    iv_scan_step<LANES, 32>(flo, fhi, glo, ghi, ok, pos);
    const uint32_t blo = lane_shift<LANES, 1, true>(flo), bhi = lane_shift<LANES, 1, true>(fhi);
    const uint32_t alo = lane_shift<LANES, 1, false>(glo), ahi = lane_shift<LANES, 1, false>(ghi);
    const bool first = pos == 0, last = pos == LANES - 1;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      ok = ok & (first | iv_touch(blo, bhi, plo[k], phi[k])) &
           (last | iv_touch(slo[k], shi[k], alo, ahi));
      plo[k] = first ? plo[k] : min(plo[k], blo);
      phi[k] = first ? phi[k] : max(phi[k], bhi);
      slo[k] = last ? slo[k] : min(slo[k], alo);
      shi[k] = last ? shi[k] : max(shi[k], ahi);
    }
    if (__all(ok)) {  // (every row is then below nrows: rows past it are empty)
#pragma unroll
      for (int k = 0; k < K; ++k) {
        pre[r0 + k] = make_uint2(plo[k], phi[k]);
        suf[r0 + k] = make_uint2(slo[k], shi[k]);
      }
      return;
    }
  }
  Ivl2 v[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    v[k] = Ivl2{x[k].x, x[k].y, SUMM_EMPTY, 0u};
    more(k, v[k]);
  }
  Ivl2 lp[K], ls[K];
  lp[0] = v[0];
#pragma unroll
  for (int k = 1; k < K; ++k) lp[k] = ivl_join(lp[k - 1], v[k]);
  ls[K - 1] = v[K - 1];
#pragma unroll
  for (int k = K - 2; k >= 0; --k) ls[k] = ivl_join(v[k], ls[k + 1]);
  // inclusive scans of the lane totals, then shifted by one lane
  Ivl2 f = lp[K - 1], g = lp[K - 1];
  ivl_scan_step<LANES, 1>(f, g, pos);
  ivl_scan_step<LANES, 2>(f, g, pos);
  ivl_scan_step<LANES, 4>(f, g, pos);
  ivl_scan_step<LANES, 8>(f, g, pos);
  ivl_scan_step<LANES, 16>(f, g, pos);
  ivl_scan_step<LANES, 32>(f, g, pos);
  Ivl2 before = ivl_shift<LANES, 1, true>(f), after = ivl_shift<LANES, 1, false>(g);
  if (pos == 0) before = none;
  if (pos == LANES - 1) after = none;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const uint64_t r = r0 + k;
    if (r < nrows) {
      ivl_store(ivl_join(before, lp[k]), pre + r, pre2 + r);
      ivl_store(ivl_join(ls[k], after), suf + r, suf2 + r);
    }
  }
}

// The first range of each of K consecutive rows with bounds rb[0 .. K]
// (empty: (empty, 0)), all loaded together under one branch; an empty
// row's load repeats a non-empty row's and is dropped.
template <int K, typename S>
__device__ __forceinline__ void load_first_ranges(const uint32_t *__restrict__ en,
                                                  const S (&rb)[K + 1], uint2 (&x)[K],
                                                  uint32_t empty) {
  S any_at = rb[K];
#pragma unroll
  for (int k = K - 1; k >= 0; --k) any_at = rb[k + 1] > rb[k] ? rb[k] : any_at;
#pragma unroll
  for (int k = 0; k < K; ++k) x[k] = make_uint2(empty, 0u);
  if (rb[K] > rb[0]) {
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] = ld_range(en, rb[k + 1] > rb[k] ? rb[k] : any_at);
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (rb[k + 1] <= rb[k]) x[k] = make_uint2(empty, 0u);
  }
}

template <typename S, int LANES>
__global__ __launch_bounds__(BLOCK) void k_interval_summary(Mat<S> B, uint64_t nrows,
                                                            uint2 *__restrict__ pre,
                                                            uint2 *__restrict__ suf,
                                                            uint4 *__restrict__ pre2,
                                                            uint4 *__restrict__ suf2,
                                                            const unsigned int *__restrict__ status) {
  if (*status) return;
  constexpr int K = SUMM_ROWS_PER_LANE;
  const int pos = threadIdx.x & (LANES - 1);  // the lane's place in its block
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK * K; base < nrows;
       base += (uint64_t)gridDim.x * BLOCK * K) {
    const uint64_t r0 = base + (uint64_t)threadIdx.x * K;
    // the rows' bounds, then every row's first range, each set loaded at
    // once (rows past the end are empty); further ranges of a row (rare)
    // are read one by one
    S rb[K + 1];
#pragma unroll
    for (int k = 0; k <= K; ++k) rb[k] = B.rm[min(r0 + k, nrows)];
    uint2 x[K];
    load_first_ranges<K>(B.en, rb, x, SUMM_EMPTY);
    bool single = true;
#pragma unroll
    for (int k = 0; k < K; ++k) single = single && rb[k + 1] - rb[k] == 2;
    summary_lane<LANES>(
        x, single,
        [&](int k, Ivl2 &v) {
          for (S p = rb[k] + 2; p < rb[k + 1] && !ivl_unknown(v); p += 2) {
            const uint2 y = ld_range(B.en, p);
            v = ivl_join(v, Ivl2{y.x, y.y, SUMM_EMPTY, 0u});
          }
        },
        r0, nrows, pos, pre, suf, pre2, suf2);
  }
}

// Fold the B rows [s, e] of an A range that crosses a block boundary through
// the summaries: rows [s, h1) by the suffix summary of s, each whole block by
// the prefix summary of its last row, rows [t0, e] by the prefix summary of e.
// The summaries are loaded SUMM_BATCH at a time before any is used, so a wide
// range costs a few round trips, not one per block.  An unknown summary falls
// back to that piece's ranges.
constexpr int SUMM_BATCH = 6;  // the items of a 151-row slice over 32-row blocks

// rows(r0, r1) folds B rows [r0, r1) from their ranges (the fallback).
template <typename Rows>
__device__ __forceinline__ void fold_summarised(const uint2 *__restrict__ pre,
                                                const uint2 *__restrict__ suf,
                                                const uint4 *__restrict__ pre2,
                                                const uint4 *__restrict__ suf2, uint32_t s,
                                                uint32_t e, uint32_t sh, Rows rows,
                                                uint32_t &cl, uint32_t &ch, Runs &r, bool &ok,
                                                uint32_t &gathered) {
  // row indices are column indices of A, at most 0xFFFFFFFD: 32-bit math
  const uint32_t W = 1u << sh;  // rows per summary block
  const uint32_t h1 = (s + W - 1) >> sh << sh;
  const uint32_t t0 = (e + 1) >> sh << sh;
  const uint32_t head = s < h1 ? 1 : 0, tail = t0 <= e ? 1 : 0;
  const uint32_t nitems = head + ((t0 - h1) >> sh) + tail;  // at most 32 (the caller's check)
  uint32_t unknown = 0;  // pieces whose summary is unknown, read after the summaries
  uint32_t two = 0;      // pieces whose summary has two pieces (in the side entries)
  for (uint32_t j0 = 0; ok && j0 < nitems; j0 += SUMM_BATCH) {
    uint2 sm[SUMM_BATCH];
#pragma unroll
    for (int q = 0; q < SUMM_BATCH; ++q) {
      const uint32_t j = j0 + q;
      sm[q] = make_uint2(SUMM_EMPTY, 0u);
      if (j < nitems) {
        if (j < head) sm[q] = suf[s];
        else if (tail && j + 1 == nitems) sm[q] = pre[e];
        else sm[q] = pre[h1 + ((j - head) << sh) + W - 1];
      }
    }
#pragma unroll
    for (int q = 0; q < SUMM_BATCH; ++q) {
      if (j0 + q >= nitems) continue;
      if (sm[q].x <= sm[q].y) {
        fold(sm[q].x, sm[q].y, cl, ch, r, ok);
        gathered += 1;
      } else if (sm[q].x == SUMM_TWO) {
        two |= 1u << (j0 + q);
      } else if (sm[q].x != SUMM_EMPTY) {
        unknown |= 1u << (j0 + q);
      }
    }
  }
  // two-piece summaries: one more read each
  while (ok && two) {
    const uint32_t j = (uint32_t)__builtin_ctz(two);
    two &= two - 1;
    const uint4 t = j < head ? suf2[s] : (tail && j + 1 == nitems) ? pre2[e]
                                                                  : pre2[h1 + ((j - head) << sh) + W - 1];
    fold(t.x, t.y, cl, ch, r, ok);
    fold(t.z, t.w, cl, ch, r, ok);
    gathered += 1;
  }
  // unknown pieces: their own ranges (the union does not depend on the order)
  while (ok && unknown) {
    const uint32_t j = (uint32_t)__builtin_ctz(unknown);
    unknown &= unknown - 1;
    const uint32_t r0 = j < head ? s : (tail && j + 1 == nitems) ? t0 : h1 + ((j - head) << sh);
    const uint32_t r1 = j < head ? h1 : (tail && j + 1 == nitems) ? e + 1 : r0 + W;
    rows(r0, r1);
  }
}

template <typename S, bool SUMM>
__global__ __launch_bounds__(BLOCK) void k_spgemm_lane(
    Mat<S> A, Mat<S> B, uint64_t nrows, uint32_t KS, uint32_t lane_max, uint32_t mid_max,
    uint32_t *__restrict__ cnt, uint32_t *__restrict__ stage,
    uint32_t *__restrict__ ovf_list, uint32_t *__restrict__ mid_list,
    uint32_t *__restrict__ big_list, OpCounters *__restrict__ ctr,
    const unsigned int *__restrict__ status, const uint2 *__restrict__ pre,
    const uint2 *__restrict__ suf, const uint4 *__restrict__ pre2,
    const uint4 *__restrict__ suf2, uint32_t summ_min, uint32_t summ_shift, uint32_t plus_i,
    unsigned long long *__restrict__ parts) {
  if (*status) return;  // an earlier step overflowed: operands are not valid
  unsigned long long nnz_acc = 0;
  uint32_t slice_acc = 0;  // per lane: slices of a few thousand rows at most
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  // Three-stage software pipeline over this lane's rows (one visit apart):
  // while the current row is folded, the next row's first A range's B row
  // bounds, the first A range of the row after, and the A row bounds of the
  // third are loaded.  A row's own chain is then a single round trip to its
  // B ranges (or summaries).
  uint64_t row = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  S c_a0 = 0, c_a1 = 0, c_bs = 0, c_be = 0;  // current row
  S p1_a0 = 0, p1_a1 = 0;                    // next row
  S p2_a0 = 0, p2_a1 = 0;                    // the one after
  uint2 c_first = make_uint2(1u, 0u), p1_first = make_uint2(1u, 0u);
  if (row < nrows) {
    c_a0 = A.rm[row];
    c_a1 = A.rm[row + 1];
  }
  if (row + stride < nrows) {
    p1_a0 = A.rm[row + stride];
    p1_a1 = A.rm[row + stride + 1];
  }
  if (row + 2 * stride < nrows) {
    p2_a0 = A.rm[row + 2 * stride];
    p2_a1 = A.rm[row + 2 * stride + 1];
  }
  if (c_a0 < c_a1) c_first = ld_range(A.en, c_a0);
  if (p1_a0 < p1_a1) p1_first = ld_range(A.en, p1_a0);
  if (c_a0 < c_a1) {
    c_bs = B.rm[c_first.x];
    c_be = B.rm[(uint64_t)c_first.y + 1];
  }
  need(c_bs);
  need(c_be);
  need(p1_first.x);
  need(p1_first.y);
  need(p2_a0);
  need(p2_a1);
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < nrows; base += stride) {
    row = base + threadIdx.x;
    const S a0 = c_a0, a1 = c_a1;
    const uint2 first = c_first;
    const S first_bs = c_bs, first_be = c_be;
    // advance the pipeline: next row's B bounds, the first range of the row
    // after, the A bounds of the third
    c_a0 = p1_a0;
    c_a1 = p1_a1;
    c_first = p1_first;
    c_bs = c_be = 0;
    if (c_a0 < c_a1) {
      c_bs = B.rm[c_first.x];
      c_be = B.rm[(uint64_t)c_first.y + 1];
    }
    p1_a0 = p2_a0;
    p1_a1 = p2_a1;
    p1_first = make_uint2(1u, 0u);
    if (p1_a0 < p1_a1) p1_first = ld_range(A.en, p1_a0);
    p2_a0 = p2_a1 = 0;
    if (row + 3 * stride < nrows) {
      p2_a0 = A.rm[row + 3 * stride];
      p2_a1 = A.rm[row + 3 * stride + 1];
    }
    int cls = 0;  // 0 done here (or no row), 1 -> medium groups, 2 -> 64-lane groups
    bool ovf = false;
    const bool live = row < nrows;
    uint64_t len = 0;  // B-slice ranges of the row, the tier choice
    bool ok = true;
    Runs r;
#pragma unroll
    for (int t = 0; t < LANE_SLOTS; ++t) {
      r.lo[t] = 0xFFFFFFFFu;
      r.hi[t] = 0u;
    }
    // The running interval [cl, ch] absorbs ranges that touch it (the
    // common case in banded slices); any other range moves it into the
    // slots and starts a new running interval.
    uint32_t cl = 0xFFFFFFFFu, ch = 0u;
    if (live) {
      for (uint64_t ai = a0; ai < a1; ai += 2) {
        uint2 ar = first;  // the first range and its B bounds: from the pipeline
        S bs = first_bs, be = first_be;
        if (ai != a0) {
          ar = ld_range(A.en, ai);
          bs = B.rm[ar.x];
          be = B.rm[(uint64_t)ar.y + 1];
          need(bs);
          need(be);
        }
        const uint64_t s = ar.x, e = ar.y;
        len += (be - bs) >> 1;
        if (len > lane_max) ok = false;
        if (!ok) continue;  // only the slice size is still needed
        if (plus_i) fold((uint32_t)s, (uint32_t)e, cl, ch, r, ok);  // C = A (B + I)
        // summaries for a range that crosses a block boundary and touches at
        // most 64 blocks
        if (SUMM && (s >> summ_shift) != (e >> summ_shift) && e - s + 1 >= summ_min &&
            (e >> summ_shift) - (s >> summ_shift) < 31) {
          fold_summarised(
              pre, suf, pre2, suf2, ar.x, ar.y, summ_shift,
              [&](uint32_t r0, uint32_t r1) {
                fold_segment<S, false>(B.en, r0 == ar.x ? bs : B.rm[r0],
                                       r1 == ar.y + 1 ? be : B.rm[r1], cl, ch, r, ok, slice_acc);
              },
              cl, ch, r, ok, slice_acc);
        } else {
          // the 64-bit summarised variant keeps 4 waves per SIMD only
          // without the fast batch
          fold_segment<S, !(SUMM && sizeof(S) == 8)>(B.en, bs, be, cl, ch, r, ok, slice_acc);
        }
      }
    }
    // the prefetches issued above have landed with the slice loads: take
    // them now, before this row's stores are issued
    need(c_bs);
    need(c_be);
    need(p1_first.x);
    need(p1_first.y);
    need(p2_a0);
    need(p2_a1);
    // live slots form a prefix (insert fills the first free slot, compact
    // moves emptied ones last): with slot 0 empty the row is [cl, ch] alone
    const bool single = ok && r.lo[0] > r.hi[0];
    if (live && single) {
      uint32_t idx = 0;
      if (cl <= ch) {
        nnz_acc += (unsigned long long)(ch - cl) + 1;
        if (0 < KS) reinterpret_cast<uint2 *>(stage)[row] = make_uint2(cl, ch);
        idx = 2;
      }
      cnt[row] = idx;
      ovf = idx > KS;
    } else if (live) {
      if (ok && cl <= ch && !insert(r, cl, ch)) {
        compact(r);
        ok = insert(r, cl, ch);
      }
      if (!ok) {
        cls = len <= mid_max ? 1 : 2;
      } else {
        compact(r);
        uint32_t idx = 0;
#pragma unroll
        for (int t = 0; t < LANE_SLOTS; ++t) {
          if (r.lo[t] <= r.hi[t]) {
            nnz_acc += (unsigned long long)(r.hi[t] - r.lo[t]) + 1;
            if (idx < KS)  // KS is even: the whole range fits
              reinterpret_cast<uint2 *>(stage)[(uint64_t)(idx >> 1) * nrows + row] =
                  make_uint2(r.lo[t], r.hi[t]);
            idx += 2;
          }
        }
        cnt[row] = idx;
        ovf = idx > KS;
      }
    }
    wave_append(ovf, (uint32_t)row, ovf_list, &ctr->overflow);
    wave_append(cls == 1, (uint32_t)row, mid_list, &ctr->n_mid);
    wave_append(cls == 2, (uint32_t)row, big_list, &ctr->n_big);
  }
  // per-block partial sums, added to the counters by the row-pointer scan:
  // one atomic per block on the same two words would serialise at the end
  // of the kernel, when every block finishes at once
  nnz_acc = block_sum_u64(nnz_acc);
  const unsigned long long slice_sum = block_sum_u64(slice_acc);
  if (threadIdx.x == 0) {
    parts[2 * blockIdx.x] = nnz_acc;
    parts[2 * blockIdx.x + 1] = slice_sum;
  }
}

// ---------------------------------------------------------------------------
// Group tiers: one row of C = A.B by a group of G lanes.  DIRECT: write the
// row's entries at c_en[c_base ...]; otherwise count, stage and list rows
// whose output exceeds the stage.
template <typename S, int G, bool DIRECT>
__device__ __forceinline__ void spgemm_row(uint64_t row, Mat<S> A, Mat<S> B, uint64_t *lds,
                                           uint32_t W, uint32_t lane, uint32_t KS,
                                           uint64_t nstage, uint32_t *cnt, uint32_t *stage,
                                           uint32_t *ovf_list, unsigned int *ovf_count,
                                           uint32_t *c_en, uint64_t c_base,
                                           unsigned long long &nnz_acc, uint32_t plus_i) {
  const uint64_t a0 = A.rm[row], a1 = A.rm[row + 1];

  // 1. column extent.  A single-segment slice of at most CACHE * G ranges is
  //    kept in registers for the scatter pass.
  uint2 rc[CACHE];
  bool cached = false;
  uint32_t lo = 0xFFFFFFFFu, hi = 0u;
  uint2 a_only = make_uint2(1u, 0u);  // the single A range, for C = A (B + I)
  if (a1 - a0 == 2) {
    const uint32_t s = A.en[a0], e = A.en[a0 + 1];
    if (plus_i) a_only = make_uint2(s, e);
    const uint64_t p0 = B.rm[s], p1 = B.rm[(uint64_t)e + 1];
    if (p1 - p0 <= 2ull * G * CACHE) {
      cached = true;
#pragma unroll
      for (int k = 0; k < CACHE; ++k) {
        const uint64_t p = p0 + 2ull * (lane + (uint32_t)k * G);
        rc[k] = p < p1 ? ld_range(B.en, p) : make_uint2(1u, 0u);  // (1, 0): empty
      }
#pragma unroll
      for (int k = 0; k < CACHE; ++k)
        if (rc[k].x <= rc[k].y) {
          lo = min(lo, rc[k].x);
          hi = max(hi, rc[k].y);
        }
    }
  }
  if (!cached) {
    for (uint64_t ai = a0; ai < a1; ai += 2) {
      const uint32_t s = A.en[ai], e = A.en[ai + 1];
      const uint64_t p1 = B.rm[(uint64_t)e + 1];
#pragma unroll 4
      for (uint64_t p = (uint64_t)B.rm[s] + 2 * lane; p < p1; p += 2 * G) {
        const uint2 r = ld_range(B.en, p);
        lo = min(lo, r.x);
        hi = max(hi, r.y);
      }
    }
  }
  if (plus_i)  // the A ranges themselves are part of the row
    for (uint64_t ai = a0; ai < a1; ai += 2) {
      lo = min(lo, (uint32_t)A.en[ai]);
      hi = max(hi, (uint32_t)A.en[ai + 1]);
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
    if (cached) {
#pragma unroll
      for (int k = 0; k < CACHE; ++k) {
        const uint64_t rs = max((uint64_t)rc[k].x, cb), rf = min((uint64_t)rc[k].y, cl);
        if (rs <= rf) window_set(lds, (uint32_t)(rs - cb), (uint32_t)(rf - cb));
      }
      if (lane == 0 && a_only.x <= a_only.y) {
        const uint64_t rs = max((uint64_t)a_only.x, cb), rf = min((uint64_t)a_only.y, cl);
        if (rs <= rf) window_set(lds, (uint32_t)(rs - cb), (uint32_t)(rf - cb));
      }
    } else {
      for (uint64_t ai = a0; ai < a1; ai += 2) {
        const uint32_t s = A.en[ai], e = A.en[ai + 1];
        if (plus_i && lane == 0) {
          const uint64_t rs = max((uint64_t)s, cb), rf = min((uint64_t)e, cl);
          if (rs <= rf) window_set(lds, (uint32_t)(rs - cb), (uint32_t)(rf - cb));
        }
        const uint64_t p1 = B.rm[(uint64_t)e + 1];
        for (uint64_t p = (uint64_t)B.rm[s] + 2 * lane; p < p1; p += 2 * G) {
          const uint2 r = ld_range(B.en, p);
          const uint64_t rs = max((uint64_t)r.x, cb), rf = min((uint64_t)r.y, cl);
          if (rs <= rf) window_set(lds, (uint32_t)(rs - cb), (uint32_t)(rf - cb));
        }
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
            stage[((uint64_t)(idx >> 1) * nstage + row) * 2 + (idx & 1u)] = val;
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

// Rows rows[0 .. n) with n = *n_dev (counting) or n_host (direct).  Groups
// take RPG consecutive list entries at a time.
template <typename S, int G, bool DIRECT>
__global__ __launch_bounds__(BLOCK) void k_spgemm_group(
    Mat<S> A, Mat<S> B, uint32_t W, uint32_t KS, uint32_t RPG, uint64_t nstage,
    const uint32_t *__restrict__ rows, const unsigned int *__restrict__ n_dev, uint32_t n_host,
    uint32_t *__restrict__ cnt, uint32_t *__restrict__ stage, uint32_t *__restrict__ ovf_list,
    OpCounters *__restrict__ ctr, const S *__restrict__ c_rm, uint32_t *__restrict__ c_en,
    const unsigned int *__restrict__ status, uint32_t plus_i) {
  extern __shared__ uint64_t lds_all[];
  if (*status) return;
  constexpr uint32_t GPB = BLOCK / G;
  const uint32_t grp = threadIdx.x / G, lane = threadIdx.x % G;
  uint64_t *lds = lds_all + (size_t)grp * W;
  const uint64_t nl = n_dev ? *n_dev : n_host;
  unsigned long long nnz_acc = 0;
  const uint64_t groups = (uint64_t)gridDim.x * GPB;
  const uint64_t gid = (uint64_t)xcd_tile(blockIdx.x, gridDim.x) * GPB + grp;
  for (uint64_t i0 = gid * RPG; i0 < nl; i0 += groups * RPG) {
    const uint64_t i1 = min(i0 + RPG, nl);
    for (uint64_t i = uniform_u64<G>(i0); i < i1; ++i) {
      const uint64_t row = uniform_u64<G>(rows[i]);
      spgemm_row<S, G, DIRECT>(row, A, B, lds, W, lane, KS, nstage, cnt, stage, ovf_list,
                               DIRECT ? nullptr : &ctr->overflow, c_en,
                               DIRECT ? (uint64_t)c_rm[row] : 0, nnz_acc, plus_i);
    }
  }
  if (!DIRECT) block_add_u64(nnz_acc, &ctr->nnz);
}

// Copy staged rows (count <= KS) into C; stage holds range k of row r at
// uint2 index k * nrows + r.
template <typename S>
__global__ __launch_bounds__(BLOCK) void k_spgemm_emit_staged(
    uint64_t nrows, const S *__restrict__ c_rm, const uint32_t *__restrict__ stage,
    uint32_t KS, uint32_t *__restrict__ c_en, const unsigned int *__restrict__ status) {
  if (*status) return;
  for (uint64_t row = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; row < nrows;
       row += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t b = c_rm[row], e = c_rm[row + 1];
    const uint64_t c = e - b;
    if (c <= KS)
      for (uint64_t k = 0; k < c; k += 2)
        *reinterpret_cast<uint2 *>(c_en + b + k) =
            reinterpret_cast<const uint2 *>(stage)[(k >> 1) * nrows + row];
  }
}

// ---------------------------------------------------------------------------
// Row-pointer scan: tiles of BLOCK * ITEMS counts.
constexpr int SCAN_ITEMS = 8;
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

// ---------------------------------------------------------------------------
// Row-pointer scan in two launches with no inter-block waiting: (1) each tile
// of SCAN_TILE rows sums its counts; (2) each tile adds up the sums of all
// tiles before it (a block reduction over at most a few thousand words, from
// L2), scans its own rows and writes them.  Loads and stores of counts and
// row pointers go through LDS so that global accesses are coalesced; the
// per-thread run of SCAN_ITEMS consecutive rows is read from a padded LDS
// layout (one pad word per 16) without bank conflicts.
// EMIT: rows whose count fits the stage are copied into C here (one uint2
// per range, rows across lanes); a tile whose entries
// would pass `cap` writes none of them (the total then flags bit 0).

// Per-tile sums of the counts.  The last block also adds the counting
// kernel's per-block (nnz, slice) partial sums, if any, to the counters.
__global__ __launch_bounds__(BLOCK) void k_tile_sums(const uint32_t *__restrict__ cnt, uint64_t n,
                                                     unsigned long long *__restrict__ tsum,
                                                     const unsigned long long *__restrict__ parts,
                                                     uint32_t nparts, OpCounters *__restrict__ ctr) {
  const uint64_t t0 = (uint64_t)blockIdx.x * SCAN_TILE;
  unsigned long long v = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const uint64_t i = t0 + (uint64_t)k * BLOCK + threadIdx.x;
    if (i < n) v += cnt[i];
  }
  v = block_sum_u64(v);
  if (threadIdx.x == 0) tsum[blockIdx.x] = v;
  if (nparts && blockIdx.x + 1 == gridDim.x) {
    unsigned long long a = 0, b = 0;
    for (uint32_t j = threadIdx.x; j < nparts; j += BLOCK) {
      a += parts[2 * j];
      b += parts[2 * j + 1];
    }
    a = block_sum_u64(a);
    b = block_sum_u64(b);
    if (threadIdx.x == 0) {
      atomicAdd(&ctr->nnz, a);
      atomicAdd(&ctr->slice, b);
    }
  }
}

// With many tiles, one block turns the tile sums into exclusive prefixes in
// place (each of its threads sums a contiguous chunk, the chunk sums are
// scanned, the chunks rewritten), so that a tile reads its prefix instead of
// adding up all earlier tiles.
constexpr int SPINE_THREADS = 1024;
constexpr uint64_t SPINE_MIN_TILES = 4096;

__global__ __launch_bounds__(SPINE_THREADS) void k_scan_spine(unsigned long long *__restrict__ tsum,
                                                              uint64_t ntiles) {
  __shared__ unsigned long long wsum[SPINE_THREADS / 64];
  const uint64_t chunk = (ntiles + SPINE_THREADS - 1) / SPINE_THREADS;
  const uint64_t b0 = min(ntiles, (uint64_t)threadIdx.x * chunk), b1 = min(ntiles, b0 + chunk);
  unsigned long long v = 0;
  for (uint64_t j = b0; j < b1; ++j) v += tsum[j];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  unsigned long long x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long t = __shfl_up(x, d, 64);
    if (lane >= d) x += t;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  unsigned long long off = 0;
  for (int w = 0; w < wid; ++w) off += wsum[w];
  unsigned long long ex = off + x - v;
  for (uint64_t j = b0; j < b1; ++j) {
    const unsigned long long t = tsum[j];
    tsum[j] = ex;
    ex += t;
  }
}

constexpr int SCAN_PAD = SCAN_TILE + SCAN_TILE / 16;
__device__ __forceinline__ uint32_t scan_pad(uint32_t i) { return i + (i >> 4); }

template <typename S, bool EMIT>
__global__ __launch_bounds__(BLOCK) void k_scan_rows(
    const uint32_t *__restrict__ cnt, uint64_t n, S *__restrict__ rm,
    const unsigned long long *__restrict__ tsum, const uint32_t *__restrict__ stage, uint32_t KS,
    uint32_t *__restrict__ c_en, OpCounters *__restrict__ ctr, unsigned long long cap,
    uint32_t sb, unsigned int *__restrict__ status, int prefixed) {
  __shared__ uint32_t s_cnt[SCAN_PAD];
  __shared__ uint32_t s_rel[SCAN_PAD];  // exclusive offsets within the tile
  const uint64_t tile = blockIdx.x;
  const uint64_t t0 = tile * SCAN_TILE;
  const uint2 *stage2 = reinterpret_cast<const uint2 *>(stage);
  // coalesced load of the tile's counts
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const uint32_t i = (uint32_t)k * BLOCK + threadIdx.x;
    s_cnt[scan_pad(i)] = t0 + i < n ? cnt[t0 + i] : 0u;
  }
  // rows before this tile: its prefix from the spine, or the sum of the
  // earlier tiles' sums
  unsigned long long texcl;
  if (prefixed) {
    __syncthreads();  // orders the s_cnt stores before the reads below
    texcl = tsum[tile];
  } else {
    unsigned long long e = 0;
    for (uint64_t j = threadIdx.x; j < tile; j += BLOCK) e += tsum[j];
    texcl = block_sum_u64(e);  // (its barrier also orders s_cnt)
  }
  // this thread's SCAN_ITEMS consecutive rows
  const uint32_t r0 = threadIdx.x * SCAN_ITEMS;
  uint32_t v[SCAN_ITEMS];
  unsigned long long tsum_t = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    v[k] = s_cnt[scan_pad(r0 + k)];
    tsum_t += v[k];
  }
  unsigned long long agg;
  const unsigned long long incl = block_scan_incl(tsum_t, agg);
  const unsigned long long lstart = incl - tsum_t;  // within the tile
  if (tile + 1 == gridDim.x && threadIdx.x == 0) {
    const unsigned long long total = texcl + agg;
    ctr->total = total;
    if (total > cap) atomicOr(status, 1u);
    if (sb == 4 && total > 0xFFFFFFFFull) atomicOr(status, 2u);
  }
  if (tile == 0 && threadIdx.x == 0) rm[0] = 0;
  const bool fits = texcl + agg <= cap && (sb == 8 || texcl + agg <= 0xFFFFFFFFull);
  if (agg > 0xFFFFFFFFull) {
    // a tile of more than 2^32 entries: write from the per-thread runs
    unsigned long long run = texcl + lstart;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
      const uint64_t row = t0 + r0 + k;
      if (row < n) {
        if (EMIT && fits && v[k] <= KS)
          for (uint32_t q = 0; q < v[k]; q += 2)
            *reinterpret_cast<uint2 *>(c_en + run + q) = stage2[(uint64_t)(q >> 1) * n + row];
        run += v[k];
        rm[row + 1] = (S)run;
      }
    }
    return;
  }
  uint32_t rel = (uint32_t)lstart;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    s_rel[scan_pad(r0 + k)] = rel;
    rel += v[k];
  }
  __syncthreads();
  // coalesced stores of the row pointers (and the staged entries)
#pragma unroll 4
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const uint32_t i = (uint32_t)k * BLOCK + threadIdx.x;
    const uint64_t row = t0 + i;
    if (row < n) {
      const uint32_t c = s_cnt[scan_pad(i)];
      const unsigned long long ex = texcl + s_rel[scan_pad(i)];
      rm[row + 1] = (S)(ex + c);
      if (EMIT && fits && c <= KS)
        for (uint32_t q = 0; q < c; q += 2)
          *reinterpret_cast<uint2 *>(c_en + ex + q) = stage2[(uint64_t)(q >> 1) * n + row];
    }
  }
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
                                                       uint32_t *__restrict__ cnt,
                                                       const unsigned int *__restrict__ status) {
  if (*status) return;
  for (uint64_t row = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; row < n;
       row += (uint64_t)gridDim.x * BLOCK)
    cnt[row] = (uint32_t)spadd_row<S, false>(A.en, A.rm[row], A.rm[row + 1], B.en,
                                             B.rm[row], B.rm[row + 1], nullptr);
}

template <typename S>
__global__ __launch_bounds__(BLOCK) void k_spadd_write(Mat<S> A, Mat<S> B, uint64_t n,
                                                       const S *__restrict__ c_rm,
                                                       uint32_t *__restrict__ c_en,
                                                       const unsigned int *__restrict__ status) {
  if (*status) return;
  for (uint64_t row = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; row < n;
       row += (uint64_t)gridDim.x * BLOCK)
    spadd_row<S, true>(A.en, A.rm[row], A.rm[row + 1], B.en, B.rm[row], B.rm[row + 1],
                       c_en + c_rm[row]);
}

// nnz = sum of (e - s + 1) over all ranges (crs_matrix.hpp:175-189).
// nr = ctr->total / 2 ranges (the entry count written by the scan).
__global__ __launch_bounds__(BLOCK) void k_range_nnz(const uint2 *__restrict__ r,
                                                     OpCounters *__restrict__ ctr,
                                                     const unsigned int *__restrict__ status) {
  if (*status) return;
  const uint64_t nr = ctr->total / 2;
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

// ---------------------------------------------------------------------------
// Serving: one launch answers a batch of (index, i, j) queries against a
// table of resident indexes (each with its own rowmap width).
__global__ __launch_bounds__(BLOCK) void k_query_multi(const IndexDesc *__restrict__ table,
                                                       uint32_t ntable,
                                                       const uint32_t *__restrict__ qidx,
                                                       const uint32_t *__restrict__ qi,
                                                       const uint32_t *__restrict__ qj, uint32_t n,
                                                       uint8_t *__restrict__ out) {
  for (uint32_t t = blockIdx.x * BLOCK + threadIdx.x; t < n; t += gridDim.x * BLOCK) {
    const uint3 v = make_uint3(qidx[t], qi[t], qj[t]);  // x: index id, y: i, z: j
    if (v.x >= ntable) {
      out[t] = 2;
      continue;
    }
    const IndexDesc d = table[v.x];
    if (v.y >= d.nrows || v.z >= d.ncols) {
      out[t] = 2;
      continue;
    }
    uint64_t b, e;
    if (d.sb == 4) {
      b = static_cast<const uint32_t *>(d.rowmap)[v.y];
      e = static_cast<const uint32_t *>(d.rowmap)[v.y + 1];
    } else {
      b = static_cast<const uint64_t *>(d.rowmap)[v.y];
      e = static_cast<const uint64_t *>(d.rowmap)[v.y + 1];
    }
    uint64_t first = b, len = e - b;
    while (len > 0) {
      const uint64_t half = len >> 1;
      if (d.entries[first + half] < v.z) {
        first += half + 1;
        len -= half + 1;
      } else {
        len = half;
      }
    }
    if ((first - b) & 1) --first;
    out[t] = (uint8_t)(first != e && d.entries[first] <= v.z);
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

// Sets bit 2 of *status (4) when a row's ranges are not canonical: each range
// must end at least two columns before the next one starts
// (e_k + 1 < s_(k+1)).  Rows of one range are canonical.
template <typename S>
__global__ __launch_bounds__(BLOCK) void k_check_canonical(Mat<S> M, uint64_t nrows,
                                                           unsigned int *__restrict__ status) {
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  bool canonical = true;
  for (uint64_t row = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; row < nrows; row += stride) {
    const uint64_t b = M.rm[row], e = M.rm[row + 1];
    if (e - b <= 2) continue;
    uint32_t prev_end = M.en[b + 1];
    for (uint64_t k = b + 2; k < e; k += 2) {
      const uint2 v = ld_range(M.en, k);
      canonical = canonical && (uint64_t)prev_end + 1 < (uint64_t)v.x;
      prev_end = v.y;
    }
  }
  if (!canonical) atomicOr(status, 4u);
}

// The reference's accumulator window for non-canonical right operands.  Its
// TeamSequential HBitVector<2048> product (range_sparse.hpp:1192-1242, the
// same scatter in the numeric phase :1349-1399) keeps, per row i of C, a
// window of 64-bit words [c_min, c_max): first the 2048-bit L1 band placed
// around i (hbitvector.hpp:278-299), then, before each B row j it visits (A's
// ranges in their stored order, rows in ascending order), grown down to the
// word of B row j's first range start and up to the word after its last
// range end.  Bits of B row j outside the window at that moment are lost (a
// later growth clears their words; the gather reads only the window).  For
// canonical B rows nothing is ever outside; for others the product is the
// union of each B range clipped to its window.  This kernel writes those
// clipped pieces (counts first, then, with WRITE, the pieces at the row
// pointers), so that I . P (set union) gives the reference's product.
__device__ __forceinline__ uint64_t hbv_l1_begin_bit(uint64_t centre, uint64_t ncols) {
  constexpr uint64_t L1_BITS = 2048, L1_WORDS = L1_BITS / 64;
  const uint64_t x_size = max((ncols + 63) & ~63ull, L1_BITS);
  const uint64_t nb = x_size >> 6, ctr = centre >> 6;
  if (L1_WORDS >= nb) return 0;
  const uint64_t lflank = (L1_WORDS >> 1) - 1, rfit = nb - L1_WORDS;
  const uint64_t pb = ctr > lflank ? ctr - lflank : 0;
  return min(rfit, pb) << 6;
}

template <typename S, bool WRITE>
__global__ __launch_bounds__(BLOCK) void k_window_pieces(Mat<S> A, Mat<S> B, uint64_t nrows,
                                                         uint64_t b_ncols,
                                                         uint32_t *__restrict__ cnt,
                                                         const S *__restrict__ p_rm,
                                                         uint32_t *__restrict__ p_en,
                                                         const unsigned int *__restrict__ status) {
  if (*status) return;
  for (uint64_t row = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; row < nrows;
       row += (uint64_t)gridDim.x * BLOCK) {
    uint64_t c_min = hbv_l1_begin_bit(row, b_ncols), c_max = c_min + 2048;
    uint64_t out = WRITE ? (uint64_t)p_rm[row] : 0, k = 0;
    for (uint64_t ai = A.rm[row]; ai < (uint64_t)A.rm[row + 1]; ai += 2) {
      const uint32_t s = A.en[ai], e = A.en[ai + 1];
      for (uint64_t j = s; j <= e; ++j) {
        const uint64_t bi = B.rm[j], be = B.rm[j + 1];
        if (bi == be) continue;
        const uint64_t b_min = B.en[bi], b_max = (uint64_t)B.en[be - 1] + 1;
        if (b_min < c_min) c_min = b_min & ~63ull;
        if (c_max < b_max) c_max = (b_max + 63) & ~63ull;
        for (uint64_t q = bi; q < be; q += 2) {
          const uint64_t x = max((uint64_t)B.en[q], c_min);
          const uint64_t y = min((uint64_t)B.en[q + 1], c_max - 1);
          if (x > y) continue;
          if (WRITE) {
            p_en[out + k] = (uint32_t)x;
            p_en[out + k + 1] = (uint32_t)y;
          }
          k += 2;
        }
      }
    }
    if (!WRITE) cnt[row] = (uint32_t)min(k, (uint64_t)0xFFFFFFFFu);
  }
}

// ---------------------------------------------------------------------------
// Block-diagonal assembly of independent components and the split back.
// Component rows [row0, row0 + n) of the assembly hold the component's rows
// with every column shifted by col0 and every rowmap offset by ent0.
template <typename SI>
__global__ __launch_bounds__(BLOCK) void k_concat(const SI *__restrict__ rm, const uint32_t *__restrict__ en,
                                                  uint64_t n, uint64_t n_entries, uint64_t row0,
                                                  uint64_t ent0, uint32_t col0,
                                                  uint64_t *__restrict__ out_rm,
                                                  uint32_t *__restrict__ out_en, int write_last) {
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  const uint64_t t0 = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  for (uint64_t i = t0; i < n + (write_last ? 1 : 0); i += stride) out_rm[row0 + i] = ent0 + rm[i];
  for (uint64_t k = t0; k < n_entries; k += stride) out_en[ent0 + k] = en[k] + col0;
}

template <typename SO>
__global__ __launch_bounds__(BLOCK) void k_extract(const uint64_t *__restrict__ rm,
                                                   const uint32_t *__restrict__ en, uint64_t row0,
                                                   uint64_t n, uint32_t col0,
                                                   SO *__restrict__ out_rm,
                                                   uint32_t *__restrict__ out_en) {
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  const uint64_t t0 = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  const uint64_t base = rm[row0], len = rm[row0 + n] - base;
  for (uint64_t i = t0; i <= n; i += stride) out_rm[i] = (SO)(rm[row0 + i] - base);
  for (uint64_t k = t0; k < len; k += stride) out_en[k] = en[base + k] - col0;
}

// out[c] = rowmap[bounds[c]] for c <= ncomp (component entry offsets).
__global__ void k_gather_bounds(const uint64_t *__restrict__ rm, const uint64_t *__restrict__ bounds,
                                uint32_t m, unsigned long long *__restrict__ out) {
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < m; c += gridDim.x * blockDim.x)
    out[c] = rm[bounds[c]];
}

// nnz of the ranges entries[e0 .. e1) (sum of e - s + 1) into *out.
__global__ __launch_bounds__(BLOCK) void k_slice_nnz(const uint2 *__restrict__ r, uint64_t nr,
                                                     unsigned long long *__restrict__ out) {
  unsigned long long s = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < nr;
       i += (uint64_t)gridDim.x * BLOCK) {
    const uint2 v = r[i];
    s += (unsigned long long)v.y - v.x + 1;
  }
  block_add_u64(s, out);
}

__global__ __launch_bounds__(BLOCK) void k_fill_u32(uint32_t *__restrict__ p, uint64_t n,
                                                    uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * BLOCK)
    p[i] = v;
}

// Two buffers filled by one launch: index i < n is p[i], the rest q[i - n].
__global__ __launch_bounds__(BLOCK) void k_fill2_u32(uint32_t *__restrict__ p, uint64_t n,
                                                     uint32_t *__restrict__ q, uint64_t m,
                                                     uint32_t v) {
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n + m;
       i += (uint64_t)gridDim.x * BLOCK) {
    if (i < n)
      p[i] = v;
    else
      q[i - n] = v;
  }
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

template <typename S, int G, bool DIRECT>
void launch_group(hipStream_t s, DView a, DView b, const GroupLaunch &g) {
  const size_t lds = (size_t)(BLOCK / G) * g.lds_words * sizeof(uint64_t);
  hipLaunchKernelGGL((k_spgemm_group<S, G, DIRECT>), dim3(g.grid), dim3(BLOCK), lds, s,
                     mat<S>(a), mat<S>(b), g.lds_words, g.stage, g.rows_per_group, g.nstage,
                     g.rows, g.n_dev, g.n_host, g.cnt, g.stage_buf, g.ovf_list, g.ctr,
                     static_cast<const S *>(g.c_rowmap), g.c_entries, g.status, g.plus_i);
}

template <typename S, bool DIRECT>
void group_by_lanes(hipStream_t s, DView a, DView b, const GroupLaunch &g) {
  switch (g.lanes) {
    case 4: launch_group<S, 4, DIRECT>(s, a, b, g); break;
    case 8: launch_group<S, 8, DIRECT>(s, a, b, g); break;
    case 16: launch_group<S, 16, DIRECT>(s, a, b, g); break;
    case 32: launch_group<S, 32, DIRECT>(s, a, b, g); break;
    default: launch_group<S, 64, DIRECT>(s, a, b, g); break;
  }
}

// ---------------------------------------------------------------------------
// Slot format: the intermediate matrices of planned index builds.  One uint2
// per row: (lo, hi) with lo <= hi for a row of that one range, (SLOT_EMPTY, 0)
// for an empty row, (SLOT_MARK, off) for a row of two or more ranges kept in
// the side array: side[off] = (count, 0), then its `count` ranges.  Nearly
// every row of a power of a banded matrix is one range, so a product reads a
// B slice as consecutive slots (no row pointers) and writes each row in place
// (no count pass, row-pointer scan or copy).  Side arrays are allocated with
// SLOT_SIDE_PAD spare entries, so a reader may load a marked row's header and
// a few entries after it at once.
constexpr uint32_t SLOT_EMPTY = 0xFFFFFFFFu, SLOT_MARK = 0xFFFFFFFEu;


struct Slots {
  const uint2 *__restrict__ slot;
  const uint2 *__restrict__ side;
};

// Side entries are reserved in SLOT_SIDE_PARTS partitions of pcap entries,
// each with its own counter (ctr[p]), so that the reservations of a launch do
// not all serialise on one word; a wave (or a group) uses one partition.
struct SideAlloc {
  unsigned long long *ctr;
  unsigned long long pcap;
};

// Reserve `need` entries of a side array for each lane (0: none) with one
// atomic per wave; returns the lane's first entry.  Every lane of the wave
// must call it.
__device__ __forceinline__ unsigned long long wave_alloc(uint32_t need,
                                                         unsigned long long *counter) {
  const unsigned long long m = __ballot(need != 0);
  if (m == 0) return 0;
  const uint32_t lane = __lane_id();
  uint32_t incl = need;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t t = (uint32_t)__shfl_up((int)incl, d, 64);
    if (lane >= (uint32_t)d) incl += t;
  }
  const uint32_t total = (uint32_t)__shfl((int)incl, 63, 64);
  const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(counter, (unsigned long long)total);
  const uint32_t blo = (uint32_t)__shfl((int)(uint32_t)base, (int)leader, 64);
  const uint32_t bhi = (uint32_t)__shfl((int)(uint32_t)(base >> 32), (int)leader, 64);
  return (((unsigned long long)bhi << 32) | blo) + incl - need;
}

// wave_alloc for needs below 8 (a lane-tier row's side entries): the
// exclusive prefix of the lanes' needs from three ballots of their bits
// (no shuffle scan), one atomic per wave, the base read back by a scalar
// broadcast.
__device__ __forceinline__ unsigned long long wave_alloc_small(uint32_t need,
                                                               unsigned long long *counter) {
  const unsigned long long m = __ballot(need != 0);
  if (m == 0) return 0;
  const uint32_t lane = __lane_id();
  const unsigned long long below = (1ull << lane) - 1;
  const unsigned long long b0 = __ballot(need & 1u), b1 = __ballot(need & 2u),
                           b2 = __ballot(need & 4u);
  const uint32_t before = (uint32_t)__popcll(b0 & below) + 2u * (uint32_t)__popcll(b1 & below) +
                          4u * (uint32_t)__popcll(b2 & below);
  const uint32_t total =
      (uint32_t)__popcll(b0) + 2u * (uint32_t)__popcll(b1) + 4u * (uint32_t)__popcll(b2);
  const uint32_t leader = (uint32_t)__ffsll((long long)m) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(counter, (unsigned long long)total);
  const uint32_t blo = (uint32_t)__shfl((int)(uint32_t)base, (int)leader, 64);
  const uint32_t bhi = (uint32_t)__shfl((int)(uint32_t)(base >> 32), (int)leader, 64);
  return (((unsigned long long)bhi << 32) | blo) + before;
}

// Fold the ranges of a marked row (side entry `off`): the header and the
// first SIDE_BATCH ranges in one batch, any further ones after.
constexpr int SIDE_BATCH = 4;
static_assert(SLOT_SIDE_PAD >= SIDE_BATCH, "side padding covers a batch");

__device__ __forceinline__ void fold_side(const uint2 *__restrict__ side, uint32_t off,
                                          uint32_t &cl, uint32_t &ch, Runs &r, bool &ok,
                                          uint32_t &gathered) {
  uint2 v[SIDE_BATCH];
  const uint2 h = side[off];
#pragma unroll
  for (int k = 0; k < SIDE_BATCH; ++k) v[k] = side[off + 1 + k];
  const uint32_t cnt = h.x;
  gathered += cnt;
#pragma unroll
  for (int k = 0; k < SIDE_BATCH; ++k)
    if ((uint32_t)k < cnt) fold(v[k].x, v[k].y, cl, ch, r, ok);
  for (uint32_t q0 = SIDE_BATCH; ok && q0 < cnt; q0 += SIDE_BATCH) {
#pragma unroll
    for (int k = 0; k < SIDE_BATCH; ++k)
      v[k] = q0 + k < cnt ? side[off + 1 + q0 + k] : make_uint2(1u, 0u);
#pragma unroll
    for (int k = 0; k < SIDE_BATCH; ++k) fold(v[k].x, v[k].y, cl, ch, r, ok);
  }
}

// Fold B rows [s, e] (slots) with LANE_BATCH loads in flight; marked rows'
// ranges are read after their batch.
__device__ __forceinline__ void fold_slots(const Slots &B, uint32_t s, uint32_t e, uint32_t &cl,
                                           uint32_t &ch, Runs &r, bool &ok,
                                           uint32_t &gathered) {
  gathered += (e - s) + 1u;
  for (uint32_t j = s; ok && j <= e;) {
    const uint32_t rem = e - j;  // rows after j in the slice (no index past e is formed)
    uint2 v[LANE_BATCH];
#pragma unroll
    for (int k = 0; k < LANE_BATCH; ++k)
      v[k] = (uint32_t)k <= rem ? B.slot[j + k] : make_uint2(SLOT_EMPTY, 0u);
    fold_batch(v, cl, ch, r, ok);  // markers and empty rows are not ranges: skipped
    uint32_t anym = 0xFFFFFFFFu;  // 0 when some slot of the batch is marked (rare)
#pragma unroll
    for (int k = 0; k < LANE_BATCH; ++k) anym = min(anym, v[k].x ^ SLOT_MARK);
    if (anym == 0) {
      uint32_t mk = 0, offs[LANE_BATCH];
#pragma unroll
      for (int k = 0; k < LANE_BATCH; ++k) {
        mk |= v[k].x == SLOT_MARK ? 1u << k : 0u;
        offs[k] = v[k].y;
      }
      while (ok && mk) {
        const uint32_t k = (uint32_t)__builtin_ctz(mk);
        mk &= mk - 1;
        uint32_t off = offs[0];
#pragma unroll
        for (int q = 1; q < LANE_BATCH; ++q) off = k == (uint32_t)q ? offs[q] : off;
        fold_side(B.side, off, cl, ch, r, ok, gathered);
      }
    }
    j = rem >= (uint32_t)LANE_BATCH ? j + LANE_BATCH : e + 1;
    if (j == 0) break;  // (e + 1 wrapped: e was the last ordinal)
  }
}

// Write a lane-tier row (slot) from its compacted runs: one range in place,
// several in the side array (reserved by the caller: off, `need` entries).
__device__ __forceinline__ void slot_store_runs(const Runs &r, uint32_t nruns, uint64_t row,
                                                uint2 *__restrict__ c_slot,
                                                uint2 *__restrict__ c_side,
                                                unsigned long long off) {
  if (nruns == 0) {
    c_slot[row] = make_uint2(SLOT_EMPTY, 0u);
  } else if (nruns == 1) {
    c_slot[row] = make_uint2(r.lo[0], r.hi[0]);
  } else {
    c_side[off] = make_uint2(nruns, 0u);
#pragma unroll
    for (int t = 0; t < LANE_SLOTS; ++t)
      if ((uint32_t)t < nruns) c_side[off + 1 + t] = make_uint2(r.lo[t], r.hi[t]);
    c_slot[row] = make_uint2(SLOT_MARK, (uint32_t)off);
  }
}

// Chains of a slot matrix written by a product.  Rows are taken in chunks of
// 64 (one wave of the product kernel).  Row x "chains" to row x + 1 when both
// are single ranges in the same chunk, lo(x + 1) >= lo(x), hi(x + 1) >= hi(x)
// and lo(x + 1) <= hi(x) + 1: then the union of rows j .. t of a chain is the
// one interval [lo(j), hi(t)].  seg[j] holds the chunk offset (0 .. 63) of the
// last row of the chain starting at row j (its own offset when row j chains
// to nothing: an empty row, a row of several ranges, a row the group tiers
// write, the chunk's last row); chunk[c] holds [lo, hi] of chunk c when its
// 64 rows form one chain, else (1, 0).  Powers of banded graphs chain almost
// everywhere (fewer than 1% of the rows end a chain at the 5e6-bp workload),
// so a B slice of any width costs a product two or three loads per 64 rows.
struct Chains {
  const uint8_t *__restrict__ seg;
  const uint2 *__restrict__ chunk;
};
constexpr uint32_t CHAIN_ROWS = 64;

// The chains of the 64 rows of this wave (one chunk): `one` when the lane's
// row is the single range [lo, hi].  Every lane of the wave calls it.
__device__ __forceinline__ void chain_write(bool one, uint32_t lo, uint32_t hi, bool live,
                                            uint64_t row, uint8_t *__restrict__ seg,
                                            uint2 *__restrict__ chunk) {
  const uint32_t lane = __lane_id();
  const uint32_t nlo = (uint32_t)__shfl_down((int)lo, 1, 64);
  const uint32_t nhi = (uint32_t)__shfl_down((int)hi, 1, 64);
  const bool none = __shfl_down((int)one, 1, 64) != 0;
  // columns are at most 0xFFFFFFFD, so hi + 1 does not wrap
  const bool g = one && none && lane != 63 && nlo >= lo && nhi >= hi && nlo <= hi + 1u;
  const unsigned long long brk = __ballot(!g);  // bit 63 is always set
  const uint32_t t = lane + (uint32_t)__builtin_ctzll(brk >> lane);
  const uint32_t hlast = (uint32_t)__shfl((int)hi, 63, 64);
  if (live) seg[row] = (uint8_t)t;
  if (lane == 0 && live) chunk[row / CHAIN_ROWS] = t == 63 ? make_uint2(lo, hlast) : make_uint2(1u, 0u);
}

// The chains of a slot matrix, from its slots (a row is single when its slot
// is one range): one row per lane, a wave per 64-row chunk.  Run after every
// kernel of a product has written its rows.
__global__ __launch_bounds__(BLOCK) void k_slot_chains(const uint2 *__restrict__ slot,
                                                       uint64_t nrows, uint8_t *__restrict__ seg,
                                                       uint2 *__restrict__ chunk,
                                                       const unsigned int *__restrict__ status) {
  if (*status) return;
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < nrows;
       base += (uint64_t)gridDim.x * BLOCK) {
    const uint64_t row = base + threadIdx.x;
    const bool live = row < nrows;
    uint2 v = make_uint2(SLOT_EMPTY, 0u);
    if (live) v = slot[row];
    chain_write(live && v.x <= v.y, v.x, v.y, live, row, seg, chunk);
  }
}

// Lane tier on slots: C = A.B (or A (B + I)), one lane per row of A, rows
// whose union has more than LANE_SLOTS runs or whose slices are long go to
// the group lists as in k_spgemm_lane.  A two-stage pipeline loads the A
// slot of the row after next while this row folds.  B slices are read slot
// by slot (LANE_BATCH at a time), through interval summaries (SUMM) or
// through B's chains or narrow joins (k_slot_fast); with c_seg the kernel
// writes C's chains.
#define DV_SLOT_LANE_PARAMS                                                                \
  Slots A, Slots B, uint64_t nrows, uint32_t lane_max, uint32_t mid_max,                   \
      uint2 *__restrict__ c_slot, uint2 *__restrict__ c_side, SideAlloc sa,                \
      uint32_t *__restrict__ mid_list, uint32_t *__restrict__ big_list,                    \
      OpCounters *__restrict__ ctr, unsigned int *__restrict__ status,                     \
      const uint2 *__restrict__ pre, const uint2 *__restrict__ suf,                        \
      const uint4 *__restrict__ pre2, const uint4 *__restrict__ suf2, uint32_t summ_min,   \
      uint32_t summ_shift, uint32_t plus_i, unsigned long long *__restrict__ parts,        \
      Chains BC, uint8_t *__restrict__ c_seg, uint2 *__restrict__ c_chunk
#define DV_SLOT_LANE_ARGS                                                                  \
  A, B, nrows, lane_max, mid_max, c_slot, c_side, sa, mid_list, big_list, ctr, status, pre, \
      suf, pre2, suf2, summ_min, summ_shift, plus_i, parts, BC, c_seg, c_chunk

enum SlotMode { SLOT_PLAIN = 0, SLOT_SUMM = 1 };

// The row folded into runs r and running interval [cl, ch] (ok false: more
// than LANE_SLOTS runs, or a slice past lane_max): written, or handed to the
// group lists.  Every lane of the wave calls it; `live` marks the lanes with
// a row.
template <bool CHAIN_OUT>
__device__ __forceinline__ void slot_row_emit(DV_SLOT_LANE_PARAMS, bool live, uint64_t row,
                                              uint64_t len, bool ok, Runs &r, uint32_t cl,
                                              uint32_t ch, unsigned long long &nnz_acc,
                                              uint32_t &ranges_acc) {
  int cls = 0;
  uint32_t nruns = 0;
  // live slots form a prefix: with slot 0 empty the row is [cl, ch] alone
  const bool single = ok && r.lo[0] > r.hi[0];
  if (live && single) {
    if (cl <= ch) {
      r.lo[0] = cl;
      r.hi[0] = ch;
      nnz_acc += (unsigned long long)(ch - cl) + 1;
      nruns = 1;
    }
    ranges_acc += nruns;
  } else if (live) {
    if (ok && cl <= ch && !insert(r, cl, ch)) {
      compact(r);
      ok = insert(r, cl, ch);
    }
    if (!ok) {
      cls = len <= mid_max ? 1 : 2;
    } else {
      compact(r);
#pragma unroll
      for (int t = 0; t < LANE_SLOTS; ++t)
        if (r.lo[t] <= r.hi[t]) {
          nnz_acc += (unsigned long long)(r.hi[t] - r.lo[t]) + 1;
          ++nruns;
        }
      ranges_acc += nruns;
    }
  }
  const uint32_t need_side = live && ok && nruns > 1 ? nruns + 1 : 0u;
  const uint32_t part = (blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6)) & (SLOT_SIDE_PARTS - 1);
  static_assert(LANE_SLOTS + 1 < 8, "a lane row needs fewer than 8 side entries");
  const unsigned long long loc = wave_alloc_small(need_side, &sa.ctr[part]);
  if (live && ok) {
    if (need_side && loc + need_side > sa.pcap) {
      atomicOr(status, 1u);  // side partition full: the build falls back
    } else {
      slot_store_runs(r, nruns, row, c_slot, c_side, part * sa.pcap + loc);
    }
  }
  wave_append(cls == 1, (uint32_t)row, mid_list, &ctr->n_mid);
  wave_append(cls == 2, (uint32_t)row, big_list, &ctr->n_big);
  if (CHAIN_OUT && c_seg)
    chain_write(live && ok && nruns == 1, r.lo[0], r.hi[0], live, row, c_seg, c_chunk);
}

// One row of C = A.B (or A (B + I)) by one lane, its A slot `a` given: the
// B slices are read slot by slot (LANE_BATCH at a time), or through interval
// summaries (SUMM); a row of one or a few runs is written, a row of more
// runs or a long slice goes to the group lists.  Every lane of the wave
// calls it (wave-wide reservations); `live` marks the lanes with a row.
// With CHAIN_OUT the row's chain data is written too.
template <int MODE, bool CHAIN_OUT>
__device__ __forceinline__ void slot_lane_row(DV_SLOT_LANE_PARAMS, bool live, uint64_t row,
                                              uint2 a, unsigned long long &nnz_acc,
                                              uint32_t &slice_acc,
                                              uint32_t &ranges_acc) {
  constexpr bool SUMM = MODE == SLOT_SUMM;
  uint64_t len = 0;
  bool ok = true;
  Runs r;
#pragma unroll
  for (int t = 0; t < LANE_SLOTS; ++t) {
    r.lo[t] = 0xFFFFFFFFu;
    r.hi[t] = 0u;
  }
  uint32_t cl = 0xFFFFFFFFu, ch = 0u;
  // A ranges: the slot itself, or the side entries of a marked row
  // (the reads of a marked row's ranges are waited on where they are
  // made, so the common row does not wait for the caller's prefetches)
  uint32_t na = 0;
  if (live) {
    if (a.x <= a.y) {
      na = 1;
    } else if (a.x == SLOT_MARK) {
      na = A.side[a.y].x;
      need(na);
    }
  }
  for (uint32_t q = 0; q < na; ++q) {
    uint2 ar = a;
    if (a.x > a.y) {
      ar = A.side[a.y + 1 + q];
      need(ar.x);
      need(ar.y);
    }
    const uint32_t s = ar.x, e = ar.y;
    len += (uint64_t)(e - s) + 1;
    if (len > lane_max) ok = false;
    if (!ok) continue;
    if (plus_i) fold(s, e, cl, ch, r, ok);  // C = A (B + I)
    if (SUMM && (s >> summ_shift) != (e >> summ_shift) && e - s + 1 >= summ_min &&
        (e >> summ_shift) - (s >> summ_shift) < 31) {
      fold_summarised(
          pre, suf, pre2, suf2, s, e, summ_shift,
          [&](uint32_t r0, uint32_t r1) { fold_slots(B, r0, r1 - 1, cl, ch, r, ok, slice_acc); },
          cl, ch, r, ok, slice_acc);
    } else {
      fold_slots(B, s, e, cl, ch, r, ok, slice_acc);
    }
  }
  slot_row_emit<CHAIN_OUT>(DV_SLOT_LANE_ARGS, live, row, len, ok, r, cl, ch, nnz_acc,
                           ranges_acc);
}

// Lane tier over the rows of A, one lane per row, persistent grid-stride
// loop; a two-stage pipeline loads the A slot of the row after next while
// this row folds.  Per-block (nnz, slice, ranges) sums go to
// parts[3 blockIdx.x ..].
template <int MODE>
__device__ __forceinline__ void slot_lane_body(DV_SLOT_LANE_PARAMS) {
  if (*status) return;
  unsigned long long nnz_acc = 0;
  uint32_t slice_acc = 0, ranges_acc = 0;  // per lane (the block sums are 64-bit)
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  uint64_t row = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  uint2 c_a = make_uint2(SLOT_EMPTY, 0u), n_a = make_uint2(SLOT_EMPTY, 0u);
  if (row < nrows) c_a = A.slot[row];
  if (row + stride < nrows) n_a = A.slot[row + stride];
  need(c_a.x);
  need(c_a.y);
  need(n_a.x);
  need(n_a.y);
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < nrows; base += stride) {
    row = base + threadIdx.x;
    const uint2 a = c_a;
    c_a = n_a;
    n_a = make_uint2(SLOT_EMPTY, 0u);
    if (row + 2 * stride < nrows) n_a = A.slot[row + 2 * stride];
    slot_lane_row<MODE, true>(DV_SLOT_LANE_ARGS, row < nrows, row, a, nnz_acc, slice_acc,
                                  ranges_acc);
    need(c_a.x);
    need(c_a.y);
    need(n_a.x);
    need(n_a.y);
  }
  nnz_acc = block_sum_u64(nnz_acc);
  const unsigned long long slice_sum = block_sum_u64(slice_acc);
  const unsigned long long ranges_sum = block_sum_u64(ranges_acc);
  if (threadIdx.x == 0) {
    parts[3 * blockIdx.x] = nnz_acc;
    parts[3 * blockIdx.x + 1] = slice_sum;
    parts[3 * blockIdx.x + 2] = ranges_sum;
  }
}

template <bool SUMM>
__global__ __launch_bounds__(BLOCK) void k_slot_lane(DV_SLOT_LANE_PARAMS) {
  slot_lane_body<SUMM ? SLOT_SUMM : SLOT_PLAIN>(DV_SLOT_LANE_ARGS);
}



// A product over B's chains: every row whose A row is one range [s, e]
// covering at most FAST_CHUNKS + 1 chunks of B and whose slice is one chain
// per part (the head s .. end of s's chunk, whole chunks, the tail start of
// e's chunk .. e), the parts touching, is one interval of C, found from a
// fixed set of loads issued together and a few min/max operations, with no
// branch that differs between lanes.  The wave queues its other rows (with
// their A slots) in LDS and, whenever 64 are queued, folds them slot by slot
// one per lane (slot_lane_row), so their round trips overlap the other
// waves' rows; what is left is folded after the loop.  Rows near an indel
// or a multi-range row defer together (1-3% of the rows of a power of a
// banded graph).  Narrow form (NB = 2 or 4: the product's A ranges span at
// most NB rows): a row whose A row is one range [s, e] with e - s < NB and
// whose NB slots s .. (repeating e) are single ranges each touching the
// running union is one interval, with the same queue for the others.
constexpr uint32_t FAST_CHUNKS = 4;
constexpr uint32_t DEFER_RING = 128;  // per wave: < 64 left + one row's worth of 64

template <int NB>  // 0: chain form; 2, 4: narrow form
__global__ __launch_bounds__(BLOCK) void k_slot_fast(DV_SLOT_LANE_PARAMS) {
  __shared__ uint4 ring[BLOCK / 64][DEFER_RING];
  // the whole block returns or none of it (the barriers of the sums)
  __shared__ unsigned int st0;
  if (threadIdx.x == 0) st0 = *status;
  __syncthreads();
  if (st0) return;
  uint4 *q = ring[threadIdx.x >> 6];
  uint32_t q_head = 0, q_tail = 0;  // wave-uniform
  unsigned long long nnz_acc = 0;
  uint32_t slice_acc = 0, ranges_acc = 0;  // per lane (the block sums are 64-bit)
  const uint32_t lane = __lane_id();
  const uint64_t stride = (uint64_t)gridDim.x * BLOCK;
  uint64_t row = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  uint2 c_a = make_uint2(SLOT_EMPTY, 0u), n_a = make_uint2(SLOT_EMPTY, 0u);
  if (row < nrows) c_a = A.slot[row];
  if (row + stride < nrows) n_a = A.slot[row + stride];
  // fold `count` (at most 64) queued rows from q_head, one per lane (in the
  // narrow form too with LANE_BATCH slots per round trip: the kernel then
  // fits 80 VGPRs, six waves per SIMD)
  auto fold_queued = [&](uint32_t count) {
    __builtin_amdgcn_wave_barrier();
    const bool on = lane < count;
    const uint4 d = q[(q_head + lane) & (DEFER_RING - 1)];
    slot_lane_row<SLOT_PLAIN, false>(DV_SLOT_LANE_ARGS, on, on ? d.x : 0u,
                                                            make_uint2(d.y, d.z), nnz_acc,
                                                            slice_acc, ranges_acc);
    q_head += count;
    __builtin_amdgcn_wave_barrier();
  };
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK; base < nrows; base += stride) {
    row = base + threadIdx.x;
    const bool live = row < nrows;
    const uint2 a = c_a;
    c_a = n_a;
    n_a = make_uint2(SLOT_EMPTY, 0u);
    if (row + 2 * stride < nrows) n_a = A.slot[row + 2 * stride];
    // lanes whose A row is not one range read row 0's data and discard it
    const bool one_a = live && a.x <= a.y;
    const uint32_t s = one_a ? a.x : 0u, e = one_a ? a.y : 0u;
    bool fast;
    uint32_t cl, ch, pieces;
    if constexpr (NB == 0) {
      const uint32_t c0 = s / CHAIN_ROWS, c1 = e / CHAIN_ROWS;
      const bool cross = c1 != c0;
      const uint32_t e0 = cross ? c0 * CHAIN_ROWS + (CHAIN_ROWS - 1) : e;
      const uint32_t s1 = cross ? c1 * CHAIN_ROWS : s;
      const uint32_t nmid = cross ? c1 - c0 - 1 : 0u;
      const uint2 v0 = B.slot[s], v1 = B.slot[e0], w0 = B.slot[s1], w1 = B.slot[e];
      const uint32_t t0 = c0 * CHAIN_ROWS + BC.seg[s], u0 = c1 * CHAIN_ROWS + BC.seg[s1];
      uint2 m[FAST_CHUNKS];
#pragma unroll
      for (int k = 0; k < (int)FAST_CHUNKS; ++k)
        m[k] = BC.chunk[nmid ? c0 + 1 + min((uint32_t)k, nmid - 1) : c0];
      // a part is one chain: several rows chained to its end, or one single
      // row.  (Bitwise tests and selects throughout: with short-circuit
      // tests the compiler sinks the load of seg[s1] into a branch, a
      // second round trip for every wave with a crossing slice.)
      const bool head_ok = ((e0 > s) & (t0 >= e0)) | ((e0 <= s) & (v0.x <= v0.y));
      const bool tail_ok = !cross | ((e > s1) & (u0 >= e)) | ((e <= s1) & (w0.x <= w0.y));
      fast = one_a & (nmid <= FAST_CHUNKS) & head_ok & tail_ok;
      // the union of the parts, in row order, each touching the running
      // union; a part that is not there joins the head's interval again,
      // which changes nothing
      cl = plus_i ? s : v0.x;
      ch = plus_i ? e : v1.y;
      auto join = [&](uint32_t lo, uint32_t hi) {
        fast = fast & (lo <= ch + 1u) & (hi + 1u >= cl);
        cl = min(cl, lo);
        ch = max(ch, hi);
      };
      if (plus_i) join(v0.x, v1.y);
#pragma unroll
      for (int k = 0; k < (int)FAST_CHUNKS; ++k) {
        const bool there = (uint32_t)k < nmid;
        fast = fast & (!there | (m[k].x <= m[k].y));
        join(there ? m[k].x : v0.x, there ? m[k].y : v1.y);
      }
      join(cross ? w0.x : v0.x, cross ? w1.y : v1.y);
      pieces = 3 + nmid + (cross ? 3 : 0);
    } else {
      // the NB slots from s, repeating e past it (a repeated range joins
      // unchanged), each one range touching the running union
      const uint32_t len = e - s;
      uint2 v[NB];
#pragma unroll
      for (int k = 0; k < NB; ++k) v[k] = B.slot[s + min((uint32_t)k, len)];
      fast = one_a & (len < (uint32_t)NB);
      cl = plus_i ? s : v[0].x;
      ch = plus_i ? e : v[0].y;
#pragma unroll
      for (int k = 0; k < NB; ++k) {
        fast = fast & (v[k].x <= v[k].y) & (v[k].x <= ch + 1u) & (v[k].y + 1u >= cl);
        cl = min(cl, v[k].x);
        ch = max(ch, v[k].y);
      }
      pieces = len + 1;
    }
    const bool empty_a = live && a.x == SLOT_EMPTY;
    if (fast) {
      c_slot[row] = make_uint2(cl, ch);
      nnz_acc += (unsigned long long)(ch - cl) + 1;
      ranges_acc += 1;
      slice_acc += pieces;
    } else if (empty_a) {
      c_slot[row] = make_uint2(SLOT_EMPTY, 0u);
    }
    // the rest to the wave's queue, folded 64 at a time
    const bool defer = live && !fast && !empty_a;
    const unsigned long long dm = __ballot(defer);
    if (dm) {
      if (defer)
        q[(q_tail + (uint32_t)__popcll(dm & ((1ull << lane) - 1))) & (DEFER_RING - 1)] =
            make_uint4((uint32_t)row, a.x, a.y, 0u);
      q_tail += (uint32_t)__popcll(dm);
      if (q_tail - q_head >= 64) fold_queued(64);
    }
  }
  if (q_tail != q_head) fold_queued(q_tail - q_head);
  nnz_acc = block_sum_u64(nnz_acc);
  const unsigned long long slice_sum = block_sum_u64(slice_acc);
  const unsigned long long ranges_sum = block_sum_u64(ranges_acc);
  if (threadIdx.x == 0) {
    parts[3 * blockIdx.x] = nnz_acc;
    parts[3 * blockIdx.x + 1] = slice_sum;
    parts[3 * blockIdx.x + 2] = ranges_sum;
  }
}

// Group tiers on slots: one row of C by G lanes with an LDS bit window, as
// spgemm_row; the row is counted first, then (several runs) its side entries
// are reserved and written in a second pass over the window.
template <int G>
__device__ __forceinline__ void slot_row(uint64_t row, const Slots &A, const Slots &B,
                                         uint64_t *lds, uint32_t W, uint32_t lane,
                                         uint2 *__restrict__ c_slot, uint2 *__restrict__ c_side,
                                         const SideAlloc &sa, OpCounters *ctr,
                                         unsigned int *status, unsigned long long &nnz_acc,
                                         uint32_t &ranges_acc, uint32_t plus_i) {
  const uint2 a = A.slot[row];
  uint32_t na = 0, aoff = 0;  // A ranges: the slot itself, or side[aoff + 1 ..]
  if (a.x <= a.y) {
    na = 1;
  } else if (a.x == SLOT_MARK) {
    aoff = a.y;
    na = A.side[aoff].x;
  }
  auto arange = [&](uint32_t q) { return na == 1 && a.x <= a.y ? a : A.side[aoff + 1 + q]; };
  // visit(x) for every range of B rows [s, e], rows strided over the group
  auto each_b = [&](uint32_t s, uint32_t e, auto &&visit) {
    for (uint64_t j = (uint64_t)s + lane; j <= e; j += G) {
      const uint2 b = B.slot[j];
      if (b.x <= b.y) {
        visit(b);
      } else if (b.x == SLOT_MARK) {
        const uint32_t nb = B.side[b.y].x;
        for (uint32_t q = 0; q < nb; ++q) visit(B.side[b.y + 1 + q]);
      }
    }
  };
  // 1. extent
  uint32_t lo = 0xFFFFFFFFu, hi = 0u;
  for (uint32_t q = 0; q < na; ++q) {
    const uint2 ar = arange(q);
    each_b(ar.x, ar.y, [&](uint2 x) {
      lo = min(lo, x.x);
      hi = max(hi, x.y);
    });
    if (plus_i) {
      lo = min(lo, ar.x);
      hi = max(hi, ar.y);
    }
  }
  lo = grp_min<G>(lo);
  hi = grp_max<G>(hi);
  if (lo > hi) {
    if (lane == 0) c_slot[row] = make_uint2(SLOT_EMPTY, 0u);
    return;
  }
  const uint64_t w0 = lo >> 6;
  const uint64_t nwords = (((uint64_t)hi + 1) >> 6) - w0 + 1;
  const uint64_t nchunks = (nwords + W - 1) / W;
  // one pass over the chunks: scatter, gather; WRITE stores entries at
  // word `base` of c_side onwards
  auto pass = [&](bool write, uint64_t base_word, bool scatter) -> uint32_t {
    uint32_t count = 0;
    uint64_t carry = 0;
    for (uint64_t c0 = 0; c0 < nwords; c0 += W) {
      const uint32_t cw = (uint32_t)min((uint64_t)W, nwords - c0);
      const uint64_t cb = (w0 + c0) << 6;
      const uint64_t clast = cb + ((uint64_t)cw << 6) - 1;
      if (scatter) {
        for (uint32_t j = lane; j < cw; j += G) lds[j] = 0;
        wave_sync();
        for (uint32_t q = 0; q < na; ++q) {
          const uint2 ar = arange(q);
          if (plus_i && lane == 0) {
            const uint64_t rs = max((uint64_t)ar.x, cb), rf = min((uint64_t)ar.y, clast);
            if (rs <= rf) window_set(lds, (uint32_t)(rs - cb), (uint32_t)(rf - cb));
          }
          each_b(ar.x, ar.y, [&](uint2 x) {
            const uint64_t rs = max((uint64_t)x.x, cb), rf = min((uint64_t)x.y, clast);
            if (rs <= rf) window_set(lds, (uint32_t)(rs - cb), (uint32_t)(rf - cb));
          });
        }
        wave_sync();
      }
      const uint32_t qn = (cw + G - 1) / G;
      const uint32_t j0 = min(lane * qn, cw), j1 = min(j0 + qn, cw);
      const uint64_t cy0 = (j0 == 0) ? carry : (lds[j0 - 1] >> 63);
      uint32_t nb = 0;
      {
        uint64_t cy = cy0;
        for (uint32_t j = j0; j < j1; ++j) {
          const uint64_t x = lds[j];
          nb += __popcll(run_bounds(x, cy));
          if (!write) nnz_acc += __popcll(x);
          cy = x >> 63;
        }
      }
      const uint32_t incl = grp_scan_incl<G>(nb, lane);
      const uint32_t tot = (uint32_t)__shfl((int)incl, G - 1, G);
      if (write) {
        uint32_t idx = count + incl - nb;
        uint32_t *words = reinterpret_cast<uint32_t *>(c_side);
        uint64_t cy = cy0;
        for (uint32_t j = j0; j < j1; ++j) {
          const uint64_t x = lds[j];
          uint64_t bnd = run_bounds(x, cy);
          cy = x >> 63;
          while (bnd) {
            const uint32_t p = (uint32_t)__builtin_ctzll(bnd);
            bnd &= bnd - 1;
            words[base_word + idx] = (uint32_t)(cb + ((uint64_t)j << 6) + p - (idx & 1u));
            ++idx;
          }
        }
      }
      count += tot;
      carry = lds[cw - 1] >> 63;
      wave_sync();
    }
    return count;
  };
  const uint32_t count = pass(false, 0, true);  // entries (2 per run)
  const uint32_t runs = count / 2;
  if (lane == 0) ranges_acc += runs;
  if (count == 2) {  // one run: the extent itself
    if (lane == 0) c_slot[row] = make_uint2(lo, hi);
    return;
  }
  const uint32_t part = blockIdx.x & (SLOT_SIDE_PARTS - 1);
  unsigned long long loc = 0;
  if (lane == 0) loc = atomicAdd(&sa.ctr[part], (unsigned long long)runs + 1);
  loc = ((unsigned long long)(uint32_t)__shfl((int)(uint32_t)(loc >> 32), 0, G) << 32) |
        (uint32_t)__shfl((int)(uint32_t)loc, 0, G);
  if (loc + runs + 1 > sa.pcap) {
    if (lane == 0) atomicOr(status, 1u);
    return;
  }
  const unsigned long long off = part * sa.pcap + loc;
  // a one-chunk window still holds the row's bits
  pass(true, 2 * (off + 1), nchunks > 1);
  if (lane == 0) {
    c_side[off] = make_uint2(runs, 0u);
    c_slot[row] = make_uint2(SLOT_MARK, (uint32_t)off);
  }
}

// One row list by groups of G lanes over blocks b of nb (nb a multiple of 8).
template <int G>
__device__ __forceinline__ void slot_group_rows(
    const Slots &A, const Slots &B, uint32_t W, uint32_t RPG, const uint32_t *__restrict__ rows,
    const unsigned int *__restrict__ n_dev, uint2 *__restrict__ c_slot,
    uint2 *__restrict__ c_side, const SideAlloc &sa, OpCounters *__restrict__ ctr,
    unsigned int *__restrict__ status, uint32_t plus_i, uint64_t *lds_all, uint32_t b,
    uint32_t nb) {
  constexpr uint32_t GPB = BLOCK / G;
  const uint32_t grp = threadIdx.x / G, lane = threadIdx.x % G;
  uint64_t *lds = lds_all + (size_t)grp * W;
  const uint64_t nl = *n_dev;
  unsigned long long nnz_acc = 0;
  uint32_t ranges_acc = 0;  // per lane (the block sums are 64-bit)
  const uint64_t groups = (uint64_t)nb * GPB;
  const uint64_t gid = (uint64_t)xcd_tile(b, nb) * GPB + grp;
  for (uint64_t i0 = gid * RPG; i0 < nl; i0 += groups * RPG) {
    const uint64_t i1 = min(i0 + RPG, nl);
    for (uint64_t i = uniform_u64<G>(i0); i < i1; ++i) {
      const uint64_t row = uniform_u64<G>(rows[i]);
      slot_row<G>(row, A, B, lds, W, lane, c_slot, c_side, sa, ctr, status, nnz_acc,
                  ranges_acc, plus_i);
    }
  }
  block_add_u64(nnz_acc, &ctr->nnz);
  block_add_u64(ranges_acc, &ctr->ranges);
}

// Both group tiers of a product in one launch: blocks [0, nb1) take the
// medium rows with groups of G1 lanes, the next nb2 blocks the large rows
// with 64-lane groups.
template <int G1>
__global__ __launch_bounds__(BLOCK) void k_slot_groups(
    Slots A, Slots B, uint32_t W1, uint32_t W2, uint32_t RPG, const uint32_t *__restrict__ rows1,
    const unsigned int *__restrict__ n1, uint32_t nb1, const uint32_t *__restrict__ rows2,
    const unsigned int *__restrict__ n2, uint32_t nb2, uint2 *__restrict__ c_slot,
    uint2 *__restrict__ c_side, SideAlloc sa, OpCounters *__restrict__ ctr,
    unsigned int *__restrict__ status, uint32_t plus_i) {
  extern __shared__ uint64_t lds_all[];
  // the status and the block's list length in one round trip (the lists
  // are usually empty, and then that is the launch's whole work)
  const bool medium = blockIdx.x < nb1;
  const unsigned int st = *status, nl = *(medium ? n1 : n2);
  if (st | (nl == 0u)) return;
  if (medium) {
    slot_group_rows<G1>(A, B, W1, RPG, rows1, n1, c_slot, c_side, sa, ctr, status, plus_i,
                        lds_all, blockIdx.x, nb1);
  } else {
    slot_group_rows<64>(A, B, W2, RPG, rows2, n2, c_slot, c_side, sa, ctr, status, plus_i,
                        lds_all, blockIdx.x - nb1, nb2);
  }
}

// Interval summaries of a slot matrix (as k_interval_summary).
template <int LANES, int K>
__global__ __launch_bounds__(BLOCK) void k_slot_summary(Slots B, uint64_t nrows,
                                                        uint2 *__restrict__ pre,
                                                        uint2 *__restrict__ suf,
                                                        uint4 *__restrict__ pre2,
                                                        uint4 *__restrict__ suf2,
                                                        const unsigned int *__restrict__ status) {
  if (*status) return;
  const int pos = threadIdx.x & (LANES - 1);
  for (uint64_t base = (uint64_t)blockIdx.x * BLOCK * K; base < nrows;
       base += (uint64_t)gridDim.x * BLOCK * K) {
    const uint64_t r0 = base + (uint64_t)threadIdx.x * K;
    uint2 sl[K], x[K];
    bool single = true;
    // every load is made (rows past the end read the last row and are then
    // emptied): a load under a branch waits at the branch's end
#pragma unroll
    for (int k = 0; k < K; ++k) sl[k] = B.slot[min(r0 + k, nrows - 1)];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      sl[k] = r0 + k < nrows ? sl[k] : make_uint2(SLOT_EMPTY, 0u);
      const bool one = sl[k].x <= sl[k].y;
      x[k] = one ? sl[k] : make_uint2(SUMM_EMPTY, 0u);
      single = single && one;
    }
    summary_lane<LANES, K>(
        x, single,
        [&](int k, Ivl2 &v) {
          if (sl[k].x != SLOT_MARK) return;
          const uint32_t nb = B.side[sl[k].y].x;
          for (uint32_t q = 0; q < nb && !ivl_unknown(v); ++q) {
            const uint2 y = B.side[sl[k].y + 1 + q];
            v = ivl_join(v, Ivl2{y.x, y.y, SUMM_EMPTY, 0u});
          }
        },
        r0, nrows, pos, pre, suf, pre2, suf2);
  }
}

// rCRS -> slots (a row of several ranges into the side array).  Four rows
// per lane: their bounds, then their first ranges, each loaded at once.
// Status bits: 8 when a wave's rows need more side entries than a partition
// holds (the input does not fit the slot form), 4 when a row's ranges are
// not canonical (each range must end at least two columns before the next
// starts): slot-form I + A and the paired schedule assume canonical rows.
template <typename S>
__global__ __launch_bounds__(BLOCK) void k_rcrs_to_slots(Mat<S> A, uint64_t nrows,
                                                         uint2 *__restrict__ slot,
                                                         uint2 *__restrict__ side, SideAlloc sa,
                                                         unsigned int *__restrict__ status) {
  constexpr int K = 4;
  for (uint64_t b0 = (uint64_t)blockIdx.x * BLOCK * K; b0 < nrows;
       b0 += (uint64_t)gridDim.x * BLOCK * K) {
    const uint64_t r0 = b0 + (uint64_t)threadIdx.x * K;
    // the rows' bounds (past the end: the last bound, so those rows are
    // empty), then their first ranges, each set loaded at once (a load under
    // a branch would wait at the branch's end)
    S rb[K + 1];
#pragma unroll
    for (int k = 0; k <= K; ++k) rb[k] = A.rm[min(r0 + k, nrows)];
    uint2 x[K];
    load_first_ranges<K>(A.en, rb, x, SLOT_EMPTY);
    uint64_t need64 = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint64_t nr = (uint64_t)(rb[k + 1] - rb[k]) / 2;
      need64 += nr > 1 ? nr + 1 : 0;
    }
    // (more than a side partition holds: flagged below without a reservation)
    const uint32_t need = need64 <= sa.pcap ? (uint32_t)need64 : 0u;
    const uint32_t part = (blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6)) & (SLOT_SIDE_PARTS - 1);
    const unsigned long long loc = wave_alloc(need, &sa.ctr[part]);
    if (need64 > sa.pcap || (need && loc + need > sa.pcap)) {
      atomicOr(status, 8u);
      continue;
    }
    unsigned long long off = part * sa.pcap + loc;
    bool canonical = true;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const uint64_t row = r0 + k;
      if (row >= nrows) continue;
      const uint64_t nr = (uint64_t)(rb[k + 1] - rb[k]) / 2;
      if (nr <= 1) {
        slot[row] = x[k];
      } else {
        side[off] = make_uint2((uint32_t)nr, 0u);
        side[off + 1] = x[k];
        uint32_t prev_end = x[k].y;
        for (uint64_t q = 1; q < nr; ++q) {
          const uint2 v = ld_range(A.en, rb[k] + 2 * q);
          canonical = canonical && (uint64_t)prev_end + 1 < (uint64_t)v.x;
          prev_end = v.y;
          side[off + 1 + q] = v;
        }
        slot[row] = make_uint2(SLOT_MARK, (uint32_t)off);
        off += nr + 1;
      }
    }
    if (!canonical) atomicOr(status, 4u);
  }
}

__global__ __launch_bounds__(BLOCK) void k_slot_identity(uint64_t n, uint2 *__restrict__ slot) {
  for (uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * BLOCK)
    slot[i] = make_uint2((uint32_t)i, (uint32_t)i);
}

// C = I + A on slots (range_sparse.hpp:112-357 for canonical rows: the
// union of A's row and {i}).  A lane takes PI_ROWS rows per visit (rows
// k * BLOCK apart, so every load is coalesced): their slots are loaded
// together, and the side entries of all of them come from one reservation
// per wave, so a visit costs one round trip for the slots, one for the
// marked rows' headers and one atomic, for four rows.
constexpr int PI_ROWS = 4;

__global__ __launch_bounds__(BLOCK) void k_slot_plus_identity(
    Slots A, uint64_t nrows, uint2 *__restrict__ c_slot, uint2 *__restrict__ c_side,
    SideAlloc sa, unsigned int *__restrict__ status, unsigned long long *__restrict__ parts) {
  if (*status) return;
  unsigned long long nnz_acc = 0;
  uint32_t ranges_acc = 0;  // per lane (the block sums are 64-bit)
  const uint32_t part = (blockIdx.x * (BLOCK / 64) + (threadIdx.x >> 6)) & (SLOT_SIDE_PARTS - 1);
  for (uint64_t b0 = (uint64_t)blockIdx.x * BLOCK * PI_ROWS; b0 < nrows;
       b0 += (uint64_t)gridDim.x * BLOCK * PI_ROWS) {
    uint2 a[PI_ROWS];
    uint32_t na[PI_ROWS], need[PI_ROWS];
    // every load is made (rows past the end read the last row and are then
    // emptied): a load under a branch waits at the branch's end
#pragma unroll
    for (int k = 0; k < PI_ROWS; ++k)
      a[k] = A.slot[min(b0 + (uint64_t)k * BLOCK + threadIdx.x, nrows - 1)];
#pragma unroll
    for (int k = 0; k < PI_ROWS; ++k) {
      if (b0 + (uint64_t)k * BLOCK + threadIdx.x >= nrows) a[k] = make_uint2(SLOT_EMPTY, 0u);
      na[k] = a[k].x <= a[k].y ? 1u : 0u;
    }
    // A's ranges of a marked row: from the side array
#pragma unroll
    for (int k = 0; k < PI_ROWS; ++k)
      if (a[k].x == SLOT_MARK) na[k] = A.side[a[k].y].x;
    // one range that touches i stays one range; otherwise the merge of A's
    // ranges with [i, i] goes to the side array (at most na + 1 runs and a
    // header)
    uint32_t total = 0;
#pragma unroll
    for (int k = 0; k < PI_ROWS; ++k) {
      const uint64_t row = b0 + (uint64_t)k * BLOCK + threadIdx.x;
      const uint32_t i = (uint32_t)row;
      const bool one = na[k] == 0 || (na[k] == 1 && iv_touch(a[k].x, a[k].y, i, i));
      need[k] = row < nrows && !one ? na[k] + 2 : 0u;
      total += need[k];
    }
    const unsigned long long loc = wave_alloc(total, &sa.ctr[part]);
    const bool fits = loc + total <= sa.pcap;
    if (!fits) atomicOr(status, 1u);
    unsigned long long off = part * sa.pcap + loc;
#pragma unroll
    for (int k = 0; k < PI_ROWS; ++k) {
      const uint64_t row = b0 + (uint64_t)k * BLOCK + threadIdx.x;
      if (row >= nrows) continue;
      const uint32_t i = (uint32_t)row;
      const uint2 x0 = a[k];
      const uint32_t n0 = na[k];
      if (need[k] == 0) {
        const uint2 u = n0 == 0 ? make_uint2(i, i) : make_uint2(min(x0.x, i), max(x0.y, i));
        c_slot[row] = u;
        nnz_acc += (unsigned long long)(u.y - u.x) + 1;
        ranges_acc += 1;
        continue;
      }
      if (!fits) continue;
      auto arange = [&](uint32_t q) {
        return n0 == 1 && x0.x <= x0.y ? x0 : A.side[x0.y + 1 + q];
      };
      // merge: i goes in order; touching neighbours fold together
      uint32_t m = 0;
      bool placed = false;
      uint32_t plo = 0, phi = 0;
      bool have = false;
      auto put = [&](uint32_t lo, uint32_t hi) {
        if (have && lo <= phi + 1u) {
          phi = max(phi, hi);
        } else {
          if (have) {
            c_side[off + 1 + m] = make_uint2(plo, phi);
            nnz_acc += (unsigned long long)(phi - plo) + 1;
            ++m;
          }
          plo = lo;
          phi = hi;
          have = true;
        }
      };
      for (uint32_t q = 0; q < n0; ++q) {
        const uint2 x = arange(q);
        if (!placed && i < x.x) {
          put(i, i);
          placed = true;
        }
        put(x.x, x.y);
      }
      if (!placed) put(i, i);
      c_side[off + 1 + m] = make_uint2(plo, phi);
      nnz_acc += (unsigned long long)(phi - plo) + 1;
      ++m;
      ranges_acc += m;
      if (m == 1) {
        c_slot[row] = make_uint2(plo, phi);
      } else {
        c_side[off] = make_uint2(m, 0u);
        c_slot[row] = make_uint2(SLOT_MARK, (uint32_t)off);
      }
      off += need[k];
    }
  }
  // per-block partial sums (as the lane tier)
  nnz_acc = block_sum_u64(nnz_acc);
  const unsigned long long ranges_sum = block_sum_u64(ranges_acc);
  if (threadIdx.x == 0) {
    parts[3 * blockIdx.x] = nnz_acc;
    parts[3 * blockIdx.x + 1] = 0;
    parts[3 * blockIdx.x + 2] = ranges_sum;
  }
}

// Per-block (nnz, slice, ranges) sums of the lane-tier launches, per_op
// triples for each of nops operations, of which the first blocks.n[i] were
// written by operation i's launch (every block of it writes its triple):
// one block per operation adds them to its counters.
__global__ __launch_bounds__(BLOCK) void k_reduce_parts(const unsigned long long *__restrict__ parts,
                                                        uint32_t per_op, PartBlocks blocks,
                                                        const unsigned long long *__restrict__ side_ctr,
                                                        OpCounters *__restrict__ ctr) {
  const unsigned long long *p = parts + 3ull * per_op * blockIdx.x;
  const uint32_t nb = blocks.all ? per_op : blocks.n[blockIdx.x];
  unsigned long long a = 0, b = 0, c = 0;
  for (uint32_t j = threadIdx.x; j < nb; j += BLOCK) {
    a += p[3 * j];
    b += p[3 * j + 1];
    c += p[3 * j + 2];
  }
  a = block_sum_u64(a);
  b = block_sum_u64(b);
  c = block_sum_u64(c);
  // side entries used: in total, and in the fullest partition
  unsigned long long used = 0, most = 0;
  if (threadIdx.x < SLOT_SIDE_PARTS) used = most = side_ctr[SLOT_SIDE_PARTS * blockIdx.x + threadIdx.x];
  used = block_sum_u64(used);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) most = max(most, (unsigned long long)__shfl_xor(most, o, 64));
  if (threadIdx.x == 0) {
    ctr[blockIdx.x].nnz += a;
    ctr[blockIdx.x].slice += b;
    ctr[blockIdx.x].ranges += c;
    ctr[blockIdx.x].total = used;
    ctr[blockIdx.x].pad = (unsigned int)min(most, 0xFFFFFFFFull);
  }
}

// Slots -> rCRS: entry counts per row, then (after the row-pointer scan) the
// entries at their row pointers.
// The entry counts of a slot matrix's rows (2 per range) and the
// row-pointer scan's tile sums in one pass: each block takes one scan tile
// of rows, writes their counts and the tile's sum.  The slots are loaded together (rows past the end read the
// last row and count nothing).
__global__ __launch_bounds__(BLOCK) void k_slot_tile_counts(Slots A, uint64_t n,
                                                            uint32_t *__restrict__ cnt,
                                                            unsigned long long *__restrict__ tsum,
                                                            const unsigned int *__restrict__ status) {
  // the whole block returns or none of it (the barriers of the sum)
  __shared__ unsigned int st0;
  if (threadIdx.x == 0) st0 = *status;
  __syncthreads();
  if (st0) return;  // a product ran out of side entries: slots not valid
  const uint64_t t0 = (uint64_t)blockIdx.x * SCAN_TILE;
  uint2 a[SCAN_ITEMS];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k)
    a[k] = A.slot[min(t0 + (uint64_t)k * BLOCK + threadIdx.x, n - 1)];
  unsigned long long v = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const uint64_t i = t0 + (uint64_t)k * BLOCK + threadIdx.x;
    if (i >= n) continue;
    const uint32_t c = a[k].x <= a[k].y ? 2u : a[k].x == SLOT_MARK ? 2u * A.side[a[k].y].x : 0u;
    cnt[i] = c;
    v += c;
  }
  v = block_sum_u64(v);
  if (threadIdx.x == 0) tsum[blockIdx.x] = v;
}

template <typename S>
__global__ __launch_bounds__(BLOCK) void k_slot_emit(Slots A, uint64_t nrows,
                                                     const S *__restrict__ rm,
                                                     uint32_t *__restrict__ en) {
  for (uint64_t row = (uint64_t)blockIdx.x * BLOCK + threadIdx.x; row < nrows;
       row += (uint64_t)gridDim.x * BLOCK) {
    const uint2 a = A.slot[row];
    const uint64_t p = rm[row];
    if (a.x <= a.y) {
      *reinterpret_cast<uint2 *>(en + p) = a;
    } else if (a.x == SLOT_MARK) {
      const uint32_t nr = A.side[a.y].x;
      for (uint32_t q = 0; q < nr; ++q)
        *reinterpret_cast<uint2 *>(en + p + 2 * q) = A.side[a.y + 1 + q];
    }
  }
}

}  // namespace

void fill_u32(hipStream_t s, uint32_t *p, uint64_t n, uint32_t v) {
  if (!n) return;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(BLOCK), 0, s, p, n, v);
}

void fill2_u32(hipStream_t s, uint32_t *p, uint64_t n, uint32_t *q, uint64_t m, uint32_t v) {
  if (n + m == 0) return;
  hipLaunchKernelGGL(k_fill2_u32, dim3(grid_for(n + m)), dim3(BLOCK), 0, s, p, n, q, m, v);
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

uint64_t scan_words(uint64_t n) { return (n + SCAN_TILE - 1) / SCAN_TILE + 1; }

void scan_tiles_rows(hipStream_t s, uint32_t sb, const uint32_t *cnt, uint64_t n, void *rowmap,
                     unsigned long long *tile_sums, OpCounters *ctr, uint64_t cap,
                     unsigned int *status, const uint32_t *stage, uint32_t stage_entries,
                     uint32_t *c_entries);

void scan_rowmap(hipStream_t s, uint32_t sb, const uint32_t *cnt, uint64_t n, void *rowmap,
                 unsigned long long *tile_sums, OpCounters *ctr, uint64_t cap,
                 unsigned int *status, const uint32_t *stage, uint32_t stage_entries,
                 uint32_t *c_entries, const unsigned long long *parts, uint32_t nparts) {
  if (n == 0) {
    fill_u32(s, static_cast<uint32_t *>(rowmap), sb / 4, 0u);
    return;
  }
  const uint64_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  hipLaunchKernelGGL(k_tile_sums, dim3((uint32_t)ntiles), dim3(BLOCK), 0, s, cnt, n, tile_sums,
                     parts, parts ? nparts : 0u, ctr);
  scan_tiles_rows(s, sb, cnt, n, rowmap, tile_sums, ctr, cap, status, stage, stage_entries,
                  c_entries);
}

// The row-pointer scan after its tile sums: the spine (many tiles), then
// the rows.
void scan_tiles_rows(hipStream_t s, uint32_t sb, const uint32_t *cnt, uint64_t n, void *rowmap,
                     unsigned long long *tile_sums, OpCounters *ctr, uint64_t cap,
                     unsigned int *status, const uint32_t *stage, uint32_t stage_entries,
                     uint32_t *c_entries) {
  const uint64_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  const int prefixed = ntiles >= SPINE_MIN_TILES;
  if (prefixed)
    hipLaunchKernelGGL(k_scan_spine, dim3(1), dim3(SPINE_THREADS), 0, s, tile_sums, ntiles);
  const bool emit = c_entries != nullptr;
#define DV_SR(S_, E_)                                                                       \
  hipLaunchKernelGGL((k_scan_rows<S_, E_>), dim3((uint32_t)ntiles), dim3(BLOCK), 0, s, cnt, n, \
                     static_cast<S_ *>(rowmap), tile_sums, stage, stage_entries, c_entries,     \
                     ctr, (unsigned long long)cap, sb, status, prefixed)
  if (sb == 4) {
    if (emit) DV_SR(uint32_t, true);
    else DV_SR(uint32_t, false);
  } else {
    if (emit) DV_SR(uint64_t, true);
    else DV_SR(uint64_t, false);
  }
#undef DV_SR
}

void spadd_count(hipStream_t s, uint32_t sb, DView a, DView b, uint32_t *cnt,
                 const unsigned int *status) {
  if (!a.nrows) return;
  if (sb == 4)
    hipLaunchKernelGGL(k_spadd_count<uint32_t>, dim3(grid_for(a.nrows)), dim3(BLOCK), 0, s,
                       mat<uint32_t>(a), mat<uint32_t>(b), a.nrows, cnt, status);
  else
    hipLaunchKernelGGL(k_spadd_count<uint64_t>, dim3(grid_for(a.nrows)), dim3(BLOCK), 0, s,
                       mat<uint64_t>(a), mat<uint64_t>(b), a.nrows, cnt, status);
}

void spadd_write(hipStream_t s, uint32_t sb, DView a, DView b, const void *c_rowmap,
                 uint32_t *c_entries, const unsigned int *status) {
  if (!a.nrows) return;
  if (sb == 4)
    hipLaunchKernelGGL(k_spadd_write<uint32_t>, dim3(grid_for(a.nrows)), dim3(BLOCK), 0, s,
                       mat<uint32_t>(a), mat<uint32_t>(b), a.nrows,
                       static_cast<const uint32_t *>(c_rowmap), c_entries, status);
  else
    hipLaunchKernelGGL(k_spadd_write<uint64_t>, dim3(grid_for(a.nrows)), dim3(BLOCK), 0, s,
                       mat<uint64_t>(a), mat<uint64_t>(b), a.nrows,
                       static_cast<const uint64_t *>(c_rowmap), c_entries, status);
}

void range_nnz(hipStream_t s, const uint32_t *entries, uint64_t max_entries, OpCounters *ctr,
               const unsigned int *status) {
  if (max_entries < 2) return;
  // few blocks: each adds its sum to one counter with one atomic
  hipLaunchKernelGGL(k_range_nnz, dim3(grid_for(max_entries / 2, BLOCK * 8, 1024)), dim3(BLOCK),
                     0, s, reinterpret_cast<const uint2 *>(entries), ctr, status);
}

void spgemm_lane(hipStream_t s, uint32_t sb, DView a, DView b, uint32_t stage, uint32_t lane_max,
                 uint32_t mid_max, uint32_t grid, uint32_t *cnt, uint32_t *stage_buf, uint32_t *ovf_list,
                 uint32_t *mid_list, uint32_t *big_list, OpCounters *ctr,
                 const unsigned int *status, const void *summ, uint32_t summ_min,
                 uint32_t summ_rows, uint32_t plus_i, unsigned long long *parts) {
  if (!a.nrows) return;
  // summary layout: pre, suf (uint2 per row), then pre2, suf2 (uint4 per row)
  const uint2 *pre = static_cast<const uint2 *>(summ);
  const uint2 *suf = pre ? pre + b.nrows : nullptr;
  const uint4 *pre2 = pre ? reinterpret_cast<const uint4 *>(pre + 2 * b.nrows) : nullptr;
  const uint4 *suf2 = pre ? pre2 + b.nrows : nullptr;
  // without summaries the kernel is built without the summary walk (fewer
  // registers, more waves per SIMD)
#define DV_LANE(S_, M_)                                                                        \
  hipLaunchKernelGGL((k_spgemm_lane<S_, M_>), dim3(grid), dim3(BLOCK), 0, s, mat<S_>(a),       \
                     mat<S_>(b), a.nrows, stage, lane_max, mid_max, cnt, stage_buf, ovf_list, \
                     mid_list, big_list, ctr, status, pre, suf, pre2, suf2, summ_min,          \
                     (uint32_t)__builtin_ctz(summ_rows ? summ_rows : 1), plus_i, parts)
  if (sb == 4) {
    if (pre) DV_LANE(uint32_t, true);
    else DV_LANE(uint32_t, false);
  } else {
    if (pre) DV_LANE(uint64_t, true);
    else DV_LANE(uint64_t, false);
  }
#undef DV_LANE
}

uint32_t lane_blocks_per_cu(uint32_t sb, bool summ) {
  // resident blocks of the lane kernel per CU (occupancy), per variant; the
  // cache is shared by every context and thread (a racing fill stores the
  // same value)
  static std::atomic<int> cache[4] = {0, 0, 0, 0};
  const int k = (sb == 4 ? 0 : 2) + (summ ? 1 : 0);
  if (!cache[k].load(std::memory_order_relaxed)) {
    int nb = 0;
    hipError_t e;
    if (sb == 4)
      e = summ ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spgemm_lane<uint32_t, true>, BLOCK, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spgemm_lane<uint32_t, false>, BLOCK, 0);
    else
      e = summ ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spgemm_lane<uint64_t, true>, BLOCK, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_spgemm_lane<uint64_t, false>, BLOCK, 0);
    if (e != hipSuccess || nb < 1) {
      (void)hipGetLastError();
      nb = 8;
    }
    cache[k] = nb;
  }
  return (uint32_t)cache[k];
}

uint64_t summary_words(uint64_t nrows) { return 6 * nrows; }

uint32_t summary_block_rows(uint64_t span) {
  uint32_t w = SUMM_ROWS_MIN;
  while (w < SUMM_ROWS_MAX && 8ull * w <= span) w <<= 1;
  return w;
}

void block_summary(hipStream_t s, uint32_t sb, DView b, void *summ, const unsigned int *status,
                   uint32_t block_rows) {
  if (!b.nrows) return;
  uint2 *pre = static_cast<uint2 *>(summ), *suf = pre + b.nrows;
  uint4 *pre2 = reinterpret_cast<uint4 *>(pre + 2 * b.nrows), *suf2 = pre2 + b.nrows;
  const uint32_t grid = grid_for(b.nrows, BLOCK * SUMM_ROWS_PER_LANE);
#define DV_SUMM(S_, L_)                                                                     \
  hipLaunchKernelGGL((k_interval_summary<S_, L_>), dim3(grid), dim3(BLOCK), 0, s, mat<S_>(b), \
                     b.nrows, pre, suf, pre2, suf2, status)
#define DV_SUMM_W(S_)                        \
  switch (block_rows / SUMM_ROWS_PER_LANE) { \
    case 8: DV_SUMM(S_, 8); break;           \
    case 16: DV_SUMM(S_, 16); break;         \
    case 32: DV_SUMM(S_, 32); break;         \
    default: DV_SUMM(S_, 64); break;         \
  }
  if (sb == 4) {
    DV_SUMM_W(uint32_t);
  } else {
    DV_SUMM_W(uint64_t);
  }
#undef DV_SUMM_W
#undef DV_SUMM
}

void spgemm_group(hipStream_t s, uint32_t sb, DView a, DView b, const GroupLaunch &g) {
  const bool direct = g.c_entries != nullptr;
  if (direct && !g.n_dev && g.n_host == 0) return;
  if (sb == 4) {
    if (direct) group_by_lanes<uint32_t, true>(s, a, b, g);
    else group_by_lanes<uint32_t, false>(s, a, b, g);
  } else {
    if (direct) group_by_lanes<uint64_t, true>(s, a, b, g);
    else group_by_lanes<uint64_t, false>(s, a, b, g);
  }
}

void spgemm_emit_staged(hipStream_t s, uint32_t sb, uint64_t nrows, const void *c_rowmap,
                        const uint32_t *stage_buf, uint32_t stage, uint32_t *c_entries,
                        const unsigned int *status) {
  if (!nrows) return;
  if (sb == 4)
    hipLaunchKernelGGL(k_spgemm_emit_staged<uint32_t>, dim3(grid_for(nrows)), dim3(BLOCK), 0,
                       s, nrows, static_cast<const uint32_t *>(c_rowmap), stage_buf, stage,
                       c_entries, status);
  else
    hipLaunchKernelGGL(k_spgemm_emit_staged<uint64_t>, dim3(grid_for(nrows)), dim3(BLOCK), 0,
                       s, nrows, static_cast<const uint64_t *>(c_rowmap), stage_buf, stage,
                       c_entries, status);
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

void concat_component(hipStream_t s, uint32_t sb, const void *rm, const uint32_t *en, uint64_t n,
                      uint64_t n_entries, uint64_t row0, uint64_t ent0, uint32_t col0,
                      uint64_t *out_rm, uint32_t *out_en, bool write_last) {
  const uint64_t work = std::max<uint64_t>(n + 1, n_entries);
  if (sb == 4)
    hipLaunchKernelGGL(k_concat<uint32_t>, dim3(grid_for(work)), dim3(BLOCK), 0, s,
                       static_cast<const uint32_t *>(rm), en, n, n_entries, row0, ent0, col0,
                       out_rm, out_en, write_last ? 1 : 0);
  else
    hipLaunchKernelGGL(k_concat<uint64_t>, dim3(grid_for(work)), dim3(BLOCK), 0, s,
                       static_cast<const uint64_t *>(rm), en, n, n_entries, row0, ent0, col0,
                       out_rm, out_en, write_last ? 1 : 0);
}

void extract_component(hipStream_t s, uint32_t sb, const uint64_t *rm, const uint32_t *en,
                       uint64_t row0, uint64_t n, uint64_t n_entries, uint32_t col0,
                       void *out_rm, uint32_t *out_en) {
  const uint64_t work = std::max<uint64_t>(n + 1, n_entries);
  if (sb == 4)
    hipLaunchKernelGGL(k_extract<uint32_t>, dim3(grid_for(work)), dim3(BLOCK), 0, s, rm, en, row0,
                       n, col0, static_cast<uint32_t *>(out_rm), out_en);
  else
    hipLaunchKernelGGL(k_extract<uint64_t>, dim3(grid_for(work)), dim3(BLOCK), 0, s, rm, en, row0,
                       n, col0, static_cast<uint64_t *>(out_rm), out_en);
}

void gather_bounds(hipStream_t s, const uint64_t *rm, const uint64_t *bounds, uint32_t m,
                   unsigned long long *out) {
  hipLaunchKernelGGL(k_gather_bounds, dim3(1), dim3(BLOCK), 0, s, rm, bounds, m, out);
}

void slice_nnz(hipStream_t s, const uint32_t *entries, uint64_t n_entries,
               unsigned long long *out) {
  if (n_entries < 2) return;
  hipLaunchKernelGGL(k_slice_nnz, dim3(grid_for(n_entries / 2, BLOCK * 8, 4096)), dim3(BLOCK), 0,
                     s, reinterpret_cast<const uint2 *>(entries), n_entries / 2, out);
}

void query_multi(hipStream_t s, const IndexDesc *table, uint32_t ntable, const uint32_t *qidx,
                 const uint32_t *qi, const uint32_t *qj, uint32_t n, uint8_t *out) {
  if (!n) return;
  hipLaunchKernelGGL(k_query_multi, dim3(grid_for(n, BLOCK, 1024)), dim3(BLOCK), 0, s, table,
                     ntable, qidx, qi, qj, n, out);
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

void check_canonical(hipStream_t s, uint32_t sb, DView m, unsigned int *status) {
  if (!m.nrows) return;
  if (sb == 4)
    hipLaunchKernelGGL(k_check_canonical<uint32_t>, dim3(grid_for(m.nrows)), dim3(BLOCK), 0, s,
                       mat<uint32_t>(m), m.nrows, status);
  else
    hipLaunchKernelGGL(k_check_canonical<uint64_t>, dim3(grid_for(m.nrows)), dim3(BLOCK), 0, s,
                       mat<uint64_t>(m), m.nrows, status);
}

void window_pieces(hipStream_t s, uint32_t sb, DView a, DView b, uint32_t *cnt,
                   const void *p_rowmap, uint32_t *p_entries, const unsigned int *status) {
  if (!a.nrows) return;
  const bool write = p_entries != nullptr;
  if (sb == 4) {
    if (write)
      hipLaunchKernelGGL((k_window_pieces<uint32_t, true>), dim3(grid_for(a.nrows)), dim3(BLOCK), 0, s,
                         mat<uint32_t>(a), mat<uint32_t>(b), a.nrows, b.ncols, cnt,
                         static_cast<const uint32_t *>(p_rowmap), p_entries, status);
    else
      hipLaunchKernelGGL((k_window_pieces<uint32_t, false>), dim3(grid_for(a.nrows)), dim3(BLOCK), 0, s,
                         mat<uint32_t>(a), mat<uint32_t>(b), a.nrows, b.ncols, cnt, nullptr, nullptr,
                         status);
  } else {
    if (write)
      hipLaunchKernelGGL((k_window_pieces<uint64_t, true>), dim3(grid_for(a.nrows)), dim3(BLOCK), 0, s,
                         mat<uint64_t>(a), mat<uint64_t>(b), a.nrows, b.ncols, cnt,
                         static_cast<const uint64_t *>(p_rowmap), p_entries, status);
    else
      hipLaunchKernelGGL((k_window_pieces<uint64_t, false>), dim3(grid_for(a.nrows)), dim3(BLOCK), 0, s,
                         mat<uint64_t>(a), mat<uint64_t>(b), a.nrows, b.ncols, cnt, nullptr, nullptr,
                         status);
  }
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


// ---------------------------------------------------------------------------
// Slot-format launchers.
namespace {
Slots slots_of(SlotView v) {
  return Slots{static_cast<const uint2 *>(v.slot), static_cast<const uint2 *>(v.side)};
}
}  // namespace

void slots_from_rcrs(hipStream_t s, uint32_t sb, DView a, void *slot, void *side,
                     uint64_t side_cap, unsigned long long *side_ctr, unsigned int *status) {
  if (!a.nrows) return;
  uint2 *sl = static_cast<uint2 *>(slot), *sd = static_cast<uint2 *>(side);
  const SideAlloc sa{side_ctr, side_cap / SLOT_SIDE_PARTS};
  if (sb == 4)
    hipLaunchKernelGGL(k_rcrs_to_slots<uint32_t>, dim3(grid_for(a.nrows, BLOCK * 4)), dim3(BLOCK), 0, s,
                       mat<uint32_t>(a), a.nrows, sl, sd, sa, status);
  else
    hipLaunchKernelGGL(k_rcrs_to_slots<uint64_t>, dim3(grid_for(a.nrows, BLOCK * 4)), dim3(BLOCK), 0, s,
                       mat<uint64_t>(a), a.nrows, sl, sd, sa, status);
}

void slots_chains(hipStream_t s, const void *slot, uint64_t n, ChainView c,
                  const unsigned int *status) {
  if (!n) return;
  hipLaunchKernelGGL(k_slot_chains, dim3(grid_for(n)), dim3(BLOCK), 0, s,
                     static_cast<const uint2 *>(slot), n,
                     static_cast<uint8_t *>(const_cast<void *>(c.seg)),
                     static_cast<uint2 *>(const_cast<void *>(c.chunk)), status);
}

void slots_identity(hipStream_t s, uint64_t n, void *slot) {
  if (!n) return;
  hipLaunchKernelGGL(k_slot_identity, dim3(grid_for(n)), dim3(BLOCK), 0, s, n,
                     static_cast<uint2 *>(slot));
}

uint32_t slots_plus_identity_blocks(uint64_t n, uint32_t max_blocks) {
  return n ? grid_for(n, (uint64_t)BLOCK * PI_ROWS, max_blocks) : 0u;
}

void slots_plus_identity(hipStream_t s, SlotView a, uint64_t n, void *c_slot, void *c_side,
                         uint64_t side_cap, unsigned long long *side_ctr, unsigned int *status,
                         unsigned long long *parts, uint32_t max_blocks) {
  if (!n) return;
  hipLaunchKernelGGL(k_slot_plus_identity, dim3(slots_plus_identity_blocks(n, max_blocks)), dim3(BLOCK), 0, s,
                     slots_of(a), n, static_cast<uint2 *>(c_slot), static_cast<uint2 *>(c_side),
                     SideAlloc{side_ctr, side_cap / SLOT_SIDE_PARTS}, status, parts);
}

void slots_summary(hipStream_t s, SlotView b, uint64_t nrows, void *summ,
                   const unsigned int *status, uint32_t block_rows) {
  if (!nrows) return;
  uint2 *pre = static_cast<uint2 *>(summ), *suf = pre + nrows;
  uint4 *pre2 = reinterpret_cast<uint4 *>(pre + 2 * nrows), *suf2 = pre2 + nrows;
  // 32-row blocks take two rows per lane (16 lanes, within one DPP row): a
  // wave then covers 128 rows, and fewer waves meet a row of several ranges,
  // which sends the whole wave down the general path
#define DV_SSUMM(L_, K_)                                                                   \
  hipLaunchKernelGGL((k_slot_summary<L_, K_>), dim3(grid_for(nrows, BLOCK * K_)), dim3(BLOCK), \
                     0, s, slots_of(b), nrows, pre, suf, pre2, suf2, status)
  switch (block_rows) {
    case 32: DV_SSUMM(16, 2); break;
    case 64: DV_SSUMM(16, 4); break;
    case 128: DV_SSUMM(32, 4); break;
    default: DV_SSUMM(64, 4); break;
  }
#undef DV_SSUMM
}

bool slots_lane(hipStream_t s, SlotView a, SlotView b, uint64_t nrows, uint64_t b_nrows,
                uint32_t lane_max, uint32_t mid_max, uint32_t grid, void *c_slot, void *c_side,
                uint64_t side_cap, unsigned long long *side_ctr, uint32_t *mid_list,
                uint32_t *big_list, OpCounters *ctr, unsigned int *status, const void *summ,
                uint32_t summ_min, uint32_t summ_rows, uint32_t plus_i,
                unsigned long long *parts, uint32_t batch, ChainView b_chains,
                ChainView c_chains, hipEvent_t ev_start, hipEvent_t ev_stop) {
  if (!nrows) return true;
  const Chains bc{static_cast<const uint8_t *>(b_chains.seg),
                  static_cast<const uint2 *>(b_chains.chunk)};
  uint8_t *c_seg = static_cast<uint8_t *>(const_cast<void *>(c_chains.seg));
  uint2 *c_chunk = static_cast<uint2 *>(const_cast<void *>(c_chains.chunk));
  const SideAlloc sa{side_ctr, side_cap / SLOT_SIDE_PARTS};
  const uint2 *pre = static_cast<const uint2 *>(summ);
  const uint2 *suf = pre ? pre + b_nrows : nullptr;
  const uint4 *pre2 = pre ? reinterpret_cast<const uint4 *>(pre + 2 * b_nrows) : nullptr;
  const uint4 *suf2 = pre ? pre2 + b_nrows : nullptr;
  const uint32_t sh = (uint32_t)__builtin_ctz(summ_rows ? summ_rows : 1);
  // with events, the launch records them itself as it starts and ends (the
  // kernel's own time; an event recorded as a packet of its own holds the
  // queue for a few microseconds more)
#define DV_SLANE(K_)                                                                         \
  do {                                                                                       \
    if (ev_start)                                                                            \
      hipExtLaunchKernelGGL(K_, dim3(grid), dim3(BLOCK), 0, s, ev_start, ev_stop, 0u,        \
                            slots_of(a), slots_of(b), nrows, lane_max, mid_max,              \
                            static_cast<uint2 *>(c_slot), static_cast<uint2 *>(c_side), sa,  \
                            mid_list, big_list, ctr, status, pre, suf, pre2, suf2, summ_min, \
                            sh, plus_i, parts, bc, c_seg, c_chunk);                          \
    else                                                                                     \
      hipLaunchKernelGGL(K_, dim3(grid), dim3(BLOCK), 0, s, slots_of(a), slots_of(b), nrows, \
                         lane_max, mid_max, static_cast<uint2 *>(c_slot),                    \
                         static_cast<uint2 *>(c_side), sa, mid_list, big_list, ctr, status,  \
                         pre, suf, pre2, suf2, summ_min, sh, plus_i, parts, bc, c_seg,       \
                         c_chunk);                                                           \
  } while (0)
  // the one-launch forms (chains, narrow slices) leave C's chains to
  // slots_chains: their deferred rows are written after their chunks
  const bool narrow = !bc.seg && !pre && batch <= 4;
  if (bc.seg || narrow) {
    c_seg = nullptr;
    c_chunk = nullptr;
  }
  if (bc.seg) DV_SLANE(k_slot_fast<0>);
  else if (pre) DV_SLANE(k_slot_lane<true>);
  else if (batch <= 2) DV_SLANE(k_slot_fast<2>);
  else if (batch <= 4) DV_SLANE(k_slot_fast<4>);
  else DV_SLANE(k_slot_lane<false>);
#undef DV_SLANE
  return c_seg != nullptr;
}

void slots_groups(hipStream_t s, SlotView a, SlotView b, uint32_t lanes1, uint32_t lds1,
                  uint32_t grid1, const uint32_t *rows1, const unsigned int *n1, uint32_t lds2,
                  uint32_t grid2, const uint32_t *rows2, const unsigned int *n2,
                  uint32_t rows_per_group, void *c_slot, void *c_side, uint64_t side_cap,
                  unsigned long long *side_ctr, OpCounters *ctr, unsigned int *status,
                  uint32_t plus_i) {
  uint2 *cs = static_cast<uint2 *>(c_slot), *cd = static_cast<uint2 *>(c_side);
  const SideAlloc sa{side_ctr, side_cap / SLOT_SIDE_PARTS};
  const size_t lds2_bytes = (size_t)(BLOCK / 64) * lds2 * 8;
#define DV_SGRP(G_)                                                                             \
  hipLaunchKernelGGL((k_slot_groups<G_>), dim3(grid1 + grid2), dim3(BLOCK),                     \
                     std::max((size_t)(BLOCK / G_) * lds1 * 8, lds2_bytes), s, slots_of(a),      \
                     slots_of(b), lds1, lds2, rows_per_group, rows1, n1, grid1, rows2, n2, grid2, \
                     cs, cd, sa, ctr, status, plus_i)
  switch (lanes1) {
    case 8: DV_SGRP(8); break;
    case 16: DV_SGRP(16); break;
    case 32: DV_SGRP(32); break;
    default: DV_SGRP(64); break;
  }
#undef DV_SGRP
}

void slots_reduce_parts(hipStream_t s, const unsigned long long *parts, uint32_t per_op,
                        const PartBlocks &used, const unsigned long long *side_ctr,
                        uint32_t nops, OpCounters *ctr) {
  if (!nops) return;
  hipLaunchKernelGGL(k_reduce_parts, dim3(nops), dim3(BLOCK), 0, s, parts, per_op, used,
                     side_ctr, ctr);
}

void slots_scan_rowmap(hipStream_t s, uint32_t sb, SlotView a, uint64_t n, uint32_t *cnt,
                       void *rowmap, unsigned long long *tile_sums, OpCounters *ctr,
                       unsigned int *status) {
  if (n == 0) {
    fill_u32(s, static_cast<uint32_t *>(rowmap), sb / 4, 0u);
    return;
  }
  const uint64_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  hipLaunchKernelGGL(k_slot_tile_counts, dim3((uint32_t)ntiles), dim3(BLOCK), 0, s, slots_of(a),
                     n, cnt, tile_sums, status);
  scan_tiles_rows(s, sb, cnt, n, rowmap, tile_sums, ctr, ~0ull, status, nullptr, 0, nullptr);
}

void slots_emit(hipStream_t s, SlotView a, uint64_t n, uint32_t sb, const void *rowmap,
                uint32_t *entries) {
  if (!n) return;
  if (sb == 4)
    hipLaunchKernelGGL(k_slot_emit<uint32_t>, dim3(grid_for(n)), dim3(BLOCK), 0, s, slots_of(a), n,
                       static_cast<const uint32_t *>(rowmap), entries);
  else
    hipLaunchKernelGGL(k_slot_emit<uint64_t>, dim3(grid_for(n)), dim3(BLOCK), 0, s, slots_of(a), n,
                       static_cast<const uint64_t *>(rowmap), entries);
}

uint32_t slot_lane_blocks_per_cu(bool summ, uint32_t batch, bool chains) {
  static std::atomic<int> cache[5] = {0, 0, 0, 0, 0};  // shared by every context and thread
  const int k = chains ? 4 : summ ? 3 : batch <= 2 ? 0 : batch <= 4 ? 1 : 2;
  if (!cache[k].load(std::memory_order_relaxed)) {
    int nb = 0;
    const hipError_t e =
        k == 4   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_slot_fast<0>, BLOCK, 0)
        : k == 3 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_slot_lane<true>, BLOCK, 0)
        : k == 0 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_slot_fast<2>, BLOCK, 0)
        : k == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_slot_fast<4>, BLOCK, 0)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_slot_lane<false>, BLOCK, 0);
    if (e != hipSuccess || nb < 1) {
      (void)hipGetLastError();
      nb = 4;
    }
    cache[k] = nb;
  }
  return (uint32_t)cache[k];
}

}  // namespace dv
