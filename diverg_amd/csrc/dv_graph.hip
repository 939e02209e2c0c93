// Character-level adjacency of a sequence graph, built on the device.
//
// Rows are loci.  Inside a node, locus l has the single column l + 1; the
// last locus of node k has the first locus of each out-neighbour, ascending.
// Runs of consecutive columns become one range, exactly as the reference's
// Basic -> Range conversion does (crs_matrix.hpp:197-242) on the rows
// adjacency.hpp:170-228 produces.
//
// The host computes each row count's ingredients and the exact output size
// (runs per node), so the build needs no readback for allocation: count per
// locus -> row-pointer scan -> write.  Threads take chunks of
// ADJ_CHUNK consecutive loci, find their node once by binary search over the
// node starts, then walk forward.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "dv_graph.hpp"
#include "dv_runtime.hpp"

namespace dv {

namespace {

constexpr uint32_t ADJ_CHUNK = 16;
constexpr uint32_t ADJ_BLOCK = 256;

// node k with start[k] <= l < start[k + 1] (start[0] = 0 <= l < start[m])
__device__ __forceinline__ uint64_t node_of(const uint64_t *__restrict__ start, uint64_t m,
                                            uint64_t l) {
  uint64_t lo = 0, hi = m;
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if (start[mid] <= l) lo = mid;
    else hi = mid;
  }
  return lo;
}

template <bool WRITE, typename S>
__global__ __launch_bounds__(ADJ_BLOCK) void k_adjacency(
    const uint64_t *__restrict__ start, uint64_t m, const uint64_t *__restrict__ eoff,
    const uint32_t *__restrict__ etgt, uint64_t n, uint32_t *__restrict__ cnt,
    const S *__restrict__ rowmap, uint32_t *__restrict__ entries,
    const unsigned int *__restrict__ status) {
  if (WRITE && *status) return;  // the scan disagreed with the sized output
  const uint64_t chunks = (n + ADJ_CHUNK - 1) / ADJ_CHUNK;
  for (uint64_t c = (uint64_t)blockIdx.x * ADJ_BLOCK + threadIdx.x; c < chunks;
       c += (uint64_t)gridDim.x * ADJ_BLOCK) {
    const uint64_t l0 = c * ADJ_CHUNK, l1 = min(n, l0 + ADJ_CHUNK);
    uint64_t k = node_of(start, m, l0);
    uint64_t kend = start[k + 1];
    for (uint64_t l = l0; l < l1; ++l) {
      while (kend <= l) {
        ++k;
        kend = start[k + 1];
      }
      if (l + 1 < kend) {  // inside the node
        if (WRITE) {
          const uint64_t p = rowmap[l];
          *reinterpret_cast<uint2 *>(entries + p) = make_uint2((uint32_t)(l + 1), (uint32_t)(l + 1));
        } else {
          cnt[l] = 2;
        }
        continue;
      }
      // last locus of node k: runs of the out-neighbours' first loci
      const uint64_t p = WRITE ? (uint64_t)rowmap[l] : 0;
      uint32_t runs = 0;
      uint64_t pend = 0;
      for (uint64_t q = eoff[k]; q < eoff[k + 1]; ++q) {
        const uint64_t col = start[etgt[q]];
        if (runs && col == pend + 1) {
          pend = col;
        } else {
          if (WRITE && runs) entries[p + 2 * runs - 1] = (uint32_t)pend;
          if (WRITE) entries[p + 2 * runs] = (uint32_t)col;
          ++runs;
          pend = col;
        }
      }
      if (WRITE && runs) entries[p + 2 * runs - 1] = (uint32_t)pend;
      if (!WRITE) cnt[l] = 2 * runs;
    }
  }
}

uint32_t adj_grid(uint64_t n) {
  const uint64_t chunks = (n + ADJ_CHUNK - 1) / ADJ_CHUNK;
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((chunks + ADJ_BLOCK - 1) / ADJ_BLOCK, 65535));
}

}  // namespace

int Ctx::adjacency_dev(const Graph &g, uint64_t lower, uint64_t upper, uint32_t sb, DevMat &out) {
  out = DevMat{};
  const uint64_t nn = g.len.size();
  if (upper == 0) upper = nn;
  if (lower > upper || upper > nn || (lower == upper && nn != 0))
    return fail("adjacency_matrix: need lower < upper <= nodes");
  const uint64_t m = upper - lower;
  const uint64_t base = g.charorder[lower];
  const uint64_t n = g.charorder[upper] - base;  // loci
  // local node starts, out-edges, exact sizes
  std::vector<uint64_t> start(m + 1), eoff(m + 1);
  std::vector<uint32_t> etgt;
  etgt.reserve(g.eoff[upper] - g.eoff[lower]);
  uint64_t ranges = 0, nnz = 0;
  for (uint64_t k = 0; k <= m; ++k) start[k] = g.charorder[lower + k] - base;
  for (uint64_t k = 0; k < m; ++k) {
    const uint64_t u = lower + k;
    eoff[k] = etgt.size();
    uint64_t runs = 0, pend = 0;
    for (uint64_t q = g.eoff[u]; q < g.eoff[u + 1]; ++q) {
      const uint64_t v = g.etgt[q];
      if (v < lower || v >= upper)
        return fail("adjacency_matrix: an edge leaves the node range (it must cover whole components)");
      etgt.push_back((uint32_t)(v - lower));
      const uint64_t col = start[v - lower];
      if (!(runs && col == pend + 1)) ++runs;
      pend = col;
    }
    ranges += (g.len[u] - 1) + runs;
    nnz += (g.len[u] - 1) + (g.eoff[u + 1] - g.eoff[u]);
  }
  eoff[m] = etgt.size();
  if (sb == 4 && 2 * ranges > 0xFFFFFFFFull)
    return fail("adjacency_matrix: entry count exceeds a 32-bit rowmap", DV_ERR_OVERFLOW);

  out.nrows = out.ncols = n;
  out.sb = sb;
  out.n_entries = 2 * ranges;
  out.nnz = nnz;
  out.owned = true;
  out.rowmap = dalloc((uint64_t)sb * (n + 1));
  out.entries = static_cast<uint32_t *>(dalloc(4 * out.n_entries));
  auto *d_start = static_cast<uint64_t *>(dalloc(8 * (m + 1)));
  auto *d_eoff = static_cast<uint64_t *>(dalloc(8 * (m + 1)));
  auto *d_etgt = static_cast<uint32_t *>(dalloc(4 * std::max<size_t>(1, etgt.size())));
  void *buf[2] = {nullptr, nullptr};
  const uint64_t sz[2] = {4 * std::max<uint64_t>(1, n), 8 * scan_words(n)};
  auto cleanup = [&]() {
    dfree(d_start);
    dfree(d_eoff);
    dfree(d_etgt);
  };
  if (!out.rowmap || !out.entries || !d_start || !d_eoff || !d_etgt || !scratch_carve(sz, buf, 2)) {
    cleanup();
    release(out);
    return fail("device allocation failed (adjacency)", DV_ERR_NOMEM);
  }
  bool ok = hip(hipMemcpyAsync(d_start, start.data(), 8 * (m + 1), hipMemcpyHostToDevice, stream),
                "adjacency upload") &&
            hip(hipMemcpyAsync(d_eoff, eoff.data(), 8 * (m + 1), hipMemcpyHostToDevice, stream),
                "adjacency upload") &&
            (etgt.empty() || hip(hipMemcpyAsync(d_etgt, etgt.data(), 4 * etgt.size(),
                                                hipMemcpyHostToDevice, stream),
                                 "adjacency upload"));
  uint32_t status = 0;
  OpCounters h{};
  if (ok) {
    auto *cnt = static_cast<uint32_t *>(buf[0]);
    auto *tiles = static_cast<unsigned long long *>(buf[1]);
    fill_u32(stream, d_status, 1, 0u);
    zero_counters(stream, d_ctr);
    if (n) {
      if (sb == 4)
        hipLaunchKernelGGL((k_adjacency<false, uint32_t>), dim3(adj_grid(n)), dim3(ADJ_BLOCK), 0,
                           stream, d_start, m, d_eoff, d_etgt, n, cnt, nullptr, nullptr, d_status);
      else
        hipLaunchKernelGGL((k_adjacency<false, uint64_t>), dim3(adj_grid(n)), dim3(ADJ_BLOCK), 0,
                           stream, d_start, m, d_eoff, d_etgt, n, cnt, nullptr, nullptr, d_status);
    }
    scan_rowmap(stream, sb, cnt, n, out.rowmap, tiles, d_ctr, out.n_entries, d_status, nullptr,
                  0, nullptr);
    if (n) {
      if (sb == 4)
        hipLaunchKernelGGL((k_adjacency<true, uint32_t>), dim3(adj_grid(n)), dim3(ADJ_BLOCK), 0,
                           stream, d_start, m, d_eoff, d_etgt, n, nullptr,
                           static_cast<const uint32_t *>(out.rowmap), out.entries, d_status);
      else
        hipLaunchKernelGGL((k_adjacency<true, uint64_t>), dim3(adj_grid(n)), dim3(ADJ_BLOCK), 0,
                           stream, d_start, m, d_eoff, d_etgt, n, nullptr,
                           static_cast<const uint64_t *>(out.rowmap), out.entries, d_status);
    }
    ok = hip(hipGetLastError(), "adjacency launch") && read_counters(h, &status);
  }
  cleanup();
  if (!ok) {
    release(out);
    return DV_ERR_DEVICE;
  }
  if (status || (n && h.total != out.n_entries)) {
    release(out);
    return fail("adjacency_matrix: device row counts disagree with the host's", DV_ERR_DEVICE);
  }
  return DV_OK;
}

}  // namespace dv
