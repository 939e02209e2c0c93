// Device-side operations of the library: identity, range SpAdd and range
// SpGEMM as enqueue-only halves (count + row-pointer scan, then write), and
// the exact forms that read the entry count back between the halves, as the
// reference's view-level drivers do (range_sparse.hpp:364-428, 2215-2281).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dv_runtime.hpp"

namespace dv {

int Ctx::identity_dev(uint64_t n, uint32_t sb, DevMat &out) {
  out = DevMat{};
  out.nrows = out.ncols = n;
  out.n_entries = 2 * n;
  out.nnz = n;  // range_sparse.hpp:108
  out.sb = sb;
  out.owned = true;
  out.rowmap = dalloc(sb * (n + 1));
  out.entries = static_cast<uint32_t *>(dalloc(4 * out.n_entries));
  if (!out.rowmap || !out.entries) {
    release(out);
    return fail("device allocation failed (identity)", DV_ERR_NOMEM);
  }
  identity(stream, sb, n, out.rowmap, out.entries);
  out.canon = 1;
  return hip(hipGetLastError(), "identity launch") ? DV_OK : DV_ERR_DEVICE;
}

bool Ctx::op_scratch(uint64_t n, OpScratch &s, bool staged) {
  const uint64_t n1 = n ? n : 1;
  const uint32_t KS = staged ? tuning.stage_entries : 0u;
  // the lane-tier grid is at most the resident blocks of 256 threads, at
  // most 8 per CU; parts for twice that
  const uint32_t parts_cap = 16 * std::max<uint32_t>(1, num_cu);
  void *buf[8];
  const uint64_t sz[8] = {4 * n1, 4ull * KS * n1, 4 * n1, 4 * n1, 4 * n1, 8 * scan_words(n),
                          8 * summary_words(n1), 16ull * parts_cap};
  if (!scratch_carve(sz, buf, 8)) return false;
  s.summ = buf[6];
  s.parts = static_cast<unsigned long long *>(buf[7]);
  s.parts_cap = parts_cap;
  s.rows = n1;
  s.cnt = static_cast<uint32_t *>(buf[0]);
  s.stage = static_cast<uint32_t *>(buf[1]);
  s.ovf = static_cast<uint32_t *>(buf[2]);
  s.mid = static_cast<uint32_t *>(buf[3]);
  s.big = static_cast<uint32_t *>(buf[4]);
  s.tmp = static_cast<unsigned long long *>(buf[5]);
  s.stage_entries = KS;
  return true;
}

GroupLaunch Ctx::group_launch(uint32_t lanes, uint64_t nstage) const {
  GroupLaunch g{};
  g.lanes = lanes;
  g.lds_words = std::max<uint32_t>(1, tuning.lds_words * lanes / 64);
  g.stage = tuning.stage_entries;
  g.rows_per_group = std::max<uint32_t>(1, tuning.rows_per_group);
  // four blocks per CU (four waves per SIMD; the group kernels fit eight):
  // a row's slice gathers are latency-bound, and an index of two ranges per
  // row hands its widest products' rows to the groups (the last product's
  // 64-lane count took 11.0 ms at one block per CU, 5.5 at four, the same
  // at eight).  Rows are taken in
  // a grid-stride loop, so an empty or short list costs no more on a large
  // grid (its blocks exit at once, or add nothing), and the grid is capped
  // by the rows there could be.  Always a multiple of 8: the kernels' XCD-
  // aware block order is a permutation of the grid only then.
  const uint64_t per_block = (uint64_t)(256 / lanes) * g.rows_per_group;  // 256-thread blocks
  const uint64_t want = (std::max<uint64_t>(1, nstage) + per_block - 1) / per_block;
  const uint64_t most = (uint64_t)std::max<uint32_t>(8, (num_cu + 7) / 8 * 8) * 4;
  g.grid = (uint32_t)std::min<uint64_t>((want + 7) / 8 * 8, most);
  g.nstage = nstage;
  return g;
}

// ---------------------------------------------------------------------------
// SpAdd halves (range_sparse.hpp:112-357)
void Ctx::enq_spadd_count(const DevMat &a, const DevMat &b, void *c_rowmap, uint32_t c_sb,
                          const OpScratch &s, OpCounters *ctr, uint64_t cap,
                          unsigned int *status) {
  spadd_count(stream, a.sb, view(a), view(b), s.cnt, status);
  scan_rowmap(stream, c_sb, s.cnt, a.nrows, c_rowmap, s.tmp, ctr, cap, status, nullptr, 0,
                nullptr);
}

void Ctx::enq_spadd_write(const DevMat &a, const DevMat &b, const void *c_rowmap,
                          uint32_t *c_entries, uint64_t cap, OpCounters *ctr,
                          const unsigned int *status) {
  spadd_write(stream, a.sb, view(a), view(b), c_rowmap, c_entries, status);
  range_nnz(stream, c_entries, cap, ctr, status);
}

// ---------------------------------------------------------------------------
// SpGEMM halves: the lane tier, the two group tiers over the rows it hands
// over, the row-pointer scan (which also copies the staged rows when C's
// entries buffer is given); then the staged copy if not done yet and the
// direct rewrite of rows whose output exceeds the stage.
void Ctx::enq_spgemm_count(const DevMat &a, const DevMat &b, void *c_rowmap, uint32_t c_sb,
                           const OpScratch &s, OpCounters *ctr, uint64_t cap,
                           unsigned int *status, hipEvent_t after_lane, uint32_t *c_entries,
                           uint32_t summ_rows, bool plus_i, hipEvent_t before_lane) {
  const uint64_t n = a.nrows;
  uint32_t nparts = 0;  // lane-tier blocks whose partial sums the scan adds up
  // the summary buffer of the scratch holds summary_words(s.rows) words
  const bool summaries = summ_rows && b.nrows <= s.rows && tuning.summary_span >= 2;
  if (n) {
    if (summaries) block_summary(stream, b.sb, view(b), s.summ, status, summ_rows);
    // persistent lane-tier grid of exactly the blocks the GPU holds at once:
    // no block waits for a slot (and runs alone at the end), and few blocks
    // keep the per-block counter atomics (nnz, slice ranges) uncontended
    const uint64_t resident = (uint64_t)num_cu * lane_blocks_per_cu(a.sb, summaries);
    const uint32_t lane_grid = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>({(n + 255) / 256, resident, (uint64_t)s.parts_cap}));
    if (before_lane) (void)hipEventRecord(before_lane, stream);
    spgemm_lane(stream, a.sb, view(a), view(b), s.stage_entries, tuning.lane_slice_max,
                tuning.small_slice_max, lane_grid, s.cnt, s.stage, s.ovf, s.mid, s.big, ctr,
                status, summaries ? s.summ : nullptr, tuning.summary_span, summ_rows,
                plus_i ? 1u : 0u, s.parts);
    nparts = lane_grid;
    if (after_lane) (void)hipEventRecord(after_lane, stream);
    GroupLaunch g = group_launch(tuning.lanes_small, n);
    g.plus_i = plus_i ? 1u : 0u;
    g.rows = s.mid;
    g.n_dev = &ctr->n_mid;
    g.cnt = s.cnt;
    g.stage_buf = s.stage;
    g.ovf_list = s.ovf;
    g.ctr = ctr;
    g.status = status;
    spgemm_group(stream, a.sb, view(a), view(b), g);
    const GroupLaunch g64 = group_launch(64, n);
    g.lanes = 64;
    g.lds_words = g64.lds_words;
    g.grid = g64.grid;
    g.rows = s.big;
    g.n_dev = &ctr->n_big;
    spgemm_group(stream, a.sb, view(a), view(b), g);
  } else {
    if (before_lane) (void)hipEventRecord(before_lane, stream);
    if (after_lane) (void)hipEventRecord(after_lane, stream);
  }
  scan_rowmap(stream, c_sb, s.cnt, n, c_rowmap, s.tmp, ctr, cap, status, s.stage,
              s.stage_entries, c_entries, s.parts, nparts);
}

void Ctx::enq_spgemm_write(const DevMat &a, const DevMat &b, const void *c_rowmap,
                           uint32_t *c_entries, const OpScratch &s, OpCounters *ctr,
                           const unsigned int *status, bool staged_copied, bool plus_i) {
  const uint64_t n = a.nrows;
  if (!n) return;
  if (!staged_copied)
    spgemm_emit_staged(stream, a.sb, n, c_rowmap, s.stage, s.stage_entries, c_entries, status);
  GroupLaunch g = group_launch(64, n);
  g.plus_i = plus_i ? 1u : 0u;
  g.rows = s.ovf;
  g.n_dev = &ctr->overflow;
  g.c_rowmap = c_rowmap;
  g.c_entries = c_entries;
  g.status = status;
  spgemm_group(stream, a.sb, view(a), view(b), g);
}

// ---------------------------------------------------------------------------
// exact forms: one readback between the halves sizes C.entries
int Ctx::spadd_dev(const DevMat &a, const DevMat &b, DevMat &c) {
  if (a.nrows != b.nrows || a.ncols != b.ncols || a.sb != b.sb)
    return fail("spadd: operands differ in shape or size type");
  const uint64_t n = a.nrows;
  c = DevMat{};
  c.nrows = n;
  c.ncols = a.ncols;
  c.sb = a.sb;
  c.owned = true;
  OpScratch s;
  const bool got = op_scratch(n, s);
  c.rowmap = dalloc(c.sb * (n + 1));
  if (!got || !c.rowmap) {
    release(c);
    return fail("device allocation failed (spadd)", DV_ERR_NOMEM);
  }
  zero_counters(stream, d_ctr);
  fill_u32(stream, d_status, 1, 0u);
  prof_mark(0);
  enq_spadd_count(a, b, c.rowmap, c.sb, s, d_ctr, ~0ull, d_status);
  OpCounters h{};
  uint32_t st = 0;
  if (!read_counters(h, &st)) {
    release(c);
    return DV_ERR_DEVICE;
  }
  if (st & 2) {
    release(c);
    return fail("spadd: entry count exceeds a 32-bit rowmap", DV_ERR_OVERFLOW);
  }
  c.n_entries = h.total;
  c.entries = static_cast<uint32_t *>(dalloc(4 * c.n_entries));
  if (!c.entries) {
    release(c);
    return fail("device allocation failed (spadd entries)", DV_ERR_NOMEM);
  }
  enq_spadd_write(a, b, c.rowmap, c.entries, c.n_entries, d_ctr, d_status);
  prof_mark(1);
  if (!read_counters(h, &st)) {
    release(c);
    return DV_ERR_DEVICE;
  }
  c.nnz = h.nnz;
  // canonical inputs give a canonical sum; otherwise it is checked when used
  c.canon = a.canon == 1 && b.canon == 1 ? 1 : 0;
  prof_add(3, 0, 1, mat_bytes(a) + mat_bytes(b) + mat_bytes(c), 0);
  stats_alg_bytes += mat_bytes(a) + mat_bytes(b) + mat_bytes(c);
  stats_n_spadd += 1;
  return DV_OK;
}

uint8_t Ctx::canonical(const DevMat &m) {
  if (m.canon) return m.canon;
  if (m.nrows == 0) return 1;
  fill_u32(stream, d_status, 1, 0u);
  check_canonical(stream, m.sb, view(m), d_status);
  if (!hip(hipMemcpyAsync(h_status, d_status, sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
           "canonical readback") ||
      !hip(hipStreamSynchronize(stream), "canonical synchronize"))
    return 0;
  return (*h_status & 4) ? 2 : 1;
}

// C = A.B.  The reference's accumulator gives the set union of the B rows
// named by each A row when B's rows are canonical; for a B with a row that is
// not, its window clips ranges (spgemm_window).
int Ctx::spgemm_dev(const DevMat &a, const DevMat &b, DevMat &c) {
  if (a.ncols != b.nrows) return fail("spgemm: a.ncols != b.nrows");
  if (a.sb != b.sb) return fail("spgemm: operands differ in size type");
  if (a.ncols > 0xFFFFFFFEull || b.ncols > 0xFFFFFFFEull)
    return fail("spgemm: ncols exceeds uint32_t max - 1", DV_ERR_OVERFLOW);
  const uint8_t cb = canonical(b);
  if (cb == 0) return DV_ERR_DEVICE;
  return cb == 2 ? spgemm_window(a, b, c) : spgemm_core(a, b, c);
}

int Ctx::spgemm_window(const DevMat &a, const DevMat &b, DevMat &c) {
  const uint64_t n = a.nrows;
  OpScratch s;
  DevMat p;
  p.nrows = n;
  p.ncols = b.ncols;
  p.sb = a.sb;
  p.owned = true;
  p.canon = 2;
  const bool got = op_scratch(n, s);
  p.rowmap = dalloc(p.sb * (n + 1));
  if (!got || !p.rowmap) {
    release(p);
    return fail("device allocation failed (spgemm window)", DV_ERR_NOMEM);
  }
  zero_counters(stream, d_ctr);
  fill_u32(stream, d_status, 1, 0u);
  window_pieces(stream, a.sb, view(a), view(b), s.cnt, nullptr, nullptr, d_status);
  scan_rowmap(stream, p.sb, s.cnt, n, p.rowmap, s.tmp, d_ctr, ~0ull, d_status, nullptr, 0,
              nullptr);
  OpCounters h{};
  uint32_t st = 0;
  if (!hip(hipGetLastError(), "window pieces launch") || !read_counters(h, &st)) {
    release(p);
    return DV_ERR_DEVICE;
  }
  if (st & 2) {
    release(p);
    return fail("spgemm: entry count exceeds a 32-bit rowmap", DV_ERR_OVERFLOW);
  }
  p.n_entries = h.total;
  p.entries = static_cast<uint32_t *>(dalloc(4 * std::max<uint64_t>(1, p.n_entries)));
  if (!p.entries) {
    release(p);
    return fail("device allocation failed (spgemm window)", DV_ERR_NOMEM);
  }
  window_pieces(stream, a.sb, view(a), view(b), nullptr, p.rowmap, p.entries, d_status);
  DevMat ident;
  int rc = identity_dev(n, a.sb, ident);
  if (!rc) rc = spgemm_core(ident, p, c);
  release(ident);
  release(p);
  return rc;
}

int Ctx::spgemm_core(const DevMat &a, const DevMat &b, DevMat &c) {
  const uint64_t n = a.nrows;
  c = DevMat{};
  c.nrows = n;
  c.ncols = b.ncols;
  c.sb = a.sb;
  c.owned = true;
  OpScratch s;
  const bool got = op_scratch(n, s);
  c.rowmap = dalloc(c.sb * (n + 1));
  if (!got || !c.rowmap) {
    release(c);
    return fail("device allocation failed (spgemm)", DV_ERR_NOMEM);
  }
  zero_counters(stream, d_ctr);
  fill_u32(stream, d_status, 1, 0u);
  prof_mark(0);
  enq_spgemm_count(a, b, c.rowmap, c.sb, s, d_ctr, ~0ull, d_status, prof ? pev[1] : nullptr,
                   nullptr, 0, false, prof ? pev[5] : nullptr);
  prof_mark(2);
  OpCounters h{};
  uint32_t st = 0;
  if (!hip(hipGetLastError(), "spgemm count launch") || !read_counters(h, &st)) {
    release(c);
    return DV_ERR_DEVICE;
  }
  if (st & 2) {
    release(c);
    return fail("spgemm: entry count exceeds a 32-bit rowmap", DV_ERR_OVERFLOW);
  }
  c.n_entries = h.total;
  c.nnz = h.nnz;
  c.entries = static_cast<uint32_t *>(dalloc(4 * c.n_entries));
  if (!c.entries) {
    release(c);
    return fail("device allocation failed (spgemm entries)", DV_ERR_NOMEM);
  }
  prof_mark(3);
  enq_spgemm_write(a, b, c.rowmap, c.entries, s, d_ctr, d_status, false);
  prof_mark(4);
  if (!hip(hipGetLastError(), "spgemm write launch")) {
    release(c);
    return DV_ERR_DEVICE;
  }
  {
    const uint64_t alg = mat_bytes(a) + mat_bytes(b) + mat_bytes(c);
    prof_add(0, 5, 1, alg, 8ull * h.slice);
    prof_add(2, 0, 5, 0, 0, false);
    prof_add(2, 1, 2, 0, 0);
    prof_add(1, 3, 4, 0, 0);
  }
  stats_alg_bytes += mat_bytes(a) + mat_bytes(b) + mat_bytes(c);
  stats_n_spgemm += 1;
  stats_overflow_rows += h.overflow;
  stats_mid_rows += h.n_mid;
  stats_big_rows += h.n_big;
  if (n) stats_max_ratio = std::max(stats_max_ratio, (double)c.n_entries / (2.0 * (double)n));
  c.canon = 1;
  return DV_OK;
}

// range_sparse.hpp:2376-2431: C = I; loop { if (k & 1) C = C.A; k >>= 1;
// if (k) A = A.A; else break; }.  `a` is not modified.
int Ctx::power_exact(const DevMat &a_in, uint32_t k, DevMat &out) {
  if (a_in.nrows != a_in.ncols) return fail("power: matrix is not square");
  DevMat a = a_in;
  a.owned = false;
  DevMat cm;
  int rc = identity_dev(a.ncols, a.sb, cm);
  if (rc) return rc;
  for (;;) {
    if (k & 1) {
      DevMat t;
      rc = spgemm_dev(cm, a, t);
      if (rc) break;
      release(cm);
      cm = t;
    }
    k >>= 1;
    if (k != 0) {
      DevMat t;
      rc = spgemm_dev(a, a, t);
      if (rc) break;
      release(a);
      a = t;
    } else {
      break;
    }
  }
  release(a);
  if (rc) {
    release(cm);
    return rc;
  }
  out = cm;
  return DV_OK;
}

// dindex.hpp:46-189 with one readback per operation.
int Ctx::dindex_exact(const DevMat &a, uint32_t dlo, uint32_t dup, DevMat &d) {
  if (!(dlo <= dup && dup != 0)) return fail("create_distance_index: need dlo <= dup and dup != 0");
  if (a.nrows != a.ncols) return fail("create_distance_index: matrix is not square");
  DevMat ident, rai, rad, raid;
  int rc = identity_dev(a.nrows, a.sb, ident);
  if (rc) return rc;
  rc = spadd_dev(ident, a, rai);
  release(ident);
  if (rc) return rc;
  if (dlo != 0) {
    rc = power_exact(a, dlo, rad);
    if (rc) {
      release(rai);
      return rc;
    }
  }
  rc = power_exact(rai, dup - dlo, raid);
  release(rai);
  if (rc) {
    release(rad);
    return rc;
  }
  if (dlo != 0) {
    rc = spgemm_dev(rad, raid, d);
    release(rad);
    release(raid);
    return rc;
  }
  d = raid;
  return DV_OK;
}

}  // namespace dv
