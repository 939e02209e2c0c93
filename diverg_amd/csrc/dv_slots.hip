// Planned index builds over slot-format intermediates.  The build's
// operations are those of Plan::dindex (the same products in the same
// order); the input is converted to slots once, every product reads and
// writes slots (one uint2 per row, rows of several ranges in a side array),
// and only the index itself is converted back to range CRS (per-row counts,
// row-pointer scan, copy).  A product thus needs no count pass, row-pointer
// scan or staged copy of its own.  Side arrays get a predicted capacity; a
// product that runs out flags it on the device, and after the single
// readback the build is repeated in exact form with a larger prediction.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <vector>

#include "dv_runtime.hpp"

namespace dv {

namespace {
SlotView view_of(const PlanSlot &s) { return SlotView{s.rowmap, s.entries}; }

// Rows of several ranges cost the slot form a dependent read each (the side
// array), where range CRS reads them with the rest of a slice: when they
// dominate (a product with more side entries than SIDE_SHARE of its rows),
// builds of that shape use range CRS.  Measured: slot form is faster up to
// ~0.6 (the 1e5-locus graph of [150, 450]: 0.70 vs 0.97 ms; the chromosome
// graph, up to 0.43: 29 vs 32 ms) and slower at 0.82 (a graph of 8% SNPs and
// 2% indels: 1.64 vs 1.40 ms).
constexpr double SIDE_SHARE = 0.75;
}  // namespace

bool Ctx::slot_parts_reserve(uint64_t words) {
  if (words <= slot_parts_words) return true;
  if (slot_parts) (void)hipFree(slot_parts);
  slot_parts = nullptr;
  slot_parts_words = 0;
  if (hipMalloc((void **)&slot_parts, 8 * words) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  slot_parts_words = words;
  return true;
}

int Ctx::dindex_slots(const DevMat &a, uint32_t dlo, uint32_t dup, DevMat &d) {
  reset_stats();
  Plan p;
  p.schedule = tuning.schedule;
  p.chain_span = tuning.chain_span;
  p.summary_span = tuning.summary_span;
  p.search_nodes = Plan::search_nodes_for(a.nrows);
  const int x = p.external(a);
  p.out = p.dindex(x, dlo, dup);
  auto exact = [&](DevMat &o) { return dindex_exact(a, dlo, dup, o); };
  const int nops = (int)p.ops.size();
  // every sum must be I + X (Plan::dindex plans no other)
  for (const PlanOp &o : p.ops)
    if (o.kind == P_SPADD && p.slots[o.a].kind != P_IDENT && p.slots[o.b].kind != P_IDENT)
      return plan_run(p, d, exact);
  if (p.out == x || nops == 0) return plan_run(p, d, exact);

  // records: one per operation, then the input conversion and the output
  // scan, then the build's status word (so one fill clears them all and one
  // copy reads them back)
  if (!plan_ctr_reserve(nops + 3)) return fail("device allocation failed (plan counters)", DV_ERR_NOMEM);
  OpCounters *c_out = d_plan_ctr + nops + 1;
  unsigned int *st_dev = reinterpret_cast<unsigned int *>(d_plan_ctr + nops + 2);
  uint64_t nmax = a.nrows;
  for (const PlanOp &o : p.ops) nmax = std::max(nmax, p.slots[o.c].nrows);
  OpScratch s;
  if (!op_scratch(nmax, s, false))
    return fail("device allocation failed (plan scratch)", DV_ERR_NOMEM);
  const uint32_t per_op = s.parts_cap;  // lane-tier blocks per operation at most
  // per-block partial sums of every operation, then the side-partition
  // counters of every operation and of the input conversion
  const uint64_t parts_words = 3ull * per_op * nops;
  const uint64_t ctr_words = (uint64_t)SLOT_SIDE_PARTS * (nops + 1);
  if (!slot_parts_reserve(parts_words + ctr_words))
    return fail("device allocation failed (plan scratch)", DV_ERR_NOMEM);
  unsigned long long *side_ctr = slot_parts + parts_words;

  // buffers: slot array (8 bytes per row) in `rowmap`, side array in `entries`
  auto release_all = [&]() {
    for (PlanSlot &sl : p.slots)
      if (sl.buf >= 0 && !sl.released) {
        bufs[sl.buf].in_use = false;
        sl.released = true;
      }
  };
  auto acquire = [&](PlanSlot &sl, uint64_t side_cap) -> bool {
    side_cap = std::min<uint64_t>(side_cap, 0xFFFFFFF0ull);  // 32-bit offsets in the slots
    sl.cap = side_cap;
    // a product's slots are followed by its chains
    const uint64_t rm = sl.kind == P_SPGEMM ? slot_chain_bytes(std::max<uint64_t>(1, sl.nrows))
                                            : 8 * std::max<uint64_t>(1, sl.nrows);
    const uint64_t ent = 2 * (side_cap + SLOT_SIDE_PAD);
    sl.chains = sl.kind == P_SPGEMM;
    sl.buf = buf_acquire(rm, ent);
    if (sl.buf < 0) {  // free the cached buffers no slot holds, then try once more
      buf_trim();
      sl.buf = buf_acquire(rm, ent);
    }
    if (sl.buf < 0) return false;
    sl.rowmap = bufs[sl.buf].rowmap;
    sl.entries = bufs[sl.buf].entries;
    return true;
  };
  // the input: every row of several ranges takes its ranges plus a header
  // (1.5 entries per range at most), with room for uneven partitions
  PlanSlot &in = p.slots[x];
  const uint64_t in_ranges = a.n_entries / 2;
  in.external = false;
  if (!acquire(in, 3 * in_ranges + 4096ull * SLOT_SIDE_PARTS)) {
    release_all();
    return fail("device allocation failed (plan buffers)", DV_ERR_NOMEM);
  }
  const double f = std::max(0.001, slot_side_factor);
  for (int i = 0; i < nops; ++i) {
    PlanSlot &c = p.slots[p.ops[i].c];
    const uint64_t cap = p.ops[i].kind == P_IDENT
                             ? 0
                             : std::max<uint64_t>(4096ull * SLOT_SIDE_PARTS,
                                                  (uint64_t)(f * (double)c.nrows));
    if (!acquire(c, cap)) {
      release_all();
      return fail("device allocation failed (plan buffers)", DV_ERR_NOMEM);
    }
    // buffers whose last reader is this operation go back to the cache
    for (int y = 0; y < (int)p.slots.size(); ++y) {
      PlanSlot &sl = p.slots[y];
      if (sl.buf >= 0 && sl.last_use == i && y != p.out && !sl.released) {
        bufs[sl.buf].in_use = false;
        sl.released = true;
      }
    }
  }
  d = DevMat{};
  d.nrows = a.nrows;
  d.ncols = a.ncols;
  d.sb = a.sb;
  d.owned = true;
  d.rowmap = dalloc(a.sb * (a.nrows + 1));
  if (!d.rowmap) {
    release_all();
    return fail("device allocation failed (output)", DV_ERR_NOMEM);
  }

  // B operands that products read through their chains (only those get a
  // chain pass)
  auto reads_chains = [&](const PlanOp &o) {
    const PlanSlot &A = p.slots[o.a], &B = p.slots[o.b];
    return o.kind == P_SPGEMM && tuning.chain_span && A.span >= tuning.chain_span &&
           B.kind == P_SPGEMM && (tuning.summary_span < 2 || A.span < tuning.summary_span);
  };
  std::vector<char> chain_reader(p.slots.size(), 0);
  for (const PlanOp &o : p.ops)
    if (reads_chains(o)) chain_reader[o.b] = 1;

  // ---- enqueue: the products (with tuning.graphs, captured once into a
  // hipGraph and replayed while everything they use is the same; timed
  // builds run eagerly), then the row-pointer scan into the output
  const bool timing = prof && plan_events_reserve(nops);
  PartBlocks used{};
  auto enqueue_products = [&]() {
    // the counters and the status, and (one launch) the side counters: the
    // per-block parts need no clearing, each launched block writes its own
    // (all parts are cleared too when the operations are too many to pass
    // their block counts)
    used = PartBlocks{};
    used.all = nops > (int)PART_BLOCKS_MAX_OPS ? 1u : 0u;
    fill2_u32(stream, reinterpret_cast<uint32_t *>(d_plan_ctr),
              (uint64_t)(nops + 3) * sizeof(OpCounters) / 4,
              reinterpret_cast<uint32_t *>(used.all ? slot_parts : side_ctr),
              used.all ? 2 * (parts_words + ctr_words) : 2 * ctr_words, 0u);
    slots_from_rcrs(stream, a.sb, view(a), in.rowmap, in.entries, in.cap,
                    side_ctr + (uint64_t)SLOT_SIDE_PARTS * nops, st_dev);
    for (int i = 0; i < nops; ++i) {
      const PlanOp &o = p.ops[i];
      PlanSlot &c = p.slots[o.c];
      OpCounters *ctr = d_plan_ctr + i;
      // timing: two events per SpAdd / SpGEMM, around the kernel it is
      // priced on (a product's lane launch records its own; every event
      // recorded as a packet of its own holds the queue for a few microseconds)
      if (o.kind == P_IDENT) {
        slots_identity(stream, c.nrows, c.rowmap);
        continue;
      }
      if (o.kind == P_SPADD) {
        const PlanSlot &xs = p.slots[p.slots[o.a].kind == P_IDENT ? o.b : o.a];
        if (timing) (void)hipEventRecord(plan_ev[2 * i], stream);
        slots_plus_identity(stream, view_of(xs), c.nrows, c.rowmap, c.entries, c.cap,
                            side_ctr + (uint64_t)SLOT_SIDE_PARTS * i, st_dev,
                            slot_parts + 3ull * per_op * i, per_op);
        if (!used.all) used.n[i] = slots_plus_identity_blocks(c.nrows, per_op);
        if (timing) (void)hipEventRecord(plan_ev[2 * i + 1], stream);
        continue;
      }
      const PlanSlot &A = p.slots[o.a], &B = p.slots[o.b];
      // B slices through B's chains, or through interval summaries, or slot by slot
      // (wide products, whose slices hold more chain ends, read summaries)
      const bool chains = reads_chains(o) && B.chains;
      uint32_t summ_rows = !chains && tuning.summary_span >= 2 && A.span >= tuning.summary_span
                               ? summary_block_rows(A.span)
                               : 0u;
      if (B.nrows > s.rows) summ_rows = 0;
      if (summ_rows) slots_summary(stream, view_of(B), B.nrows, s.summ, st_dev, summ_rows);
      const uint64_t n = c.nrows;
      // slice loads per round trip: about the widest A range
      const uint32_t batch = A.span <= 2 ? 2u : A.span <= 4 ? 4u : 8u;
      // the chains of C, for the products that read it through them: written
      // by the lane kernel, or by a pass once the group tiers have run
      const bool want_chains = c.chains && chain_reader[o.c];
      const uint64_t resident =
          (uint64_t)num_cu * slot_lane_blocks_per_cu(summ_rows != 0, batch, chains);
      const uint32_t grid =
          (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>({(n + 255) / 256, resident, per_op}));
      unsigned long long *op_side = side_ctr + (uint64_t)SLOT_SIDE_PARTS * i;
      const bool chains_done = slots_lane(
          stream, view_of(A), view_of(B), n, B.nrows, tuning.lane_slice_max, tuning.small_slice_max,
          grid, c.rowmap, c.entries, c.cap, op_side, s.mid, s.big, ctr, st_dev,
          summ_rows ? s.summ : nullptr, tuning.summary_span, summ_rows, o.plus_i ? 1u : 0u,
          slot_parts + 3ull * per_op * i, batch,
          chains ? slot_chains(B.rowmap, B.nrows) : ChainView{nullptr, nullptr},
          want_chains ? slot_chains(c.rowmap, c.nrows) : ChainView{nullptr, nullptr},
          timing && n ? plan_ev[2 * i] : nullptr, timing && n ? plan_ev[2 * i + 1] : nullptr);
      if (timing && !n) {
        (void)hipEventRecord(plan_ev[2 * i], stream);
        (void)hipEventRecord(plan_ev[2 * i + 1], stream);
      }
      if (!used.all) used.n[i] = n ? grid : 0u;
      const GroupLaunch gm = group_launch(tuning.lanes_small, n), gb = group_launch(64, n);
      slots_groups(stream, view_of(A), view_of(B), gm.lanes, gm.lds_words, gm.grid, s.mid, &ctr->n_mid,
                   gb.lds_words, gb.grid, s.big, &ctr->n_big, gm.rows_per_group, c.rowmap, c.entries,
                   c.cap, op_side, ctr, st_dev, o.plus_i ? 1u : 0u);
      if (want_chains && !chains_done)
        slots_chains(stream, c.rowmap, c.nrows, slot_chains(c.rowmap, c.nrows), st_dev);
    }
    slots_reduce_parts(stream, slot_parts, per_op, used, side_ctr, (uint32_t)nops, d_plan_ctr);
  };
  if (tuning.graphs && !timing) {
    // everything the captured launches read: operations, buffers, sizes,
    // spans (they choose the kernels), scratch, counters, tuning
    std::vector<uint64_t> key;
    key.reserve(16 * p.slots.size() + 8 * nops + 64);
    auto put = [&](uint64_t v) { key.push_back(v); };
    put((uint64_t)nops);
    for (const PlanOp &o : p.ops) {
      put(o.kind);
      put(o.plus_i);
      put((uint64_t)(int64_t)o.a);
      put((uint64_t)(int64_t)o.b);
      put((uint64_t)(int64_t)o.c);
    }
    for (const PlanSlot &sl : p.slots) {
      put((uint64_t)sl.rowmap);
      put((uint64_t)sl.entries);
      put(sl.cap);
      put(sl.nrows);
      put(sl.ncols);
      put(sl.sb);
      put(sl.span);
      put(sl.kind);
      put(sl.chains);
    }
    for (uint64_t v : {(uint64_t)a.rowmap, (uint64_t)a.entries, a.nrows, a.ncols, (uint64_t)a.sb,
                       a.n_entries, (uint64_t)s.mid, (uint64_t)s.big, (uint64_t)s.summ, s.rows,
                       (uint64_t)slot_parts, (uint64_t)side_ctr, (uint64_t)d_plan_ctr,
                       (uint64_t)st_dev, (uint64_t)per_op, (uint64_t)num_cu})
      put(v);
    const uint32_t *tw = reinterpret_cast<const uint32_t *>(&tuning);
    for (size_t k = 0; k < sizeof(dv_tuning) / 4; ++k) put(tw[k]);
    if (!slot_graph_exec || key != slot_graph_key) {
      if (slot_graph_exec) (void)hipGraphExecDestroy(slot_graph_exec);
      slot_graph_exec = nullptr;
      slot_graph_key.clear();
      hipGraph_t g = nullptr;
      bool ok = hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture");
      if (ok) {
        enqueue_products();
        ok = hip(hipStreamEndCapture(stream, &g), "end capture");
      }
      ok = ok && hip(hipGraphInstantiate(&slot_graph_exec, g, nullptr, nullptr, 0), "instantiate");
      if (g) (void)hipGraphDestroy(g);
      if (!ok) {
        slot_graph_exec = nullptr;
        release_all();
        release(d);
        return DV_ERR_DEVICE;
      }
      slot_graph_key = std::move(key);
      graph_captures += 1;
    }
    if (!hip(hipGraphLaunch(slot_graph_exec, stream), "graph launch")) {
      release_all();
      release(d);
      return DV_ERR_DEVICE;
    }
  } else {
    enqueue_products();
  }
  const PlanSlot &O = p.slots[p.out];
  slots_scan_rowmap(stream, d.sb, view_of(O), O.nrows, s.cnt, d.rowmap, s.tmp, c_out, st_dev);

  // ---- one readback
  bool ok = hip(hipGetLastError(), "slot plan launch") &&
            hip(hipMemcpyAsync(h_plan_ctr, d_plan_ctr, (nops + 3) * sizeof(OpCounters),
                               hipMemcpyDeviceToHost, stream),
                "plan readback") &&
            hip(hipStreamSynchronize(stream), "plan synchronize");
  if (!ok) {
    release_all();
    release(d);
    return DV_ERR_DEVICE;
  }
  // plan_mode 2 always takes the fallback (tests)
  const uint32_t st = *reinterpret_cast<const uint32_t *>(h_plan_ctr + nops + 2) |
                      (tuning.plan_mode == 2 ? 1u : 0u);
  // side use per row (all entries, and the fullest partition's share), for
  // the layout choice and the next prediction
  double use = 0, part_use = 0;
  for (int i = 0; i < nops; ++i)
    if (p.slots[p.ops[i].c].nrows) {
      const double n = (double)p.slots[p.ops[i].c].nrows;
      use = std::max(use, (double)h_plan_ctr[i].total / n);
      part_use = std::max(part_use, (double)h_plan_ctr[i].pad * SLOT_SIDE_PARTS / n);
    }
  const std::array<uint64_t, 6> key{a.nrows, a.n_entries, a.nnz, a.sb, dlo, dup};
  if (st & 12) {
    // the input does not suit the slot form: a row that is not canonical (4;
    // the reference's I + A merge of such a row is not the set union, and its
    // products clip such rows to their accumulator window) or rows of more
    // ranges than a side partition holds (8).  Built in the exact form (4)
    // or in range CRS (8), now and for later builds of this shape.
    release_all();
    release(d);
    if (st & 4) ref_order_shapes.insert(key);
    rcrs_shapes[key] = true;
    if ((st & 4) || tuning.layout == 2) {  // the exact form
      reset_stats();
      stats_retried = 1;
      return exact(d);
    }
    const int rc = dindex_planned(a, dlo, dup, d);
    stats_retried = 1;
    return rc;
  }
  if (st & 1) {
    release_all();
    release(d);
    plan_retries += 1;
    if (tuning.layout == 0 && tuning.plan_mode == 0 && use > SIDE_SHARE) {
      // rows of several ranges dominate, as a completed build would have
      // found (below): this build and later ones of its shape in range CRS,
      // without growing the side prediction build by build
      rcrs_shapes[key] = true;
      const int rc = dindex_planned(a, dlo, dup, d);
      stats_retried = 1;
      return rc;
    }
    const int rc = exact(d);
    stats_retried = 1;
    if (!rc && tuning.plan_mode != 2) {
      // a product ran out of side entries: predict more, up to
      // SIDE_FACTOR_MAX entries per row; a shape that needs more is built in
      // range CRS from now on
      constexpr double SIDE_FACTOR_MAX = 8.0;
      const double want = std::max(2.0 * f, 1.5 * part_use);
      if (want > SIDE_FACTOR_MAX)
        rcrs_shapes[key] = true;
      else
        slot_side_factor = want;
    }
    return rc;
  }
  // the output scan's 32-bit check (2) holds only for a build whose every
  // product completed: after a failed one (the cases above) the counts it
  // scanned were never written
  if (st & 2) {
    release_all();
    release(d);
    return fail("entry count exceeds a 32-bit rowmap", DV_ERR_OVERFLOW);
  }
  // ---- the index's entries
  const uint64_t total = h_plan_ctr[nops + 1].total;
  d.n_entries = total;
  d.entries = static_cast<uint32_t *>(dalloc(4 * std::max<uint64_t>(1, total)));
  if (!d.entries) {
    release_all();
    release(d);
    return fail("device allocation failed (output)", DV_ERR_NOMEM);
  }
  slots_emit(stream, view_of(O), O.nrows, d.sb, d.rowmap, d.entries);
  const PlanOp &last = p.ops[O.producer];
  d.nnz = last.kind == P_IDENT ? O.nrows : h_plan_ctr[O.producer].nnz;
  if (!hip(hipGetLastError(), "slot plan output")) {
    release_all();
    release(d);
    return DV_ERR_DEVICE;
  }

  // ---- statistics (bytes(M) = sizeof(S) (n + 1) + 8 ranges(M))
  std::vector<uint64_t> ranges(p.slots.size(), 0);
  ranges[x] = in_ranges;
  for (int i = 0; i < nops; ++i)
    ranges[p.ops[i].c] = p.ops[i].kind == P_IDENT ? p.slots[p.ops[i].c].nrows : h_plan_ctr[i].ranges;
  auto bytes = [&](int sl) { return (uint64_t)p.slots[sl].sb * (p.slots[sl].nrows + 1) + 8 * ranges[sl]; };
  int prev_ev = -1;  // the last timed event recorded, after a SpAdd or a lane
  for (int i = 0; i < nops; ++i) {
    const PlanOp &o = p.ops[i];
    if (o.kind == P_IDENT) continue;
    const OpCounters &h = h_plan_ctr[i];
    const uint64_t alg = bytes(o.a) + bytes(o.b) + bytes(o.c);
    stats_alg_bytes += alg;
    if (o.kind == P_SPGEMM) {
      stats_n_spgemm += 1;
      stats_mid_rows += h.n_mid;
      stats_big_rows += h.n_big;
      if (p.slots[o.c].nrows)
        stats_max_ratio = std::max(stats_max_ratio, (double)ranges[o.c] / (double)p.slots[o.c].nrows);
    } else {
      stats_n_spadd += 1;
    }
    if (timing) {
      // own kernel, and the kernels since the previous timed one (the
      // previous product's group tiers, this product's summaries)
      float own = 0, between = 0;
      (void)hipEventElapsedTime(&own, plan_ev[2 * i], plan_ev[2 * i + 1]);
      if (prev_ev >= 0) (void)hipEventElapsedTime(&between, plan_ev[prev_ev], plan_ev[2 * i]);
      prev_ev = 2 * i + 1;
      prof_t[2].total_ms += between;
      if (o.kind == P_SPGEMM) {
        prof_t[0].launches += 1;
        prof_t[0].total_ms += own;
        prof_t[0].alg_bytes += alg;
        prof_t[0].slice_bytes += 8ull * h.slice;
        prof_t[2].launches += 1;
      } else {
        prof_t[3].launches += 1;
        prof_t[3].total_ms += own;
        prof_t[3].alg_bytes += alg;
      }
    }
  }
  stats_planned = 1;
  stats_layout = 0;
  release_all();
  // (SIDE_SHARE: later builds of a shape whose rows of several ranges
  // dominate use range CRS)
  if (tuning.layout == 0 && use > SIDE_SHARE)
    rcrs_shapes[{a.nrows, a.n_entries, a.nnz, a.sb, dlo, dup}] = true;
  return DV_OK;
}

}  // namespace dv
