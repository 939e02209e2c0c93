// Planned builds: a distance-index build (or a power) is laid out on the host
// as a fixed sequence of operations over matrix slots, every SpGEMM output
// gets an entries buffer of a predicted capacity, and the whole sequence is
// enqueued on the stream with no host round trip.  The scans flag a capacity
// overflow on the device; every later kernel that reads operands then does
// nothing, and after the single readback the build is repeated in exact
// (read-back-per-operation) form with a larger prediction.
//
// Exponentiation order and results follow range_power_inplace
// (range_sparse.hpp:2376-2431) and create_distance_index (dindex.hpp:46-189).
// One deviation with no effect on any output array: the first product of a
// power chain, I.P, is not computed when P is itself an SpGEMM output (which
// is canonical, so I.P == P element for element); C starts as P.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <array>
#include <functional>
#include <map>
#include <mutex>

#include "dv_runtime.hpp"

namespace dv {

int Plan::add_slot(uint64_t nrows, uint64_t ncols, uint32_t sb) {
  PlanSlot s;
  s.nrows = nrows;
  s.ncols = ncols;
  s.sb = sb;
  slots.push_back(s);
  return (int)slots.size() - 1;
}

int Plan::external(const DevMat &m) {
  const int s = add_slot(m.nrows, m.ncols, m.sb);
  slots[s].rowmap = m.rowmap;
  slots[s].entries = m.entries;
  slots[s].cap = m.n_entries;
  slots[s].n_entries = m.n_entries;
  slots[s].nnz = m.nnz;
  slots[s].external = true;
  slots[s].canon = m.canon;
  return s;
}

int Plan::op(PlanKind kind, int a, int b, uint64_t nrows, uint64_t ncols, uint32_t sb) {
  const int c = add_slot(nrows, ncols, sb);
  slots[c].producer = (int)ops.size();
  slots[c].kind = kind;
  // how many rows a range of the result may cover: I and the input 1,
  // I + A one more, and a product the sum of its operands' spans less one
  if (kind == P_SPADD) slots[c].span = std::max(slots[a].span, slots[b].span) + 1;
  if (kind == P_SPGEMM) slots[c].span = slots[a].span + slots[b].span - 1;
  ops.push_back(PlanOp{kind, a, b, c});
  if (a >= 0) slots[a].last_use = (int)ops.size() - 1;
  if (b >= 0) slots[b].last_use = (int)ops.size() - 1;
  return c;
}

int Plan::identity(uint64_t n, uint32_t sb) { return op(P_IDENT, -1, -1, n, n, sb); }

int Plan::spadd(int a, int b) {
  return op(P_SPADD, a, b, slots[a].nrows, slots[a].ncols, slots[a].sb);
}

int Plan::spgemm(int a, int b) {
  return op(P_SPGEMM, a, b, slots[a].nrows, slots[b].ncols, slots[a].sb);
}

int Plan::spgemm_plus_i(int a, int b) {
  const int c = op(P_SPGEMM, a, b, slots[a].nrows, slots[b].ncols, slots[a].sb);
  ops.back().plus_i = true;
  slots[c].span += 1;  // rows of b + I cover one more column than b's
  return c;
}

// An addition chain for a power P^k: exponents 1 = e[0] < e[1] < ... <
// e[L] = k, step i forming e[i] = e[a[i]] + e[b[i]], i.e. the product of the
// powers e[a[i]] and e[b[i]] (a[i] <= b[i]).  Powers of one matrix commute,
// and Boolean products and canonical range rows are unique, so every chain
// gives the same arrays.
struct PowerChain {
  std::vector<uint32_t> e{1u};
  std::vector<uint16_t> a{0}, b{0};
  double cost = 0;
};

// The binary method's chain (range_sparse.hpp:2376-2431: squarings, and the
// accumulated product times the square at each further set bit).
static PowerChain binary_chain(uint32_t k) {
  PowerChain c;
  int acc = -1, sq = 0;  // indices into c.e
  for (uint32_t r = k;;) {
    if (r & 1) {
      if (acc < 0) {
        acc = sq;
      } else {
        c.e.push_back(c.e[acc] + c.e[sq]);
        c.a.push_back((uint16_t)acc);
        c.b.push_back((uint16_t)sq);
        acc = (int)c.e.size() - 1;
      }
    }
    r >>= 1;
    if (r == 0) break;
    c.e.push_back(2 * c.e[sq]);
    c.a.push_back((uint16_t)sq);
    c.b.push_back((uint16_t)sq);
    sq = (int)c.e.size() - 1;
  }
  return c;
}

// Estimated time (microseconds at the 5e6-locus graph) of the product
// P^x . P^y (x <= y) when P's ranges span `span1` rows: the slice width is
// the span of P^x, w = x (span1 - 1) + 1, and the lane kernel it selects
// costs about 36 us for narrow slices (w <= 4), 52 + 3.1 (w - 1) us through
// chains (the rows near breaks, which it defers, grow with w; the chain
// pass over P^y included), 125 us with interval summaries (the summary pass
// included), 60 us slot by slot.
double Plan::product_cost(uint32_t x, uint64_t span1) const {
  const uint64_t w = (uint64_t)x * (span1 ? span1 - 1 : 0) + 1;
  if (w <= 4) return 36.0;
  if (summary_span >= 2 && w >= summary_span) return 125.0;
  if (chain_span && w >= chain_span) return 52.0 + 3.1 * (double)(w - 1);
  return 60.0;
}

// The cheapest star chain (each step adds an earlier exponent to the last)
// for P^k under product_cost, no longer than the binary method's: a
// depth-first search, largest step first, bounded by the cheapest step times
// the steps still needed (the exponent at most doubles per step), limited to
// search_nodes nodes (about 12 ns each); the binary chain when nothing
// cheaper is found.  Chains are cached per (k, span, thresholds, limit), so
// a build shape searches once.
PowerChain Plan::power_chain(uint32_t k, uint64_t span1) const {
  PowerChain best = binary_chain(k);
  for (size_t i = 1; i < best.e.size(); ++i)
    best.cost += product_cost(std::min(best.e[best.a[i]], best.e[best.b[i]]), span1);
  if (k < 3 || schedule != 0) return best;
  static std::mutex mu;
  static std::map<std::array<uint64_t, 5>, PowerChain> cache;
  const std::array<uint64_t, 5> key{k, span1, chain_span, summary_span, search_nodes};
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  const size_t max_len = best.e.size() - 1;
  double minc = 1e30;
  for (uint32_t x : {1u, 2u, 3u, 4u}) minc = std::min(minc, product_cost(x, span1));
  minc = std::min(minc, 36.0);
  PowerChain cur;
  uint64_t nodes = 0;
  std::function<void()> dfs = [&]() {
    const uint32_t last = cur.e.back();
    if (last == k) {
      if (cur.cost < best.cost - 1e-9) best = cur;
      return;
    }
    if (++nodes > search_nodes) return;
    uint32_t need = 0;
    for (uint64_t v = last; v < k; v <<= 1) ++need;
    if (cur.e.size() - 1 + need > max_len || cur.cost + need * minc >= best.cost - 1e-9) return;
    // steps last + e[j], largest first (distinct sums only)
    const uint16_t li = (uint16_t)(cur.e.size() - 1);
    for (int j = (int)cur.e.size() - 1; j >= 0; --j) {
      const uint32_t x = cur.e[j];
      if ((uint64_t)last + x > k) continue;
      if (j + 1 < (int)cur.e.size() && cur.e[j + 1] == x) continue;
      const double c = product_cost(std::min(x, last), span1);
      cur.e.push_back(last + x);
      cur.a.push_back((uint16_t)j);
      cur.b.push_back(li);
      cur.cost += c;
      dfs();
      cur.cost -= c;
      cur.e.pop_back();
      cur.a.pop_back();
      cur.b.pop_back();
    }
  };
  dfs();
  std::lock_guard<std::mutex> lk(mu);
  cache.emplace(key, best);
  return best;
}

// C = P^k along the cheapest addition chain; with schedule 1 by the
// reference's binary exponentiation (range_sparse.hpp:2376-2431: I.P first
// when bit 0 is set and P is not an SpGEMM output, then the accumulated
// product times each further square).
int Plan::power(int p, uint32_t k) {
  const uint64_t n = slots[p].ncols;
  const uint32_t sb = slots[p].sb;
  if (k == 0) return identity(n, sb);
  if (schedule != 0) {
    int c = -1;
    for (;;) {
      if (k & 1) {
        if (c >= 0)
          c = spgemm(c, p);
        else if (slots[p].kind == P_SPGEMM)
          c = p;  // I.P == P for a canonical SpGEMM output
        else
          c = spgemm(identity(n, sb), p);
      }
      k >>= 1;
      if (k == 0) break;
      p = spgemm(p, p);
    }
    return c;
  }
  if (k == 1) {
    // I.P == P for a canonical SpGEMM output
    return slots[p].kind == P_SPGEMM ? p : spgemm(identity(n, sb), p);
  }
  const PowerChain ch = power_chain(k, slots[p].span);
  std::vector<int> at(ch.e.size(), p);
  for (size_t i = 1; i < ch.e.size(); ++i) {
    // the narrower power on the left: its ranges select the slices
    const int x = at[ch.a[i]], y = at[ch.b[i]];
    at[i] = ch.e[ch.a[i]] <= ch.e[ch.b[i]] ? spgemm(x, y) : spgemm(y, x);
  }
  return at.back();
}

// SpGEMMs that power(p, k) plans for a P of this kind and span: the chain's
// steps, I.P alone for k = 1 when P is not an SpGEMM output; with schedule
// 1, floor(log2 k) squarings, one product per further set bit, and I.P
// first when bit 0 is set and P is not an SpGEMM output.
uint32_t Plan::power_spgemms(uint32_t k, bool p_is_product, uint64_t span) const {
  if (k == 0) return 0;
  if (schedule != 0) {
    uint32_t sq = 0;
    while ((k >> (sq + 1)) != 0) ++sq;
    return sq + (uint32_t)__builtin_popcount(k) - 1 + ((k & 1) && !p_is_product ? 1 : 0);
  }
  if (k == 1) return p_is_product ? 0u : 1u;
  return (uint32_t)power_chain(k, span).e.size() - 1;
}

// D = A^d1 . (A + I)^(d2 - d1) (dindex.hpp:46-189), or (A + I)^d2 when d1 = 0.
// A and A + I commute, so with m = min(d1, d2 - d1) the same matrix is also
// (A (A + I))^m times A^(d1 - m) or (A + I)^(d2 - d1 - m): M = A (A + I) holds
// the walks of length 1 or 2, M^m those of length m .. 2m.  Boolean products
// and canonical range rows are unique, so both orders give the same arrays;
// the plan takes the one with fewer SpGEMMs (for d2 = 2 d1, as [300, 600],
// M^300: 12 SpGEMMs instead of 23).
int Plan::dindex(int a, uint32_t dlo, uint32_t dup) {
  const uint64_t n = slots[a].nrows;
  const uint32_t sb = slots[a].sb;
  int rai_slot = -1;
  auto rai = [&]() {  // A + I, planned when first needed
    if (rai_slot < 0) rai_slot = spadd(identity(n, sb), a);
    return rai_slot;
  };
  if (dlo == 0) return power(rai(), dup);
  const uint32_t b = dup - dlo;
  // spans: A, A + I one more, M = A (A + I) twice A's
  const uint64_t sa = slots[a].span;
  const uint32_t classic =
      power_spgemms(dlo, false, sa) + power_spgemms(b, false, sa + 1) + 1;
  const uint32_t m = std::min(dlo, b);
  const uint32_t rest = std::max(dlo, b) - m;
  const uint32_t paired =
      m == 0 ? ~0u
             : 1 + power_spgemms(m, true, 2 * sa) +
                   (rest ? power_spgemms(rest, false, dlo > b ? sa : sa + 1) + 1 : 0);
  if (paired < classic && schedule != 1) {
    // M = A (A + I) straight from A: the product kernels add each A range
    // itself, so A + I is only planned when a rest factor needs it
    const int mm = spgemm_plus_i(a, a);
    const int mp = power(mm, m);
    if (!rest) return mp;
    return spgemm(mp, dlo > b ? power(a, rest) : power(rai(), rest));
  }
  const int rad = power(a, dlo);
  const int raid = power(rai(), b);
  return spgemm(rad, raid);
}

// Pull a buffer pair from the context's cache (best fit) or allocate one.
int Ctx::buf_acquire(uint64_t rm_bytes, uint64_t cap) {
  int best = -1;
  for (int i = 0; i < (int)bufs.size(); ++i) {
    const PlanBuf &b = bufs[i];
    if (b.in_use || b.rm_bytes < rm_bytes || b.cap < cap) continue;
    if (best < 0 || b.rm_bytes + 4 * b.cap < bufs[best].rm_bytes + 4 * bufs[best].cap) best = i;
  }
  if (best >= 0) {
    bufs[best].in_use = true;
    return best;
  }
  PlanBuf b;
  b.rm_bytes = rm_bytes;
  b.cap = std::max<uint64_t>(cap, 2);
  b.rowmap = dalloc(rm_bytes);
  b.entries = static_cast<uint32_t *>(dalloc(4 * b.cap));
  if (!b.rowmap || !b.entries) {
    dfree(b.rowmap);
    dfree(b.entries);
    return -1;
  }
  b.in_use = true;
  bufs.push_back(b);
  return (int)bufs.size() - 1;
}

void Ctx::buf_trim() {
  std::vector<PlanBuf> keep;
  for (const PlanBuf &b : bufs) {
    if (b.in_use) {
      keep.push_back(b);
    } else {
      dfree(b.rowmap);
      dfree(b.entries);
    }
  }
  bufs = keep;
}

uint64_t Ctx::predict_cap(const Plan &p, const PlanSlot &s) const {
  switch (s.kind) {
    case P_IDENT:
      return 2 * s.nrows;
    case P_SPADD:  // a merge never has more entries than its two inputs
      return p.slots[p.ops[s.producer].a].cap + p.slots[p.ops[s.producer].b].cap;
    default: {
      if (tuning.plan_mode == 2) return 1;  // forces the exact fallback
      const double f = std::max(1.0, plan_factor);
      return (uint64_t)std::ceil(2.0 * f * (double)s.nrows) + 64;
    }
  }
}

// Assign buffers to every slot (reusing those whose last reader comes
// earlier), reserve scratch and counters.  All allocation happens here, so
// plan_launch only enqueues kernels and can be captured into a graph.
int Ctx::plan_prepare(Plan &p) {
  const int nops = (int)p.ops.size();
  if (!plan_ctr_reserve(nops)) return fail("device allocation failed (plan counters)", DV_ERR_NOMEM);
  uint64_t nmax = 0;
  for (const PlanOp &o : p.ops) nmax = std::max(nmax, p.slots[o.c].nrows);
  if (!op_scratch(nmax, p.scratch))
    return fail("device allocation failed (plan scratch)", DV_ERR_NOMEM);
  for (int i = 0; i < nops; ++i) {
    PlanSlot &c = p.slots[p.ops[i].c];
    c.cap = predict_cap(p, c);
    c.buf = buf_acquire((uint64_t)c.sb * (c.nrows + 1), c.cap);
    if (c.buf < 0) return fail("device allocation failed (plan buffers)", DV_ERR_NOMEM);
    c.rowmap = bufs[c.buf].rowmap;
    c.entries = bufs[c.buf].entries;
    // c.cap stays the prediction (a reused buffer may be larger)
    if (p.ops[i].kind == P_IDENT) {
      c.n_entries = 2 * c.nrows;
      c.nnz = c.nrows;
    }
    // buffers whose last reader is this operation go back to the cache
    for (int x = 0; x < (int)p.slots.size(); ++x) {
      PlanSlot &sl = p.slots[x];
      if (!sl.external && sl.buf >= 0 && sl.last_use == i && x != p.out && !sl.released) {
        bufs[sl.buf].in_use = false;
        sl.released = true;
      }
    }
  }
  return DV_OK;
}

// Enqueue the plan's kernels (no allocation, no synchronisation).
void Ctx::plan_launch(Plan &p, bool timing) {
  const int nops = (int)p.ops.size();
  const OpScratch &s = p.scratch;
  fill_u32(stream, reinterpret_cast<uint32_t *>(d_plan_ctr),
           (uint32_t)(nops * sizeof(OpCounters) / 4), 0u);
  fill_u32(stream, d_status, 1, 0u);
  // The plan's products are set unions, which is the reference's product
  // only for canonical right operands (its accumulator window clips the
  // ranges of other rows), and A (A + I) from A equals A . SpAdd(I, A) only
  // for a canonical A.  A caller's matrix not known to be canonical is
  // checked first; a non-canonical one sets bit 2 of the status, so no later
  // kernel runs, and the build is redone in the exact form.
  for (int x = 0; x < (int)p.slots.size(); ++x) {
    const PlanSlot &sl = p.slots[x];
    if (!sl.external || sl.canon == 1) continue;
    bool read = false;
    for (const PlanOp &o : p.ops) read = read || (o.kind == P_SPGEMM && (o.b == x || o.plus_i));
    if (read) check_canonical(stream, sl.sb, view(sl.as_mat()), d_status);
  }
  for (int i = 0; i < nops; ++i) {
    const PlanOp &o = p.ops[i];
    PlanSlot &c = p.slots[o.c];
    OpCounters *ctr = d_plan_ctr + i;
    if (timing) (void)hipEventRecord(plan_ev[5 * i], stream);
    if (o.kind == P_IDENT) {
      identity(stream, c.sb, c.nrows, c.rowmap, c.entries);
      if (timing)
        for (int e = 1; e < 5; ++e) (void)hipEventRecord(plan_ev[5 * i + e], stream);
      continue;
    }
    const DevMat a = p.slots[o.a].as_mat(), b = p.slots[o.b].as_mat();
    if (o.kind == P_SPADD) {
      enq_spadd_count(a, b, c.rowmap, c.sb, s, ctr, c.cap, d_status);
      if (timing)
        for (int e = 1; e < 4; ++e) (void)hipEventRecord(plan_ev[5 * i + e], stream);
      enq_spadd_write(a, b, c.rowmap, c.entries, c.cap, ctr, d_status);
    } else {
      const uint32_t summ_rows = tuning.summary_span && p.slots[o.a].span >= tuning.summary_span
                                     ? summary_block_rows(p.slots[o.a].span)
                                     : 0u;
      enq_spgemm_count(a, b, c.rowmap, c.sb, s, ctr, c.cap, d_status,
                       timing ? plan_ev[5 * i + 2] : nullptr, c.entries, summ_rows, o.plus_i,
                       timing ? plan_ev[5 * i + 1] : nullptr);
      if (timing) (void)hipEventRecord(plan_ev[5 * i + 3], stream);
      enq_spgemm_write(a, b, c.rowmap, c.entries, s, ctr, d_status, true, o.plus_i);
    }
    if (timing) (void)hipEventRecord(plan_ev[5 * i + 4], stream);
  }
}

// A fingerprint of everything a captured plan depends on: operations,
// buffers, shapes, scratch, counters and the launch tuning.
static uint64_t plan_key(const Plan &p, const dv_tuning &t, const void *ctr, const void *status) {
  uint64_t h = 1469598103934665603ull;
  auto mix = [&](uint64_t v) {
    h ^= v;
    h *= 1099511628211ull;
  };
  for (const PlanOp &o : p.ops) {
    mix(o.kind);
    mix(o.plus_i);
    mix((uint64_t)(int64_t)o.a);
    mix((uint64_t)(int64_t)o.b);
  }
  for (const PlanSlot &sl : p.slots) {
    mix((uint64_t)sl.rowmap);
    mix((uint64_t)sl.entries);
    mix(sl.cap);
    mix(sl.nrows);
    mix(sl.ncols);
    mix(sl.sb);
    mix(sl.n_entries);
  }
  const OpScratch &s = p.scratch;
  mix((uint64_t)s.cnt);
  mix((uint64_t)s.stage);
  mix((uint64_t)s.tmp);
  mix((uint64_t)s.parts);
  mix(s.stage_entries);
  const uint32_t *tw = reinterpret_cast<const uint32_t *>(&t);
  for (size_t i = 0; i < sizeof(dv_tuning) / 4; ++i) mix(tw[i]);
  mix((uint64_t)ctr);
  mix((uint64_t)status);
  return h;
}

// Enqueue every operation of the plan; no host synchronisation.  With
// use_graph the launches are captured once into a hipGraph and replayed while
// the fingerprint stays the same.
int Ctx::plan_enqueue(Plan &p) {
  int rc = plan_prepare(p);
  if (rc) return rc;
  const int nops = (int)p.ops.size();
  const bool timing = prof && plan_events_reserve(nops);
  if (!tuning.graphs || timing) {
    plan_launch(p, timing);
    return hip(hipGetLastError(), "plan launch") ? DV_OK : DV_ERR_DEVICE;
  }
  const uint64_t key = plan_key(p, tuning, d_plan_ctr, d_status);
  if (!graph_exec || key != graph_key) {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    graph_exec = nullptr;
    hipGraph_t g = nullptr;
    if (!hip(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "begin capture"))
      return DV_ERR_DEVICE;
    plan_launch(p, false);
    if (!hip(hipStreamEndCapture(stream, &g), "end capture")) return DV_ERR_DEVICE;
    const bool ok = hip(hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0), "instantiate");
    (void)hipGraphDestroy(g);
    if (!ok) {
      graph_exec = nullptr;
      return DV_ERR_DEVICE;
    }
    graph_key = key;
    graph_captures += 1;
  }
  return hip(hipGraphLaunch(graph_exec, stream), "graph launch") ? DV_OK : DV_ERR_DEVICE;
}

// Read the plan's counters back (one synchronisation) and fill in the slots.
int Ctx::plan_finish(Plan &p, uint32_t &status_out) {
  const int nops = (int)p.ops.size();
  bool ok = hip(hipMemcpyAsync(h_plan_ctr, d_plan_ctr, nops * sizeof(OpCounters),
                               hipMemcpyDeviceToHost, stream),
                "plan readback") &&
            hip(hipMemcpyAsync(h_status, d_status, sizeof(uint32_t), hipMemcpyDeviceToHost,
                               stream),
                "plan status") &&
            hip(hipStreamSynchronize(stream), "plan synchronize");
  if (!ok) return DV_ERR_DEVICE;
  status_out = *h_status;
  const bool timing = prof && (int)plan_ev.size() >= 5 * nops;
  for (int i = 0; i < nops; ++i) {
    const PlanOp &o = p.ops[i];
    PlanSlot &c = p.slots[o.c];
    const OpCounters &h = h_plan_ctr[i];
    if (o.kind != P_IDENT) {
      c.n_entries = h.total;
      c.nnz = h.nnz;
      const uint64_t alg = p.slots[o.a].bytes() + p.slots[o.b].bytes() + c.bytes();
      stats_alg_bytes += alg;
      if (o.kind == P_SPGEMM) {
        stats_n_spgemm += 1;
        stats_overflow_rows += h.overflow;
        stats_mid_rows += h.n_mid;
        stats_big_rows += h.n_big;
        if (c.nrows)
          stats_max_ratio = std::max(stats_max_ratio, (double)c.n_entries / (2.0 * (double)c.nrows));
      } else {
        stats_n_spadd += 1;
      }
      if (timing && !status_out) {
        // events: op start, before the lane kernel, after it, after the
        // count phase (group tiers, scan), after the write phase
        float t[4] = {0, 0, 0, 0};
        for (int e = 0; e < 4; ++e)
          (void)hipEventElapsedTime(&t[e], plan_ev[5 * i + e], plan_ev[5 * i + e + 1]);
        if (o.kind == P_SPGEMM) {
          prof_t[0].launches += 1;
          prof_t[0].total_ms += t[1];
          prof_t[0].alg_bytes += alg;
          prof_t[0].slice_bytes += 8ull * h.slice;
          prof_t[1].launches += 1;
          prof_t[1].total_ms += t[3];
          prof_t[2].launches += 1;
          prof_t[2].total_ms += t[0] + t[2];
        } else {
          prof_t[3].launches += 1;
          prof_t[3].total_ms += t[0] + t[1] + t[2] + t[3];
          prof_t[3].alg_bytes += alg;
        }
      }
    }
  }
  return DV_OK;
}

// Run a plan; on a capacity overflow fall back to `exact` and learn a larger
// capacity prediction.  `out` receives a copy of the output slot in
// exactly sized buffers.
int Ctx::plan_run(Plan &p, DevMat &out, const std::function<int(DevMat &)> &exact) {
  reset_stats();
  int rc = plan_enqueue(p);
  uint32_t st = 0;
  if (!rc) rc = plan_finish(p, st);
  // whatever happens, all plan buffers return to the cache
  auto release_all = [&]() {
    for (PlanSlot &x : p.slots)
      if (!x.external && x.buf >= 0 && !x.released) {
        bufs[x.buf].in_use = false;
        x.released = true;
      }
  };
  if (rc) {
    release_all();
    return rc;
  }
  if (st & 4) {  // a non-canonical input: the exact form
    release_all();
    reset_stats();
    if (p.ref_key_set) ref_order_shapes.insert(p.ref_key);
    rc = exact(out);
    stats_retried = 1;
    return rc;
  }
  if (st & 1) {
    release_all();
    plan_retries += 1;
    reset_stats();
    stats_max_ratio = 0;
    stats_retried = 1;
    rc = exact(out);
    if (!rc && tuning.plan_mode != 2) plan_factor = std::max(plan_factor, stats_max_ratio * 1.25);
    return rc;
  }
  // a scan's 32-bit check (2) holds only when no product failed: after a
  // failed one (the cases above, which the exact form checks again) the
  // counts the later scans read were never written
  if (st & 2) {
    release_all();
    return fail("entry count exceeds a 32-bit rowmap", DV_ERR_OVERFLOW);
  }
  stats_planned = 1;
  const PlanSlot &o = p.slots[p.out];
  out = DevMat{};
  out.nrows = o.nrows;
  out.ncols = o.ncols;
  out.sb = o.sb;
  out.n_entries = o.n_entries;
  out.nnz = o.nnz;
  out.owned = true;
  out.rowmap = dalloc(o.sb * (o.nrows + 1));
  out.entries = static_cast<uint32_t *>(dalloc(4 * o.n_entries));
  if (!out.rowmap || !out.entries) {
    release_all();
    release(out);
    return fail("device allocation failed (output)", DV_ERR_NOMEM);
  }
  bool ok = hip(hipMemcpyAsync(out.rowmap, o.rowmap, o.sb * (o.nrows + 1),
                               hipMemcpyDeviceToDevice, stream),
                "output rowmap copy") &&
            (o.n_entries == 0 ||
             hip(hipMemcpyAsync(out.entries, o.entries, 4 * o.n_entries,
                                hipMemcpyDeviceToDevice, stream),
                 "output entries copy"));
  release_all();
  if (!ok) {
    release(out);
    return DV_ERR_DEVICE;
  }
  return DV_OK;
}

bool Ctx::plan_ctr_reserve(int nops) {
  if (nops <= plan_ctr_cap) return true;
  if (d_plan_ctr) (void)hipFree(d_plan_ctr);
  if (h_plan_ctr) (void)hipHostFree(h_plan_ctr);
  d_plan_ctr = nullptr;
  h_plan_ctr = nullptr;
  plan_ctr_cap = 0;
  const int want = std::max(64, nops);
  if (hipMalloc((void **)&d_plan_ctr, want * sizeof(OpCounters)) != hipSuccess ||
      hipHostMalloc((void **)&h_plan_ctr, want * sizeof(OpCounters), hipHostMallocDefault) !=
          hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  plan_ctr_cap = want;
  return true;
}

bool Ctx::plan_events_reserve(int nops) {
  while ((int)plan_ev.size() < 5 * nops) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    plan_ev.push_back(e);
  }
  return true;
}

int Ctx::dindex_planned(const DevMat &a, uint32_t dlo, uint32_t dup, DevMat &d) {
  const std::array<uint64_t, 6> key{a.nrows, a.n_entries, a.nnz, a.sb, dlo, dup};
  // a shape whose input had a non-canonical row is built in the exact form
  // (the reference's order, its accumulator window for non-canonical right
  // operands)
  if (ref_order_shapes.count(key)) {
    reset_stats();
    stats_retried = 1;
    return dindex_exact(a, dlo, dup, d);
  }
  if (tuning.layout == 2 || (tuning.layout == 0 && !rcrs_shapes.count(key)))
    return dindex_slots(a, dlo, dup, d);
  Plan p;
  p.schedule = tuning.schedule;
  p.chain_span = tuning.chain_span;
  p.summary_span = tuning.summary_span;
  p.search_nodes = Plan::search_nodes_for(a.nrows);
  p.ref_key = key;
  p.ref_key_set = true;
  p.out = p.dindex(p.external(a), dlo, dup);
  return plan_run(p, d, [&](DevMat &o) { return dindex_exact(a, dlo, dup, o); });
}

// Launch-shape autotuner, in place of the reference's SparseConfig / ExecGrid
// heuristics (range_sparse_base.hpp:99-533): coordinate descent over the
// launch parameters, each candidate timed as a whole planned build of this
// exact shape (best of two runs), the winner cached per build shape.  Every
// candidate computes the same index; only its speed differs.
int Ctx::autotune_dindex(const DevMat &a, uint32_t dlo, uint32_t dup, dv_tuning &best) {
  const dv_tuning user = tuning;
  best = user;
  const std::array<uint64_t, 6> key{a.nrows, a.n_entries, a.nnz, a.sb, dlo, dup};
  auto copy_tuned = [](dv_tuning &to, const dv_tuning &from) {
    to.lds_words = from.lds_words;
    to.stage_entries = from.stage_entries;
    to.rows_per_group = from.rows_per_group;
    to.lanes_small = from.lanes_small;
    to.small_slice_max = from.small_slice_max;
    to.lane_slice_max = from.lane_slice_max;
    to.summary_span = from.summary_span;
    to.chain_span = from.chain_span;
  };
  auto hit = tuned.find(key);
  if (hit != tuned.end()) {
    copy_tuned(best, hit->second);
    return DV_OK;
  }
  auto timed = [&](const dv_tuning &t, double &ms) -> int {
    tuning = t;
    tuning.graphs = 0;
    tuning.autotune = 0;
    ms = 1e300;
    for (int rep = 0; rep < 2; ++rep) {
      DevMat o;
      if (!hip(hipStreamSynchronize(stream), "autotune synchronize")) return DV_ERR_DEVICE;
      const auto t0 = std::chrono::steady_clock::now();
      const int rc = dindex_planned(a, dlo, dup, o);
      const bool planned = stats_planned != 0;
      const double dt =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
      release(o);
      if (rc) return rc;
      autotune_builds += 1;
      if (planned) ms = std::min(ms, dt);
    }
    return DV_OK;
  };
  const uint32_t lds_max = (uint32_t)std::min<uint64_t>(1024, lds_per_block / 32);
  const std::vector<std::pair<uint32_t dv_tuning::*, std::vector<uint32_t>>> dims = {
      {&dv_tuning::stage_entries, {4, 8, 16}},
      {&dv_tuning::lanes_small, {8, 16, 32}},
      {&dv_tuning::small_slice_max, {32, 64, 128, 256}},
      {&dv_tuning::lane_slice_max, {1024, 4096, 16384}},
      {&dv_tuning::summary_span, {0, 16, 32, 64}},
      {&dv_tuning::chain_span, {0, 2, 3, 5, 9}},
      {&dv_tuning::lds_words, {lds_max / 4, lds_max / 2, lds_max}},
      {&dv_tuning::rows_per_group, {1, 4, 16}},
  };
  double best_ms = 0;
  int rc = timed(best, best_ms);
  for (size_t d = 0; !rc && d < dims.size(); ++d) {
    const uint32_t keep = best.*(dims[d].first);
    for (uint32_t v : dims[d].second) {
      if (v == keep || (v == 0 && dims[d].first != &dv_tuning::summary_span &&
                        dims[d].first != &dv_tuning::chain_span))
        continue;
      dv_tuning cand = best;
      cand.*(dims[d].first) = v;
      double ms = 0;
      rc = timed(cand, ms);
      if (rc) break;
      if (ms < best_ms) {
        best_ms = ms;
        best = cand;
      }
    }
  }
  tuning = user;
  if (rc) {
    best = user;
    return rc;
  }
  best.graphs = user.graphs;
  best.autotune = user.autotune;
  tuned[key] = best;
  return DV_OK;
}

int Ctx::dindex_dev(const DevMat &a, uint32_t dlo, uint32_t dup, DevMat &d) {
  if (!(dlo <= dup && dup != 0)) return fail("create_distance_index: need dlo <= dup and dup != 0");
  if (a.nrows != a.ncols) return fail("create_distance_index: matrix is not square");
  last_tuning = tuning;
  if (tuning.plan_mode == 1) {
    reset_stats();
    return dindex_exact(a, dlo, dup, d);
  }
  if (!tuning.autotune || tuning.plan_mode != 0) return dindex_planned(a, dlo, dup, d);
  dv_tuning best;
  int rc = autotune_dindex(a, dlo, dup, best);
  if (rc) return rc;
  const dv_tuning user = tuning;
  tuning = best;
  last_tuning = best;
  rc = dindex_planned(a, dlo, dup, d);
  tuning = user;
  return rc;
}

int Ctx::power_dev(const DevMat &a, uint32_t k, DevMat &out) {
  if (a.nrows != a.ncols) return fail("power: matrix is not square");
  if (tuning.plan_mode == 1) {
    reset_stats();
    return power_exact(a, k, out);
  }
  Plan p;
  p.out = p.power(p.external(a), k);
  return plan_run(p, out, [&](DevMat &o) { return power_exact(a, k, o); });
}

}  // namespace dv
