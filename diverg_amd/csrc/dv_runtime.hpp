// Device context and device-resident matrices of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <functional>
#include <map>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include "dv_internal.hpp"
#include "rcrs_kernels.hpp"

namespace dv {

struct Graph;

// A device-resident rCRS matrix.  `owned` matrices were allocated from the
// context's stream-ordered pool; others alias caller memory.
struct DevMat {
  uint64_t nrows = 0, ncols = 0, n_entries = 0;
  uint64_t nnz = 0;  // exact, 64-bit
  uint32_t sb = 8;
  void *rowmap = nullptr;
  uint32_t *entries = nullptr;
  bool owned = false;
  // 1: every row canonical (ranges ascending, a gap of at least one column
  // between neighbours), 2: some row is not, 0: not known
  uint8_t canon = 0;
};

uint64_t mat_bytes(const DevMat &m);
DView view(const DevMat &m);

// Scratch of one SpAdd / SpGEMM (carved from the context's arena).
struct OpScratch {
  uint32_t *cnt = nullptr;    // per-row entry counts
  uint32_t *stage = nullptr;  // staged entries: range k of row r at uint2 k * n + r
  uint32_t *ovf = nullptr;    // rows whose output exceeds the stage
  uint32_t *mid = nullptr;    // rows for the medium groups
  uint32_t *big = nullptr;    // rows for the 64-lane groups
  unsigned long long *tmp = nullptr;  // scan tile words
  void *summ = nullptr;               // block summaries of the right operand
  unsigned long long *parts = nullptr;  // lane-tier per-block (nnz, slice) sums
  uint32_t parts_cap = 0;               // blocks that parts has room for
  uint64_t rows = 0;                    // rows the per-row buffers (and summ) hold
  uint32_t stage_entries = 0;
};

// ---------------------------------------------------------------------------
// planned builds (dv_plan.hip)
enum PlanKind { P_EXT, P_IDENT, P_SPADD, P_SPGEMM };

struct PlanSlot {
  uint64_t nrows = 0, ncols = 0;
  uint32_t sb = 8;
  PlanKind kind = P_EXT;
  void *rowmap = nullptr;
  uint32_t *entries = nullptr;
  uint64_t cap = 0;         // entries capacity
  uint64_t n_entries = 0;   // known after the readback (static for I and inputs)
  uint64_t nnz = 0;
  uint64_t span = 1;        // rows an A range of this matrix may cover (static bound)
  int producer = -1;        // index of the op writing the slot
  int last_use = -1;        // index of the last op reading the slot
  int buf = -1;             // context buffer backing the slot
  bool external = false;    // caller memory
  bool released = false;
  bool chains = false;      // slot form: the buffer also holds the matrix's chains
  uint8_t canon = 0;        // external: DevMat::canon of the caller's matrix
  DevMat as_mat() const {
    DevMat m;
    m.nrows = nrows;
    m.ncols = ncols;
    m.sb = sb;
    m.rowmap = rowmap;
    m.entries = entries;
    m.n_entries = n_entries;
    m.nnz = nnz;
    return m;
  }
  uint64_t bytes() const { return (uint64_t)sb * (nrows + 1) + 4 * n_entries; }
};

struct PlanOp {
  PlanKind kind;
  int a, b, c;
  bool plus_i = false;  // SpGEMM: c = a (b + I)
};

struct Plan {
  std::vector<PlanSlot> slots;
  std::vector<PlanOp> ops;
  int out = -1;
  OpScratch scratch;  // set by Ctx::plan_prepare
  int add_slot(uint64_t nrows, uint64_t ncols, uint32_t sb);
  int external(const DevMat &m);
  int op(PlanKind kind, int a, int b, uint64_t nrows, uint64_t ncols, uint32_t sb);
  int identity(uint64_t n, uint32_t sb);
  int spadd(int a, int b);
  int spgemm(int a, int b);
  int spgemm_plus_i(int a, int b);  // a (b + I)
  int power(int p, uint32_t k);
  int dindex(int a, uint32_t dlo, uint32_t dup);
  double product_cost(uint32_t x, uint64_t span1) const;
  struct PowerChain power_chain(uint32_t k, uint64_t span1) const;
  uint32_t power_spgemms(uint32_t k, bool p_is_product, uint64_t span) const;
  uint32_t schedule = 0;  // 0: fewer SpGEMMs; 1: the reference's order
  // the lane kernels' span thresholds (dv_tuning), for the cost of a product
  uint32_t chain_span = 0, summary_span = 0;
  // power chain search: nodes at most (a larger build affords a longer one)
  uint64_t search_nodes = 5000;
  static uint64_t search_nodes_for(uint64_t nrows) {
    return std::min<uint64_t>(20000, std::max<uint64_t>(5000, nrows / 20));
  }
  // index builds: the build shape, remembered when the input turns out not
  // to be canonical
  std::array<uint64_t, 6> ref_key{};
  bool ref_key_set = false;
};

// A rowmap + entries buffer pair kept by the context across builds.
struct PlanBuf {
  void *rowmap = nullptr;
  uint32_t *entries = nullptr;
  uint64_t rm_bytes = 0, cap = 0;
  bool in_use = false;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t num_cu = 256;
  uint32_t lds_per_block = 65536;
  dv_tuning tuning{512, 16, 1, 16, 64, 4096, 0, 0, 0, 32, 0, 0, 5};
  std::string last_error;

  // pool accounting
  std::unordered_map<void *, uint64_t> sizes;
  uint64_t cur_bytes = 0, peak_bytes = 0;

  // scratch arena reused by consecutive operations on the stream
  void *scratch = nullptr;
  uint64_t scratch_cap = 0;
  // Carve `n` buffers of the given byte sizes (256-byte aligned) from the
  // arena, growing it when needed; false on allocation failure.
  bool scratch_carve(const uint64_t *bytes, void **out, int n);

  // one-word readbacks
  OpCounters *d_ctr = nullptr, *h_ctr = nullptr;
  uint32_t *d_flag = nullptr, *h_flag = nullptr;
  unsigned int *d_status = nullptr, *h_status = nullptr;

  // planned builds
  std::vector<PlanBuf> bufs;
  double plan_factor = 2.0;  // predicted SpGEMM output entries per row, halved
  uint64_t plan_retries = 0;
  OpCounters *d_plan_ctr = nullptr, *h_plan_ctr = nullptr;
  int plan_ctr_cap = 0;
  // slot-format builds (dv_slots.hip): lane-tier partial sums of every
  // operation, and the predicted side-array entries per row
  unsigned long long *slot_parts = nullptr;
  uint64_t slot_parts_words = 0;
  double slot_side_factor = 1.0;
  // build shapes (rows, entries, nnz, size type, d1, d2) whose slot-form
  // products had many rows of several ranges: built in range CRS from then on
  std::map<std::array<uint64_t, 6>, bool> rcrs_shapes;
  // build shapes whose input had a non-canonical row: built in range CRS in
  // the reference's order (I + A by the reference merge) from then on
  std::set<std::array<uint64_t, 6>> ref_order_shapes;
  std::vector<hipEvent_t> plan_ev;
  // hipGraph of the last planned build (replayed while its fingerprint holds)
  hipGraphExec_t graph_exec = nullptr;
  // hipGraph of the last slot-form build's products (replayed while every
  // pointer, size and launch parameter they use, slot_graph_key, holds)
  hipGraphExec_t slot_graph_exec = nullptr;
  std::vector<uint64_t> slot_graph_key;
  uint64_t graph_key = 0;
  uint64_t graph_captures = 0;
  // autotuned launch shapes per build shape (rows, entries, nnz, size type,
  // d1, d2) and the tuning the last index build ran with
  std::map<std::array<uint64_t, 6>, dv_tuning> tuned;
  dv_tuning last_tuning{};
  uint64_t autotune_builds = 0;

  // per-kernel timing (dv_profile)
  bool prof = false;
  dv_kernel_time prof_t[DV_KERNEL_KINDS] = {};
  hipEvent_t pev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  void prof_mark(int i) {
    if (prof) (void)hipEventRecord(pev[i], stream);
  }
  void prof_add(int kind, int e0, int e1, uint64_t alg, uint64_t slice, bool launch = true) {
    if (!prof) return;
    float ms = 0;
    (void)hipEventSynchronize(pev[e1]);
    (void)hipEventElapsedTime(&ms, pev[e0], pev[e1]);
    prof_t[kind].launches += launch ? 1 : 0;
    prof_t[kind].total_ms += ms;
    prof_t[kind].alg_bytes += alg;
    prof_t[kind].slice_bytes += slice;
  }

  // per-operation statistics
  uint64_t stats_alg_bytes = 0, stats_overflow_rows = 0, stats_mid_rows = 0, stats_big_rows = 0;
  uint32_t stats_n_spgemm = 0, stats_n_spadd = 0;
  double stats_max_ratio = 0;
  uint32_t stats_planned = 0, stats_retried = 0, stats_layout = 1;
  void reset_stats() {
    stats_alg_bytes = stats_overflow_rows = stats_mid_rows = stats_big_rows = 0;
    stats_n_spgemm = stats_n_spadd = 0;
    stats_planned = stats_retried = 0;
    stats_layout = 1;
  }

  bool hip(hipError_t e, const char *what);
  int fail(const std::string &msg, int code = DV_ERR_INVALID);
  void *dalloc(uint64_t bytes);
  void dfree(void *p);
  void forget(void *p);
  void release(DevMat &m);
  bool read_counters(OpCounters &out, uint32_t *status = nullptr);

  int validate(const DevMat &m);
  int clone(const DevMat &src, DevMat &dst);

  // dv_ops.hip
  // staged = false: no stage (slot-form builds keep no staged rows)
  bool op_scratch(uint64_t n, OpScratch &s, bool staged = true);
  GroupLaunch group_launch(uint32_t lanes, uint64_t nstage) const;
  void enq_spadd_count(const DevMat &a, const DevMat &b, void *c_rowmap, uint32_t c_sb,
                       const OpScratch &s, OpCounters *ctr, uint64_t cap, unsigned int *status);
  void enq_spadd_write(const DevMat &a, const DevMat &b, const void *c_rowmap,
                       uint32_t *c_entries, uint64_t cap, OpCounters *ctr,
                       const unsigned int *status);
  // before_lane / after_lane (optional) are recorded right before and right
  // after the lane-tier kernel (so they bracket it alone); with c_entries the
  // scan also copies the staged rows into C; summ_rows > 0: interval
  // summaries of b in blocks of that many rows
  void enq_spgemm_count(const DevMat &a, const DevMat &b, void *c_rowmap, uint32_t c_sb,
                        const OpScratch &s, OpCounters *ctr, uint64_t cap, unsigned int *status,
                        hipEvent_t after_lane = nullptr, uint32_t *c_entries = nullptr,
                        uint32_t summ_rows = 0, bool plus_i = false,
                        hipEvent_t before_lane = nullptr);
  // plus_i: the product is a (b + I)
  void enq_spgemm_write(const DevMat &a, const DevMat &b, const void *c_rowmap,
                        uint32_t *c_entries, const OpScratch &s, OpCounters *ctr,
                        const unsigned int *status, bool staged_copied, bool plus_i = false);
  int identity_dev(uint64_t n, uint32_t sb, DevMat &out);
  int spadd_dev(const DevMat &a, const DevMat &b, DevMat &c);
  int spgemm_dev(const DevMat &a, const DevMat &b, DevMat &c);
  // set-union product (canonical rows of b, or by design)
  int spgemm_core(const DevMat &a, const DevMat &b, DevMat &c);
  // the reference's product for a non-canonical b (its accumulator window)
  int spgemm_window(const DevMat &a, const DevMat &b, DevMat &c);
  // m.canon when known, else checked on the device (1 or 2; 0 on error)
  uint8_t canonical(const DevMat &m);
  int power_exact(const DevMat &a, uint32_t k, DevMat &out);
  int dindex_exact(const DevMat &a, uint32_t dlo, uint32_t dup, DevMat &d);

  // dv_plan.hip
  int buf_acquire(uint64_t rm_bytes, uint64_t cap);
  void buf_trim();
  uint64_t predict_cap(const Plan &p, const PlanSlot &s) const;
  bool plan_ctr_reserve(int nops);
  bool plan_events_reserve(int nops);
  int plan_prepare(Plan &p);
  void plan_launch(Plan &p, bool timing);
  int plan_enqueue(Plan &p);
  int plan_finish(Plan &p, uint32_t &status);
  int plan_run(Plan &p, DevMat &out, const std::function<int(DevMat &)> &exact);
  int power_dev(const DevMat &a, uint32_t k, DevMat &out);
  int dindex_dev(const DevMat &a, uint32_t dlo, uint32_t dup, DevMat &d);
  int dindex_planned(const DevMat &a, uint32_t dlo, uint32_t dup, DevMat &d);
  // dv_slots.hip
  int dindex_slots(const DevMat &a, uint32_t dlo, uint32_t dup, DevMat &d);
  bool slot_parts_reserve(uint64_t words);
  int autotune_dindex(const DevMat &a, uint32_t dlo, uint32_t dup, dv_tuning &best);

  // dv_graph.hip
  int adjacency_dev(const Graph &g, uint64_t lower, uint64_t upper, uint32_t sb, DevMat &out);

  // dv_batch.hip
  int dindex_batched(const std::vector<DevMat> &comps, uint32_t dlo, uint32_t dup,
                     std::vector<DevMat> &outs);
  int dindex_batched_classes(const std::vector<DevMat> &comps, uint32_t dlo, uint32_t dup,
                             std::vector<DevMat> &outs);
  int dindex_batched_one(const std::vector<DevMat> &comps, uint32_t dlo, uint32_t dup,
                         std::vector<DevMat> &outs);
  uint32_t stats_classes = 1;  // builds a batched build ran (one per size class)
};

// Resolve the allocation-free entry points' scratch array on a device
// (dv_noalloc.hip; dv_init calls it).
int noalloc_setup(int device);

}  // namespace dv

// The C ABI's opaque context is the library context.
struct dv_context : public dv::Ctx {};
