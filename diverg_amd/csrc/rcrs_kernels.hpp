// Launchers of the gfx950 range-CRS kernels (rcrs_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace dv {

// Device view of an rCRS operand; `rowmap` holds S = uint32_t or uint64_t.
struct DView {
  const void *rowmap;
  const uint32_t *entries;
  uint64_t nrows;
  uint64_t ncols;
};

// Device counters of one SpGEMM / SpAdd, read back once per operation.
struct OpCounters {
  unsigned long long nnz;
  unsigned long long total;     // entries of C (rowmap[n])
  unsigned long long slice;     // B-slice ranges gathered by the counting pass
  unsigned int overflow;        // rows whose output exceeds the stage
  unsigned int n_mid;           // rows handed to the 16-lane groups
  unsigned int n_big;           // rows handed to the 64-lane groups
  unsigned int pad;
  unsigned long long ranges;    // ranges of C (slot-format products)
};

// One launch of the group tiers over a row list.  Counting mode when
// c_entries is NULL, direct-write mode otherwise.  The list length is *n_dev
// (read on the device) when n_dev is set, else n_host.
struct GroupLaunch {
  uint32_t lanes;           // 4, 8, 16, 32 or 64
  uint32_t lds_words;       // bit-window words per group
  uint32_t stage;           // entries per row kept by the counting pass
  uint32_t rows_per_group;  // consecutive list entries per group visit
  uint32_t grid;            // workgroups (multiple of 8)
  uint64_t nstage;          // rows of the staging buffer (its column stride)
  const uint32_t *rows;
  const unsigned int *n_dev;
  uint32_t n_host;
  uint32_t *cnt;
  uint32_t *stage_buf;
  uint32_t *ovf_list;
  OpCounters *ctr;
  const void *c_rowmap;
  uint32_t *c_entries;
  const unsigned int *status;  // kernels exit at once when *status != 0
  uint32_t plus_i;             // 1: the product is A (B + I)
};

void fill_u32(hipStream_t s, uint32_t *p, uint64_t n, uint32_t v);
// p[0, n) and q[0, m) set to v in one launch
void fill2_u32(hipStream_t s, uint32_t *p, uint64_t n, uint32_t *q, uint64_t m, uint32_t v);
void zero_counters(hipStream_t s, OpCounters *c);

void identity(hipStream_t s, uint32_t sb, uint64_t n, void *rowmap, uint32_t *entries);

// Row-pointer scan: rowmap[0] = 0, rowmap[i+1] = cnt[0] + ... + cnt[i];
// total -> ctr->total; a total above `cap` sets bit 0 of *status, above a
// 32-bit rowmap bit 1.  Two launches (tile sums, then rows) and, when
// c_entries is given, the copy of staged rows into C (stage of
// stage_entries per row).  tile_sums needs scan_words(n) words.  `parts`
// holds nparts (nnz, slice) pairs of the lane-tier kernel's blocks, added to
// ctr->nnz and ctr->slice.
uint64_t scan_words(uint64_t n);
void scan_rowmap(hipStream_t s, uint32_t sb, const uint32_t *cnt, uint64_t n, void *rowmap,
                 unsigned long long *tile_sums, OpCounters *ctr, uint64_t cap,
                 unsigned int *status, const uint32_t *stage, uint32_t stage_entries,
                 uint32_t *c_entries, const unsigned long long *parts = nullptr,
                 uint32_t nparts = 0);

// Kernels that read operands take `status` and do nothing once it is set.
void spadd_count(hipStream_t s, uint32_t sb, DView a, DView b, uint32_t *cnt,
                 const unsigned int *status);
void spadd_write(hipStream_t s, uint32_t sb, DView a, DView b,
                 const void *c_rowmap, uint32_t *c_entries, const unsigned int *status);
// nnz of the ctr->total entries (at most max_entries) added to ctr->nnz.
void range_nnz(hipStream_t s, const uint32_t *entries, uint64_t max_entries,
               OpCounters *ctr, const unsigned int *status);

// Lane tier of the counting pass: one lane per row of A.  Rows it cannot
// finish (more than lane_max slice ranges, or a union of more than 4 runs)
// are appended to mid_list (slices of at most mid_max ranges) or big_list.
// Writes per-row entry counts, the first `stage` entries of each row into
// stage_buf (range k of row r at uint2 k * nrows + r), rows with more output to ovf_list,
// and per block the set cells and gathered slice ranges to parts[2 b], parts[2 b + 1]
// (grid pairs; scan_rowmap adds them to the counters).
void spgemm_lane(hipStream_t s, uint32_t sb, DView a, DView b, uint32_t stage, uint32_t lane_max,
                 uint32_t mid_max,
                 uint32_t grid, uint32_t *cnt, uint32_t *stage_buf, uint32_t *ovf_list,
                 uint32_t *mid_list, uint32_t *big_list, OpCounters *ctr,
                 const unsigned int *status, const void *summ, uint32_t summ_min,
                 uint32_t summ_rows, uint32_t plus_i, unsigned long long *parts);
// Interval summaries of B in blocks of block_rows rows (a power of two, 32 to
// 256): per row, the union of the rows from its block's start to it (prefix)
// and from it to its block's end (suffix), as its one or two pieces (else
// "unknown"); summary_words(nrows) uint2 words (8-byte prefixes, suffixes,
// then 16-byte two-piece side entries of each).  The lane tier (summ, summ_rows = block_rows) folds an A range of
// at least summ_min rows that crosses a block boundary as a suffix, whole
// blocks and a prefix: 2 + span / block_rows reads instead of the rows' ranges.
uint64_t summary_words(uint64_t nrows);
// Blocks of the lane-tier kernel resident per CU (its occupancy), for its
// persistent grid.
uint32_t lane_blocks_per_cu(uint32_t sb, bool summaries);
void block_summary(hipStream_t s, uint32_t sb, DView b, void *summ, const unsigned int *status,
                   uint32_t block_rows);
// Summary block rows for products whose left operand's ranges cover up to
// `span` rows of the right one: the power of two between an eighth and a
// quarter of the span (so a range reads 4 to 8 block summaries), 32 to 256.
uint32_t summary_block_rows(uint64_t span);
// Group tiers (counting or direct write), see GroupLaunch.
void spgemm_group(hipStream_t s, uint32_t sb, DView a, DView b, const GroupLaunch &g);
// Copy staged rows (count <= stage) to their place in C.
void spgemm_emit_staged(hipStream_t s, uint32_t sb, uint64_t nrows,
                        const void *c_rowmap, const uint32_t *stage_buf,
                        uint32_t stage, uint32_t *c_entries, const unsigned int *status);

// out[q] = M(qi[q], qj[q]); 2 and *bad = 1 for out-of-range pairs.
void query(hipStream_t s, uint32_t sb, DView m, const uint32_t *qi,
           const uint32_t *qj, uint8_t *out, uint64_t n, uint32_t *bad);

// Component `c` of a block-diagonal assembly (uint64 rowmap): its rows go to
// [row0, row0 + n) with rowmap offsets + ent0 and columns + col0;
// write_last also writes rowmap[row0 + n].
void concat_component(hipStream_t s, uint32_t sb, const void *rm, const uint32_t *en, uint64_t n,
                      uint64_t n_entries, uint64_t row0, uint64_t ent0, uint32_t col0,
                      uint64_t *out_rm, uint32_t *out_en, bool write_last);
// The inverse: rows [row0, row0 + n) of the assembly, columns - col0, into a
// component matrix with a size_bytes = sb rowmap.
void extract_component(hipStream_t s, uint32_t sb, const uint64_t *rm, const uint32_t *en,
                       uint64_t row0, uint64_t n, uint64_t n_entries, uint32_t col0,
                       void *out_rm, uint32_t *out_en);
// out[c] = rm[bounds[c]] for c < m.
void gather_bounds(hipStream_t s, const uint64_t *rm, const uint64_t *bounds, uint32_t m,
                   unsigned long long *out);
// *out += sum of (e - s + 1) over entries[0 .. n_entries).
void slice_nnz(hipStream_t s, const uint32_t *entries, uint64_t n_entries,
               unsigned long long *out);

// A resident index as the serving query kernel sees it.
struct IndexDesc {
  const void *rowmap;
  const uint32_t *entries;
  uint64_t nrows;
  uint64_t ncols;
  uint32_t sb;
  uint32_t pad;
};
// out[t] = D_q(i, j) for query t = (qidx[t], qi[t], qj[t]) (index id, row,
// column); 2 for an unknown index or (i, j) outside it.
void query_multi(hipStream_t s, const IndexDesc *table, uint32_t ntable, const uint32_t *qidx,
                 const uint32_t *qi, const uint32_t *qj, uint32_t n, uint8_t *out);

// Sets *flag when rowmap[0] != 0, rowmap[n] != n_entries, a row length is
// odd or negative, or a range is not s <= e < ncols.
void validate_matrix(hipStream_t s, uint32_t sb, DView m, uint64_t n_entries, uint32_t *flag);
// Sets bit 2 (4) of *status when a row of m is not canonical (its ranges are
// not ascending with a gap of at least one column between neighbours).
void check_canonical(hipStream_t s, uint32_t sb, DView m, unsigned int *status);
// For C = A.B with a non-canonical B: per row of A, B's ranges clipped to the
// reference accumulator's window as it grows (range_sparse.hpp:1192-1242,
// hbitvector.hpp:278-299).  Without p_entries: the entry count of each row
// into cnt; with it: the pieces at the rows of p_rowmap.  The union of a
// row's pieces is the reference's row of C.
void window_pieces(hipStream_t s, uint32_t sb, DView a, DView b, uint32_t *cnt,
                   const void *p_rowmap, uint32_t *p_entries, const unsigned int *status);

// Number of B ranges each row of A gathers (sum over the row's ranges of the
// B slice length) -- the work measure behind the size classes.
void slice_lengths(hipStream_t s, uint32_t sb, DView a, DView b, uint32_t *len);

// --- Slot format (intermediate matrices of planned index builds) -----------
// One uint2 per row: (lo, hi) for a row of one range, (0xFFFFFFFF, 0) for an
// empty row, (0xFFFFFFFE, off) for a row of several ranges in the side array
// (side[off] = (count, 0), then the ranges).  Side arrays hold side_cap
// entries plus SLOT_SIDE_PAD spare ones (side_cap below 2^32 - 16: offsets
// are 32-bit), reserved in SLOT_SIDE_PARTS partitions of side_cap /
// SLOT_SIDE_PARTS entries, each counted in its own word of side_ctr (a
// product that would overfill a partition sets bit 0 of *status, and every
// later kernel of the build then does nothing); ctr->ranges counts the ranges
// written.
constexpr uint64_t SLOT_SIDE_PAD = 8;
constexpr uint32_t SLOT_SIDE_PARTS = 64;
struct SlotView {
  const void *slot;
  const void *side;
};
// Chains of a slot matrix written by a product (seg: one byte per row, the
// chunk offset of the last row of the chain starting there; chunk: per 64
// rows, their union when they form one chain, else (1, 0)); null when the
// matrix has none.
struct ChainView {
  const void *seg;
  const void *chunk;
};
// Bytes of a product's slot buffer: slots, then chain offsets, then chunks.
inline uint64_t slot_chain_bytes(uint64_t n) {
  return 8 * n + ((n + 7) & ~7ull) + 8 * ((n + 63) / 64) + 64;
}
inline ChainView slot_chains(void *slot, uint64_t n) {
  uint8_t *seg = static_cast<uint8_t *>(slot) + 8 * n;
  return ChainView{seg, seg + ((n + 7) & ~7ull)};
}
void slots_from_rcrs(hipStream_t s, uint32_t sb, DView a, void *slot, void *side,
                     uint64_t side_cap, unsigned long long *side_ctr, unsigned int *status);
void slots_identity(hipStream_t s, uint64_t n, void *slot);
// The chains of a slot matrix (after every kernel of its product ran).
void slots_chains(hipStream_t s, const void *slot, uint64_t n, ChainView c,
                  const unsigned int *status);
// C = I + A; parts as slots_lane (at most max_blocks blocks: the launch
// has slots_plus_identity_blocks(n, max_blocks))
uint32_t slots_plus_identity_blocks(uint64_t n, uint32_t max_blocks);
void slots_plus_identity(hipStream_t s, SlotView a, uint64_t n, void *c_slot, void *c_side,
                         uint64_t side_cap, unsigned long long *side_ctr, unsigned int *status,
                         unsigned long long *parts, uint32_t max_blocks);
// Interval summaries of B (layout and block_rows as block_summary).
void slots_summary(hipStream_t s, SlotView b, uint64_t nrows, void *summ,
                   const unsigned int *status, uint32_t block_rows);
// Lane tier of C = A.B (A (B + I) with plus_i): rows it finishes are
// written; others go to mid_list / big_list (counts in ctr->n_mid, n_big)
// for slots_group.  parts receives 3 words per block (nnz, slice, ranges).
// batch (without summaries): 2, 4 or 8, about the widest A range; 2 and 4
// select the narrow form, which loads that many B slots per row.  With
// b_chains.seg set, B slices are read through B's chains (no summaries);
// with c_chains.seg set, the kernel writes C's chains (the group tiers' rows
// end chains).  In the chain and narrow forms the rows the branch-free join
// cannot take are folded slot by slot within the same launch.  Returns
// whether C's chains were written (false when c_chains is null, and in the
// chain and narrow forms, whose deferred rows are written after their
// chunks: then slots_chains makes them).  With ev_start / ev_stop the lane
// kernel's launch records them as it starts and ends.
bool slots_lane(hipStream_t s, SlotView a, SlotView b, uint64_t nrows, uint64_t b_nrows,
                uint32_t lane_max, uint32_t mid_max, uint32_t grid, void *c_slot, void *c_side,
                uint64_t side_cap, unsigned long long *side_ctr, uint32_t *mid_list,
                uint32_t *big_list, OpCounters *ctr, unsigned int *status, const void *summ,
                uint32_t summ_min, uint32_t summ_rows, uint32_t plus_i,
                unsigned long long *parts, uint32_t batch, ChainView b_chains,
                ChainView c_chains, hipEvent_t ev_start = nullptr,
                hipEvent_t ev_stop = nullptr);
// Both group tiers of a product in one launch: rows1 (*n1 of them) by groups
// of lanes1 lanes (LDS window lds1 words) on grid1 blocks, rows2 by 64-lane
// groups (lds2) on grid2 blocks; grids multiples of 8.
void slots_groups(hipStream_t s, SlotView a, SlotView b, uint32_t lanes1, uint32_t lds1,
                  uint32_t grid1, const uint32_t *rows1, const unsigned int *n1, uint32_t lds2,
                  uint32_t grid2, const uint32_t *rows2, const unsigned int *n2,
                  uint32_t rows_per_group, void *c_slot, void *c_side, uint64_t side_cap,
                  unsigned long long *side_ctr, OpCounters *ctr, unsigned int *status,
                  uint32_t plus_i);
// Blocks that wrote parts, per operation (a kernel argument): n[i] for
// operation i, or per_op for every operation when `all` is set (then the
// unused parts must have been cleared).
constexpr uint32_t PART_BLOCKS_MAX_OPS = 64;
struct PartBlocks {
  uint32_t all;
  uint32_t n[PART_BLOCKS_MAX_OPS];
};
// ctr[i] += the sums of parts[3 per_op i ..] (the first used.n[i] blocks'
// words); ctr[i].total = the side entries used (side_ctr[SLOT_SIDE_PARTS i
// ..]), ctr[i].pad those of the fullest partition (one block per operation).
void slots_reduce_parts(hipStream_t s, const unsigned long long *parts, uint32_t per_op,
                        const PartBlocks &used, const unsigned long long *side_ctr,
                        uint32_t nops, OpCounters *ctr);
// Slots -> rCRS: the row pointers (2 entries per range; cnt receives the
// per-row counts, tile_sums the scan's tile sums: scan_rowmap's buffers,
// no capacity), then the entries at their row pointers.
void slots_scan_rowmap(hipStream_t s, uint32_t sb, SlotView a, uint64_t n, uint32_t *cnt,
                       void *rowmap, unsigned long long *tile_sums, OpCounters *ctr,
                       unsigned int *status);
void slots_emit(hipStream_t s, SlotView a, uint64_t n, uint32_t sb, const void *rowmap,
                uint32_t *entries);
// Resident blocks per CU of the lane kernel that slots_lane runs for these
// arguments.
uint32_t slot_lane_blocks_per_cu(bool summaries, uint32_t batch, bool chains);

}  // namespace dv
