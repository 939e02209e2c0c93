/*
 * diverg_amd.h -- C ABI of the MI355X (gfx950) range-CRS Boolean sparse
 * engine: range SpAdd, range SpGEMM, range power, distance-index
 * construction and locus-pair queries, plus a batched build over many graph
 * components and a serving loop.
 *
 * Each entry point names the DiVerG interface it replaces (paths relative to
 * the DiVerG source tree).  Signatures use plain pointers and sizes only.
 *
 * Conventions
 *  - Matrices are described by `dv_rcrs`: rows of closed column intervals
 *    [s, e], stored flat as entries = s0,e0,s1,e1,... with `rowmap[i]` ..
 *    `rowmap[i+1]` delimiting row i (the layout of
 *    CRSMatrix<RangeDynamic,bool,uint32_t,S>, crs_matrix.hpp:1873-2004).
 *    `size_bytes` is sizeof(S): 4 (uint32_t rowmap) or 8 (uint64_t rowmap).
 *  - Inputs may live on the host or on the device (`location`); they are not
 *    modified.  Device inputs are validated on the device before use, except
 *    descriptors returned by this library (`owner` set).  Outputs are allocated by the library at the requested
 *    location and released with dv_free().
 *  - `nnz` is the number of set cells.  With size_bytes == 4 it is reduced
 *    modulo 2^32, which is what the reference's uint32_t accumulator yields
 *    (range_sparse.hpp:1183-1264).
 *  - Errors are returned as dv_status codes; the reference instead calls
 *    exit(134) through DIVERG_ASSERT (utils.hpp:80-88).  dv_last_error()
 *    gives a message for the last failing call on a context.
 *  - Calls are blocking unless named *_async.  A context is used by one host
 *    thread at a time.
 */
#ifndef DIVERG_AMD_H
#define DIVERG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DV_ABI_VERSION 1

typedef enum dv_status {
  DV_OK = 0,
  DV_ERR_INVALID = 1,   /* bad argument: shape mismatch, dlo > dup, dup == 0,
                           odd entry count, unsupported size_bytes */
  DV_ERR_OVERFLOW = 2,  /* ncols > UINT32_MAX - 1 (range_sparse.hpp:2235-2236)
                           or an entry count that does not fit a 32-bit rowmap */
  DV_ERR_NOMEM = 3,     /* host or device allocation failed */
  DV_ERR_DEVICE = 4,    /* HIP runtime error */
  DV_ERR_BUDGET = 5,    /* serving: request exceeds the HBM budget; allocation-free
                           entry points: input beyond their scratch */
  DV_ERR_RANGE = 6,     /* query (i, j) outside the matrix (crs_matrix.hpp:1222-1243) */
  DV_ERR_BUSY = 7,      /* serving: no free slot; retry */
  DV_ERR_NOT_FOUND = 8, /* serving: unknown ticket or index id */
  DV_ERR_IO = 9         /* file cannot be opened, read or written, or is malformed */
} dv_status;

typedef enum dv_location { DV_HOST = 0, DV_DEVICE = 1 } dv_location;

typedef struct dv_rcrs {
  uint64_t nrows;
  uint64_t ncols;
  void *rowmap;        /* nrows + 1 elements of size_bytes each; rowmap[0] == 0 */
  uint32_t *entries;   /* n_entries elements (n_entries == rowmap[nrows], even) */
  uint64_t n_entries;
  uint64_t nnz;
  uint32_t size_bytes; /* 4 or 8 */
  uint32_t location;   /* dv_location */
  void *owner;         /* library-private; NULL for caller-owned storage */
} dv_rcrs;

typedef struct dv_context dv_context;

/* Device work of one build, measured with HIP events on the context stream.
 * alg_bytes sums bytes(X) + bytes(Y) + bytes(C) over every SpGEMM and SpAdd
 * of the build, with bytes(M) = size_bytes * (nrows + 1) + 4 * n_entries. */
typedef struct dv_build_stats {
  double build_ms;       /* first kernel after the input is resident .. last kernel writing D */
  double h2d_ms;         /* host -> device copy of the input (0 if already on the device) */
  double d2h_ms;         /* device -> host copy of the output (0 if kept on the device) */
  uint64_t alg_bytes;
  uint64_t peak_device_bytes;  /* HBM held at the build's high-water mark: the input,
                                  the context's cached buffers and scratch, the output */
  uint32_t n_spgemm;
  uint32_t n_spadd;
  uint64_t overflow_rows;  /* rows recomputed by the direct-write SpGEMM pass */
  uint64_t deferred_rows;  /* rows the lane-per-row tier handed to lane groups */
  uint32_t planned;        /* 1: enqueued as one planned sequence, one readback */
  uint32_t retried;        /* 1: the plan overflowed a predicted capacity and the
                              build was repeated with a readback per operation */
  uint32_t layout;         /* products' intermediate form: 0 slots, 1 range CRS */
  uint32_t builds;         /* planned builds the call ran: 1, or for a batched call
                              with autotune on, one per size class of components */
} dv_build_stats;

/* --- runtime ------------------------------------------------------------
 * Replaces diverg::initialize_runtime / finalize_runtime / ScopedRuntime
 * (runtime.hpp:25-55, src/eti/runtime.cpp:22-49). */
int dv_init(int device, dv_context **ctx);
int dv_finalize(dv_context *ctx);
/* Wait for every operation the context has enqueued (the reference's
 * execution_space{}.fence(), e.g. dindex.hpp:181-187). */
int dv_synchronize(dv_context *ctx);
const char *dv_last_error(const dv_context *ctx);
const char *dv_status_string(int status);
int dv_abi_version(void);
/* The HIP stream (hipStream_t) the context's work runs on. */
void *dv_stream(dv_context *ctx);

/* Launch-shape tuning (the reference's SparseConfig / ExecGrid,
 * range_sparse_base.hpp:99-498).  dv_set_tuning replaces the whole set;
 * read it with dv_get_tuning first.
 *   lds_words       LDS bit-window words of a 64-lane group (a G-lane group
 *                   gets lds_words * G / 64); wider windows are chunked
 *   stage_entries   entries per row kept by the counting pass (even)
 *   rows_per_group  consecutive rows a lane group takes per visit
 *   lanes_small     lanes per row for medium rows: 4, 8, 16, 32 or 64
 *   small_slice_max medium rows: at most this many B-slice ranges
 *   lane_slice_max  lane-per-row tier: at most this many slice ranges
 *                   (0 sends every non-empty row to the groups)
 *   autotune        1: the first planned dv_create_distance_index* build of
 *                   each shape (rows, entries, nnz, size type, d1, d2) times
 *                   candidate launch shapes -- coordinate descent over
 *                   stage_entries, lanes_small, small_slice_max,
 *                   lane_slice_max, summary_span, lds_words, rows_per_group,
 *                   each a whole build, best of two -- and later builds of
 *                   that shape reuse the winner (dv_last_tuning shows it)
 *   plan_mode       0: index builds and powers are enqueued as one sequence
 *                   with predicted output capacities and a single readback
 *                   (exact fallback on overflow); 1: one readback per
 *                   operation, as the reference's drivers; 2: as 0 with a
 *                   capacity prediction too small to hold any product, which
 *                   always takes the fallback (for tests)
 *   graphs          1: the launches of a planned build are captured once into
 *                   a hipGraph and replayed while the build's operations,
 *                   buffers and tuning stay the same (not while profiling)
 *   summary_span    planned SpGEMMs whose left operand's ranges may span at
 *                   least this many rows (>= 2; 0: never) first summarise the
 *                   right operand in blocks of 32 to 256 rows (per row, the
 *                   union of the rows from its block's start / to its block's
 *                   end when it has at most two pieces); a left range of at
 *                   least this many rows that crosses a block boundary is
 *                   then read as one suffix, whole blocks and one prefix
 *                   instead of row by row
 *   schedule        0: a planned index build takes whichever of
 *                   A^d1 (A + I)^(d2 - d1) and (A (A + I))^m . rest
 *                   (m = min(d1, d2 - d1)) needs fewer SpGEMMs -- the same
 *                   index, as the factors commute; 1: always the reference's
 *                   order (dindex.hpp:131-177)
 *   layout          2: the products of a planned index build keep their
 *                   results in slot form (one range per row in place, rows of
 *                   several ranges in a side array), so no product needs a
 *                   row-pointer scan; only the index is converted to range
 *                   CRS; 1: every product writes range CRS; 0: slot form,
 *                   except that a build shape whose products had side
 *                   entries for more than 75% of their rows is built in
 *                   range CRS from then on (a row of several ranges costs the
 *                   slot form a dependent read)
 *   chain_span      slot-form products whose left operand's ranges may span
 *                   at least this many rows (0: never) read the right operand
 *                   through its chains, which every slot-form product writes:
 *                   per 64-row chunk, where each row's run of consecutive
 *                   single-range rows with non-decreasing, touching ends
 *                   stops, and the chunk's union when all 64 rows are one
 *                   such run; a slice costs a few loads per 64 rows instead of
 *                   one per row (no interval summaries are computed) */
typedef struct dv_tuning {
  uint32_t lds_words;
  uint32_t stage_entries;
  uint32_t rows_per_group;
  uint32_t lanes_small;
  uint32_t small_slice_max;
  uint32_t lane_slice_max;
  uint32_t autotune;
  uint32_t plan_mode;
  uint32_t graphs;
  uint32_t summary_span;
  uint32_t schedule;
  uint32_t layout;
  uint32_t chain_span;
} dv_tuning;
int dv_get_tuning(const dv_context *ctx, dv_tuning *t);
int dv_set_tuning(dv_context *ctx, const dv_tuning *t);
/* The launch shape the last dv_create_distance_index* call ran with (the
 * autotuned one when autotune is on).  No counterpart in the reference. */
int dv_last_tuning(const dv_context *ctx, dv_tuning *t);

/* Per-kernel timing with HIP events on the context stream (no counterpart in
 * the reference, whose DIVERG_STATS timers print symbolic/numeric times after
 * a fence, range_sparse.hpp:2243-2278).  Kinds: 0 the SpGEMM lane-per-row
 * counting kernel (k_spgemm_lane, bracketed alone), 1 SpGEMM write passes
 * (staged copy + direct rows), 2 the rest of a SpGEMM's counting phase
 * (interval summaries, group-tier kernels, row-pointer scan), 3 SpAdd.  alg_bytes sums bytes(X) + bytes(Y) + bytes(C) of the timed
 * SpGEMMs; slice_bytes sums 8 bytes per B-slice range the lane and group
 * tiers gathered. */
#define DV_KERNEL_KINDS 4
typedef struct dv_kernel_time {
  uint64_t launches;
  double total_ms;
  uint64_t alg_bytes;
  uint64_t slice_bytes;
} dv_kernel_time;
int dv_profile(dv_context *ctx, int enable); /* enabling also clears the totals */
int dv_profile_read(const dv_context *ctx, dv_kernel_time out[DV_KERNEL_KINDS]);

/* --- matrices ------------------------------------------------------------ */
/* Copy `src` to `location` (H2D, D2H, D2D or H2H); replaces
 * entries_device_view / rowmap_device_view (crs_matrix_device.hpp:192-238)
 * and make_crs_matrix (crs_matrix_device.hpp:265-290). */
int dv_copy(dv_context *ctx, const dv_rcrs *src, uint32_t location, dv_rcrs *dst);
/* Release a matrix allocated by the library. */
int dv_free(dv_context *ctx, dv_rcrs *m);

/* --- hot path ------------------------------------------------------------ */
/* create_range_identity_matrix (range_sparse.hpp:66-110). */
int dv_range_identity(dv_context *ctx, uint64_t n, uint32_t size_bytes,
                      uint32_t out_location, dv_rcrs *out);

/* C = A v B.  RangeSpAdd<M,Space>::spadd (range_sparse.hpp:430-477) and the
 * view-level range_spadd (range_sparse.hpp:364-428). */
int dv_range_spadd(dv_context *ctx, const dv_rcrs *a, const dv_rcrs *b,
                   uint32_t out_location, dv_rcrs *c);

/* C = A . B (Boolean).  RangeSpGEMM<M,Config>::spgemm
 * (range_sparse.hpp:2283-2332) and the view-level range_spgemm
 * (range_sparse.hpp:2215-2281). */
int dv_range_spgemm(dv_context *ctx, const dv_rcrs *a, const dv_rcrs *b,
                    uint32_t out_location, dv_rcrs *c);

/* C = A^k.  RangePower<M,Config>::power (range_sparse.hpp:2334-2368) and
 * range_power_inplace (range_sparse.hpp:2370-2431). */
int dv_range_power(dv_context *ctx, const dv_rcrs *a, uint32_t k,
                   uint32_t out_location, dv_rcrs *c);

/* D = A^dlo . (A + I)^(dup - dlo), or (A + I)^dup when dlo == 0.
 * util::create_distance_index (dindex.hpp:46-189; compiled boundary
 * dindex_eti.hpp:44-53, src/eti/dindex.cpp:25-39).  `stats` may be NULL. */
int dv_create_distance_index(dv_context *ctx, const dv_rcrs *a, uint32_t dlo,
                             uint32_t dup, uint32_t out_location, dv_rcrs *d,
                             dv_build_stats *stats);

/* Distance indexes of `ncomp` independent graph components in one planned
 * build over their block-diagonal assembly: each power step is one launch
 * sequence covering every component (replayed from a hipGraph with
 * dv_tuning.graphs).  outs[c] is bit-identical to dv_create_distance_index
 * on comps[c] (the block assembly of test_dindex.cpp:111-173).  The
 * components together may hold at most UINT32_MAX - 1 loci.  No counterpart
 * entry point in the reference, which builds one graph per call. */
int dv_create_distance_index_batched(dv_context *ctx, const dv_rcrs *comps, uint32_t ncomp,
                                     uint32_t dlo, uint32_t dup, uint32_t out_location,
                                     dv_rcrs *outs, dv_build_stats *stats);

/* out[q] = D(qi[q], qj[q]) for q < n.  CRSMatrixBase::operator()(i, j)
 * (crs_matrix.hpp:1197-1218) via binary_search(..., RangeGroup)
 * (crs_matrix.hpp:149-169).  qi, qj, out are host or device pointers
 * (`query_location`); the matrix may be host or device resident.
 * Out-of-range pairs give out[q] = 2 and DV_ERR_RANGE (at(), crs_matrix.hpp:
 * 1222-1243, throws). */
int dv_query_batch(dv_context *ctx, const dv_rcrs *m, const uint32_t *qi,
                   const uint32_t *qj, uint8_t *out, uint64_t n,
                   uint32_t query_location);

/* Merge two distance indices of the same shape: the union of their cells.
 * util::merge_distance_index<TMutable>(d1, d2) for the Range group
 * (dindex.hpp:556-579): for in-core RangeDynamic operands the reference takes
 * range_spadd (dindex.hpp:435-456), and so does this entry point (the same
 * kernels as dv_range_spadd).  Its streaming variant for buffered or
 * compressed operands (dindex.hpp:290-420) gives the same arrays on canonical
 * indices.  nnz is recounted from the merged entries, as crs_matrix::nnz
 * (crs_matrix.hpp:175-189). */
int dv_merge_distance_index(dv_context *ctx, const dv_rcrs *d1, const dv_rcrs *d2,
                            uint32_t out_location, dv_rcrs *out);

/* --- persistence ------------------------------------------------------------
 * The byte layout of CRSMatrix<RangeDynamic,bool,uint32_t,S>::serialize /
 * load (crs_matrix.hpp:1392-1409, 1986-2000; vectors as utils.hpp:191-289):
 * u64 count + uint32 entries, u64 count + S rowmap, u64 ncols, u64 nnz.
 * Files are interchangeable with the reference's.  dv_save accepts host or
 * device matrices (ctx may be NULL for a host matrix).  dv_load validates the
 * structure (rowmap from 0, even, non-decreasing, ending at the entry count;
 * entries below ncols) before returning it; size_bytes 0 takes the size type
 * from the file length. */
int dv_save(dv_context *ctx, const dv_rcrs *m, const char *path);
int dv_load(dv_context *ctx, const char *path, uint32_t size_bytes, uint32_t location,
            dv_rcrs *out);

/* --- allocation-free SpAdd -------------------------------------------------
 * No counterpart in the reference: a variant of dv_range_spadd for callers
 * that cannot allocate device memory (code under stream capture, serving
 * loops, harnesses that forbid it).  C = A + B as dv_range_spadd computes it
 * (range_sparse.hpp:112-428), with A and B device resident (same shape and
 * size type) and C written to caller memory: c_rowmap (nrows + 1 words of
 * the size type) and c_entries (c_capacity words; len(A) + len(B) always
 * suffices).  *c_n_entries receives C's entry count (also when it exceeds
 * c_capacity: DV_ERR_OVERFLOW, C not written), *c_nnz its nnz (wrapped to 32
 * bits for a 32-bit rowmap).  stream: a hipStream_t, or NULL for the
 * context's; the call returns when C is complete.  Its scratch is one
 * library-owned device array of DV_NOALLOC_SCRATCH_BYTES per GPU (held until
 * the process ends): at most dv_range_spadd_noalloc_max_rows() rows, more is
 * DV_ERR_BUDGET.  Calls that overlap on one GPU share that array and race. */
#define DV_NOALLOC_SCRATCH_BYTES (32ull << 20)
uint64_t dv_range_spadd_noalloc_max_rows(void);
int dv_range_spadd_noalloc(dv_context *ctx, const dv_rcrs *a, const dv_rcrs *b, void *c_rowmap,
                           uint32_t *c_entries, uint64_t c_capacity, uint64_t *c_n_entries,
                           uint64_t *c_nnz, void *stream);

/* --- serving --------------------------------------------------------------
 * A serving loop over HBM-resident indexes (no counterpart in the reference,
 * whose queries are host binary searches, crs_matrix.hpp:149-169).
 * Continuous batching: queries that arrive while earlier batches are on the
 * GPU are packed into the next launch (up to max_inflight batches in flight),
 * on a query stream of its own.
 * Index builds run on a second stream; a build is admitted when its predicted
 * HBM footprint fits the budget, and its index joins the query table when it
 * completes.  Answers equal dv_query_batch on the same index. */
typedef struct dv_server dv_server;

typedef struct dv_server_opts {
  uint64_t hbm_budget_bytes; /* indexes + running builds; 0: 75% of free HBM */
  uint32_t max_batch;        /* queries per launch; 0: 65536 */
  uint32_t max_inflight;     /* batches in flight; 0: 2 */
  uint32_t query_cus;        /* reserved, must be 0 (CU-masked query streams are
                                switched off; nonzero is DV_ERR_INVALID) */
  uint32_t pad;
} dv_server_opts;

typedef struct dv_request_info {
  uint32_t kind;      /* 0 queries, 1 build */
  int32_t status;     /* dv_status of the finished request; DV_ERR_BUSY while pending */
  uint32_t index_id;  /* queries: the index asked; build: the index produced */
  uint32_t pad;
  uint64_t n;         /* queries in the request */
  double submit_us;   /* server clock, microseconds */
  double start_us;    /* first launch (queries) / admission (build) */
  double done_us;
} dv_request_info;

int dv_server_start(int device, const dv_server_opts *opts, dv_server **srv);
int dv_server_stop(dv_server *srv);
/* Copy an index to HBM and make it queryable; *index_id names it. */
int dv_server_add_index(dv_server *srv, const dv_rcrs *index, uint32_t *index_id);
/* Queue n locus pairs against one index; returns at once. */
int dv_server_submit_queries(dv_server *srv, uint32_t index_id, const uint32_t *qi,
                             const uint32_t *qj, uint64_t n, uint64_t *ticket);
/* Queue a distance-index build of `a` (host or device resident; copied before
 * the call returns); DV_ERR_BUDGET if it can never fit the budget. */
int dv_server_submit_build(dv_server *srv, const dv_rcrs *a, uint32_t dlo, uint32_t dup,
                           uint64_t *ticket);
/* DV_OK and *info when the request is finished, DV_ERR_BUSY while pending. */
int dv_server_poll(dv_server *srv, uint64_t ticket, dv_request_info *info);
int dv_server_wait(dv_server *srv, uint64_t ticket, dv_request_info *info);
/* Answers of a finished query request (n bytes: 1 hit, 0 miss, 2 out of range). */
int dv_server_results(dv_server *srv, uint64_t ticket, uint8_t *out, uint64_t n);
/* Copy an index the server holds (e.g. a finished build) to the host. */
int dv_server_get_index(dv_server *srv, uint32_t index_id, dv_rcrs *out);

/* Load generator: replays a query stream against the server in real time.
 * Query q (index qidx[q], pair (qi[q], qj[q])) arrives at arrival_us[q]
 * (ascending, microseconds from the start); build b of builds[b] arrives at
 * build_us[b].  Queries are submitted as they fall due, in arrival order;
 * the call returns when every query and build has finished.  answers[q] gets
 * the answer, latency_us[q] the time from arrival to the answer being on the
 * host; build_done_us[b] the completion time of build b and build_index[b]
 * the index it produced.  *elapsed_us is the replay's wall time.  The build
 * inputs are read in place (no copy on the submitting thread), so they must
 * stay valid until the call returns. */
int dv_server_replay(dv_server *srv, uint64_t n, const uint32_t *qidx, const uint32_t *qi,
                     const uint32_t *qj, const double *arrival_us, uint8_t *answers,
                     double *latency_us, uint32_t nbuilds, const dv_rcrs *builds,
                     const uint32_t *build_dlo, const uint32_t *build_dup,
                     const double *build_us, double *build_done_us, uint32_t *build_index,
                     double *elapsed_us);

#ifdef __cplusplus
}
#endif

#endif /* DIVERG_AMD_H */
