/*
 * rcrs_oracle.c -- CPU restatement of DiVerG's range-CRS (rCRS) Boolean
 * sparse path.  TEST INFRASTRUCTURE ONLY: this file is the parity checker and
 * the timed CPU baseline.  Nothing in the product library links or calls it.
 *
 * Every function restates the reference's OpenMP path and names the
 * reference file:line it follows (paths relative to the DiVerG source tree,
 * include/diverg/...).  It is plain C11 + OpenMP so that it builds with gcc
 * alone; the reference itself needs Kokkos, which is not available.
 *
 * Pinning (what checks each function, besides restating the code):
 *   orc_identity      closed form (entries[2i]=entries[2i+1]=i).
 *   orc_spadd         union of sets (canonical inputs) + literal restatement
 *                     for the verbatim-tail behaviour on non-canonical rows.
 *   orc_spgemm        dense Boolean product on small inputs, the HBitVector
 *                     known-answer test (test_hbitvector.cpp:569-661) as
 *                     I * M, and the manuscript's Example 26.
 *   orc_power         repeated dense products; k = 0, 1, 2.
 *   orc_dindex        layered walk-length BFS on small graphs and closed
 *                     forms on chains (single range [i+d1, min(i+d2, L-1)]).
 *   orc_query         the 10x10 matrix of test_crsmatrix.cpp:398-422 and
 *                     brute-force membership.
 *   orc_walk_rows     independent forward BFS by exact walk length; used to
 *                     spot-check rows of full-size indexes.
 *   orc_adjacency     dense adjacency built from its definition on small
 *                     graphs, nnz = loci - nodes + edges
 *                     (adjacency.hpp:192-194), and the block-diagonal
 *                     assembly of components (test_dindex.cpp:84-173).
 *   orc_merge_stream  union of sets on canonical indices (the merge tests of
 *                     test_crsmatrix.cpp:911-1165) and equality with
 *                     orc_spadd there.
 * The checks live in tests/test_oracle.py.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ORC_OK 0
#define ORC_EINVAL 1
#define ORC_ENOMEM 3

/* A range-CRS matrix: rows of closed column intervals [s, e], stored flat
 * as entries = s0,e0,s1,e1,...; rowmap[i]..rowmap[i+1] delimits row i.
 * Mirrors CRSMatrix<RangeDynamic,bool,uint32_t,S> (crs_matrix.hpp:1873-2004).
 * nnz is kept exact in 64 bits; a caller with a 32-bit size type narrows it,
 * which equals the reference's modular uint32 accumulation. */
typedef struct {
  uint64_t nrows;
  uint64_t ncols;
  uint64_t *rowmap;
  uint32_t *entries;
  uint64_t nnz;
} orc_rcrs;

void orc_free(orc_rcrs *m) {
  if (!m) return;
  free(m->rowmap);
  free(m->entries);
  m->rowmap = NULL;
  m->entries = NULL;
  m->nrows = m->ncols = m->nnz = 0;
}

static int alloc_rcrs(orc_rcrs *m, uint64_t nrows, uint64_t ncols,
                      uint64_t nent) {
  m->nrows = nrows;
  m->ncols = ncols;
  m->nnz = 0;
  m->rowmap = (uint64_t *)malloc(sizeof(uint64_t) * (nrows + 1));
  m->entries = (uint32_t *)malloc(sizeof(uint32_t) * (nent ? nent : 1));
  if (!m->rowmap || !m->entries) {
    orc_free(m);
    return ORC_ENOMEM;
  }
  m->rowmap[0] = 0;
  return ORC_OK;
}

/* crs_matrix.hpp:175-189 -- sum of (e - s + 1) over all ranges. */
uint64_t orc_nnz(const orc_rcrs *m) {
  uint64_t n = 0;
  uint64_t len = m->rowmap[m->nrows];
  for (uint64_t idx = 0; idx < len; idx += 2)
    n += (uint64_t)m->entries[idx + 1] - m->entries[idx] + 1;
  return n;
}

/* range_sparse.hpp:66-88 (create_range_identity_matrix). */
int orc_identity(uint64_t n, orc_rcrs *out) {
  int rc = alloc_rcrs(out, n, n, 2 * n);
  if (rc) return rc;
  for (uint64_t i = 0; i < n; ++i) {
    out->entries[2 * i] = (uint32_t)i;
    out->entries[2 * i + 1] = (uint32_t)i;
    out->rowmap[i + 1] = 2 * (i + 1);
  }
  out->nnz = n; /* range_sparse.hpp:108 passes n as nnz */
  return ORC_OK;
}

/* ---------------------------------------------------------------------------
 * Range SpAdd: range_sparse.hpp:112-357.  One row at a time, two-pointer
 * merge over the two sorted range lists; a range is merged into the pending
 * one when rs <= hi + 1.  Once one side runs out, ranges of the other side
 * are merged until the first one that does not touch the pending range; that
 * one and everything after it are copied verbatim (symbolic :199-223,
 * numeric :328-355).  `emit` NULL means count only (symbolic phase).
 * ------------------------------------------------------------------------- */
static uint64_t spadd_row(const uint32_t *ae, uint64_t a_idx, uint64_t a_end,
                          const uint32_t *be, uint64_t b_idx, uint64_t b_end,
                          uint32_t *emit) {
  uint64_t c = 0;
  if (a_idx < a_end && b_idx < b_end) {
    uint32_t lo, hi;
    if (ae[a_idx] <= be[b_idx]) {
      lo = ae[a_idx];
      hi = ae[a_idx + 1];
      a_idx += 2;
    } else {
      lo = be[b_idx];
      hi = be[b_idx + 1];
      b_idx += 2;
    }
    while (a_idx < a_end && b_idx < b_end) {
      uint32_t rs, re;
      if (ae[a_idx] <= be[b_idx]) {
        rs = ae[a_idx];
        re = ae[a_idx + 1];
        a_idx += 2;
      } else {
        rs = be[b_idx];
        re = be[b_idx + 1];
        b_idx += 2;
      }
      if (rs <= hi + 1u) {
        lo = lo < rs ? lo : rs;
        hi = hi > re ? hi : re;
        continue;
      }
      if (emit) {
        emit[c] = lo;
        emit[c + 1] = hi;
      }
      c += 2;
      lo = rs;
      hi = re;
    }
    while (a_idx < a_end || b_idx < b_end) {
      uint32_t rs, re;
      if (a_idx == a_end) {
        rs = be[b_idx];
        re = be[b_idx + 1];
        b_idx += 2;
      } else {
        rs = ae[a_idx];
        re = ae[a_idx + 1];
        a_idx += 2;
      }
      if (rs <= hi + 1u) {
        lo = lo < rs ? lo : rs;
        hi = hi > re ? hi : re;
        continue;
      }
      if (emit) {
        emit[c] = lo;
        emit[c + 1] = hi;
      }
      c += 2;
      lo = rs;
      hi = re;
      break;
    }
    if (emit) {
      emit[c] = lo;
      emit[c + 1] = hi;
    }
    c += 2;
  }
  for (; a_idx < a_end; ++a_idx) {
    if (emit) emit[c] = ae[a_idx];
    ++c;
  }
  for (; b_idx < b_end; ++b_idx) {
    if (emit) emit[c] = be[b_idx];
    ++c;
  }
  return c;
}

/* range_sparse.hpp:364-428 (range_spadd) + :438-468 (RangeSpAdd::spadd,
 * nnz by crs_matrix::nnz). */
int orc_spadd(const orc_rcrs *a, const orc_rcrs *b, orc_rcrs *c) {
  if (a->nrows != b->nrows || a->ncols != b->ncols) return ORC_EINVAL;
  uint64_t n = a->nrows;
  uint64_t *cnt = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
  if (!cnt) return ORC_ENOMEM;
#pragma omp parallel for schedule(static)
  for (uint64_t row = 0; row < n; ++row)
    cnt[row + 1] = spadd_row(a->entries, a->rowmap[row], a->rowmap[row + 1],
                             b->entries, b->rowmap[row], b->rowmap[row + 1],
                             NULL);
  cnt[0] = 0;
  for (uint64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  int rc = alloc_rcrs(c, n, a->ncols, cnt[n]);
  if (rc) {
    free(cnt);
    return rc;
  }
  memcpy(c->rowmap, cnt, sizeof(uint64_t) * (n + 1));
  free(cnt);
#pragma omp parallel for schedule(static)
  for (uint64_t row = 0; row < n; ++row)
    spadd_row(a->entries, a->rowmap[row], a->rowmap[row + 1], b->entries,
              b->rowmap[row], b->rowmap[row + 1], c->entries + c->rowmap[row]);
  c->nnz = orc_nnz(c);
  return ORC_OK;
}

/* ---------------------------------------------------------------------------
 * HBitVector word operations: hbitvector.hpp:523-597 (64-bit words, the
 * host trait at :35-39).  sel(x, i) is the position of the i-th set bit
 * (1-based), written here from its definition: clear the i-1 lowest set bits
 * and take the index of the lowest remaining one.
 * ------------------------------------------------------------------------- */
static inline uint64_t w_cnt(uint64_t x) { return (uint64_t)__builtin_popcountll(x); }
static inline uint64_t w_msb(uint64_t x) { return x >> 63; }
static inline uint64_t w_map01(uint64_t x, uint64_t c) { return (x ^ ((x << 1) | c)) & x; }
static inline uint64_t w_map10(uint64_t x, uint64_t c) { return ((x << 1) | c) & ~x; }
static inline uint64_t w_sel(uint64_t x, uint64_t i) {
  for (uint64_t k = 1; k < i; ++k) x &= x - 1;
  return (uint64_t)__builtin_ctzll(x);
}

#define L1_BITS 2048u /* DefaultSparseConfiguration: HBitVectorAccumulator<2048> */
#define L1_WORDS (L1_BITS / 64u)

static inline uint64_t aligned_index(uint64_t i) { return i & ~(uint64_t)63; }
static inline uint64_t aligned_index_ceil(uint64_t i) { return (i + 63) & ~(uint64_t)63; }

/* hbitvector.hpp:395-402 aligned_size: at least the L1 size. */
static inline uint64_t hbv_aligned_size(uint64_t n) {
  uint64_t a = aligned_index_ceil(n);
  return a > L1_BITS ? a : L1_BITS;
}

/* hbitvector.hpp:278-299 compute_l1_begin_bidx: the L1 band is placed around
 * the row index (the "centre") and pushed left near the end of the vector.
 * The reference asserts centre < aligned size; the clamp keeps the band in
 * range for non-square products (placement never changes the output). */
static uint64_t hbv_l1_begin_bidx(uint64_t centre, uint64_t ncols) {
  uint64_t x_size = hbv_aligned_size(ncols);
  uint64_t nb = x_size >> 6;
  uint64_t ctr = centre >> 6;
  if (L1_WORDS < nb) {
    uint64_t lflank = (L1_WORDS >> 1) - 1;
    uint64_t rfit = nb - L1_WORDS;
    uint64_t pb = ctr > lflank ? ctr - lflank : 0;
    return rfit < pb ? rfit : pb;
  }
  return 0;
}

/* hbitvector.hpp:1145-1233 set(tm, s, f, UnsafeAtBoundaries): OR the bits
 * [s, f] into the words (first word mask, full middle words, last word mask).
 * The reference stores L1/L2 words under a rotated index (hbitvector.hpp:
 * 100-123); that is a storage map only, so words are indexed globally here. */
static inline void hbv_set(uint64_t *w, uint64_t s, uint64_t f) {
  uint64_t sb = s >> 6, fb = f >> 6;
  if (sb != fb) {
    w[sb] |= ~(uint64_t)0 << (s & 63);
    for (uint64_t k = sb + 1; k < fb; ++k) w[k] = ~(uint64_t)0;
    w[fb] |= ~(uint64_t)0 >> (63 - (f & 63));
  } else {
    uint64_t mask = (uint64_t)1 << ((f & 63) - (s & 63));
    mask = ((mask << 1) - 1) << (s & 63);
    w[sb] |= mask;
  }
}

/* The shared scatter loop of the TeamSequential symbolic and numeric
 * kernels (range_sparse.hpp:1192-1242 and :1349-1399): for each A range
 * [b_row, b_last_row], for each B row in it, grow the window [c_min, c_max)
 * to word-aligned bounds of the B row (clearing newly exposed words), then
 * OR in the B row's ranges. */
static void hbv_scatter_row(const orc_rcrs *a, const orc_rcrs *b, uint64_t row,
                            uint64_t *w, uint64_t *c_min_p, uint64_t *c_max_p) {
  uint64_t c_min = hbv_l1_begin_bidx(row, b->ncols) << 6;
  uint64_t c_max = c_min + L1_BITS;
  memset(w + (c_min >> 6), 0, sizeof(uint64_t) * L1_WORDS);
  for (uint64_t a_idx = a->rowmap[row]; a_idx != a->rowmap[row + 1]; a_idx += 2) {
    uint64_t b_row = a->entries[a_idx];
    uint64_t b_last_row = a->entries[a_idx + 1];
    for (; b_row <= b_last_row; ++b_row) {
      uint64_t b_idx = b->rowmap[b_row];
      uint64_t b_end = b->rowmap[b_row + 1];
      if (b_idx == b_end) continue;
      uint64_t b_min = b->entries[b_idx];
      if (b_min < c_min) {
        b_min = aligned_index(b_min);
        memset(w + (b_min >> 6), 0, sizeof(uint64_t) * ((c_min - b_min) >> 6));
        c_min = b_min;
      }
      uint64_t b_max = (uint64_t)b->entries[b_end - 1] + 1;
      if (c_max < b_max) {
        b_max = aligned_index_ceil(b_max);
        memset(w + (c_max >> 6), 0, sizeof(uint64_t) * ((b_max - c_max) >> 6));
        c_max = b_max;
      }
      for (uint64_t j = b_idx; j < b_end; j += 2)
        hbv_set(w, b->entries[j], b->entries[j + 1]);
    }
  }
  *c_min_p = c_min;
  *c_max_p = c_max;
}

/* range_sparse.hpp:1132-1277 (_range_spgemm_symbolic, TeamSequential with
 * HBitVector<2048>) + :2215-2281 (range_spgemm driver: symbolic, scan,
 * allocate, numeric) + :1287-1442 (_range_spgemm_numeric).  Rows are spread
 * over OpenMP threads, each with its own word array of the aligned column
 * count (the reference's L1 + L2 scratch per team, hbitvector.hpp:485-498). */
int orc_spgemm(const orc_rcrs *a, const orc_rcrs *b, orc_rcrs *c) {
  if (a->ncols != b->nrows) return ORC_EINVAL;
  if (b->ncols > 0xFFFFFFFEull || a->ncols > 0xFFFFFFFEull) return ORC_EINVAL;
  uint64_t n = a->nrows;
  uint64_t xwords = hbv_aligned_size(b->ncols) >> 6;
  uint64_t *cnt = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
  if (!cnt) return ORC_ENOMEM;
  uint64_t nnz = 0;
  int fail = 0;

#pragma omp parallel reduction(+ : nnz)
  {
    uint64_t *w = (uint64_t *)malloc(sizeof(uint64_t) * xwords);
    if (!w) {
#pragma omp atomic write
      fail = 1;
    }
#pragma omp for schedule(dynamic, 256)
    for (uint64_t row = 0; row < n; ++row) {
      if (!w) continue;
      uint64_t c_min, c_max;
      hbv_scatter_row(a, b, row, w, &c_min, &c_max);
      uint64_t c_lbs = c_min >> 6, c_rbs = c_max >> 6;
      uint64_t count = 0, r_nnz = 0;
      for (uint64_t j = c_lbs; j < c_rbs; ++j) {
        uint64_t cy = (j != c_lbs) ? w_msb(w[j - 1]) : 0;
        count += 2 * w_cnt(w_map01(w[j], cy));
        r_nnz += w_cnt(w[j]);
      }
      cnt[row + 1] = count;
      nnz += r_nnz;
    }
    free(w);
  }
  if (fail) {
    free(cnt);
    return ORC_ENOMEM;
  }
  cnt[0] = 0;
  for (uint64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  int rc = alloc_rcrs(c, n, b->ncols, cnt[n]);
  if (rc) {
    free(cnt);
    return rc;
  }
  memcpy(c->rowmap, cnt, sizeof(uint64_t) * (n + 1));
  free(cnt);

#pragma omp parallel
  {
    uint64_t *w = (uint64_t *)malloc(sizeof(uint64_t) * xwords);
    if (!w) {
#pragma omp atomic write
      fail = 1;
    }
#pragma omp for schedule(dynamic, 256)
    for (uint64_t row = 0; row < n; ++row) {
      if (!w) continue;
      uint64_t c_min, c_max;
      hbv_scatter_row(a, b, row, w, &c_min, &c_max);
      uint64_t c_lbs = c_min >> 6, c_rbs = c_max >> 6;
      uint64_t base = c->rowmap[row];
      uint64_t prefix = 0; /* sum of cnt01 + cnt10 of the words before j */
      for (uint64_t j = c_lbs; j < c_rbs; ++j) {
        uint64_t cy = (j != c_lbs) ? w_msb(w[j - 1]) : 0;
        uint64_t bounds = w_map01(w[j], cy) | w_map10(w[j], cy);
        uint64_t nb = w_cnt(bounds);
        for (uint64_t k = 0; k < nb; ++k) {
          uint64_t lidx = base + prefix + k;
          c->entries[lidx] = (uint32_t)((j << 6) + w_sel(bounds, k + 1) - (lidx % 2));
        }
        prefix += nb;
      }
      /* range_sparse.hpp:1435-1440: a run reaching the window end. */
      if (c_rbs > c_lbs && w_msb(w[c_rbs - 1]))
        c->entries[c->rowmap[row + 1] - 1] = (uint32_t)((c_rbs << 6) - 1);
    }
    free(w);
  }
  if (fail) {
    orc_free(c);
    return ORC_ENOMEM;
  }
  c->nnz = nnz;
  return ORC_OK;
}

static void move_rcrs(orc_rcrs *dst, orc_rcrs *src) {
  *dst = *src;
  src->rowmap = NULL;
  src->entries = NULL;
}

static int copy_rcrs(orc_rcrs *dst, const orc_rcrs *src) {
  int rc = alloc_rcrs(dst, src->nrows, src->ncols, src->rowmap[src->nrows]);
  if (rc) return rc;
  memcpy(dst->rowmap, src->rowmap, sizeof(uint64_t) * (src->nrows + 1));
  memcpy(dst->entries, src->entries, sizeof(uint32_t) * src->rowmap[src->nrows]);
  dst->nnz = src->nnz;
  return ORC_OK;
}

/* range_sparse.hpp:2370-2431 (range_power_inplace): C starts as I with
 * nnz = ncols; loop { if (k & 1) C = C * A; k >>= 1; if (k) A = A * A; else
 * break; }.  The first C * A is I * A and is executed, as in the reference. */
int orc_power(const orc_rcrs *a_in, uint32_t k, orc_rcrs *out) {
  if (a_in->nrows != a_in->ncols) return ORC_EINVAL;
  orc_rcrs a, cm, tmp;
  int rc = copy_rcrs(&a, a_in);
  if (rc) return rc;
  rc = orc_identity(a_in->ncols, &cm);
  if (rc) {
    orc_free(&a);
    return rc;
  }
  while (1) {
    if (k & 1) {
      rc = orc_spgemm(&cm, &a, &tmp);
      if (rc) goto fail;
      orc_free(&cm);
      move_rcrs(&cm, &tmp);
    }
    k >>= 1;
    if (k != 0) {
      rc = orc_spgemm(&a, &a, &tmp);
      if (rc) goto fail;
      orc_free(&a);
      move_rcrs(&a, &tmp);
    } else
      break;
  }
  orc_free(&a);
  move_rcrs(out, &cm);
  return ORC_OK;
fail:
  orc_free(&a);
  orc_free(&cm);
  return rc;
}

/* dindex.hpp:46-189 (util::create_distance_index):
 *   DIVERG_ASSERT(dlo <= dup && dup != 0)      (:66; an error code here)
 *   rai = I + A                                (:122-136, a = I, b = A)
 *   rad = A^dlo            if dlo != 0         (:151-157)
 *   raid = (A + I)^(dup - dlo)                 (:160-164)
 *   d = rad * raid if dlo != 0, else raid      (:167-177)                 */
int orc_dindex(const orc_rcrs *a, uint32_t dlo, uint32_t dup, orc_rcrs *d) {
  if (!(dlo <= dup && dup != 0)) return ORC_EINVAL;
  if (a->nrows != a->ncols) return ORC_EINVAL;
  orc_rcrs ident, rai, rad, raid;
  int rc = orc_identity(a->nrows, &ident);
  if (rc) return rc;
  rc = orc_spadd(&ident, a, &rai);
  orc_free(&ident);
  if (rc) return rc;
  if (dlo != 0) {
    rc = orc_power(a, dlo, &rad);
    if (rc) {
      orc_free(&rai);
      return rc;
    }
  }
  rc = orc_power(&rai, dup - dlo, &raid);
  orc_free(&rai);
  if (rc) {
    if (dlo != 0) orc_free(&rad);
    return rc;
  }
  if (dlo != 0) {
    rc = orc_spgemm(&rad, &raid, d);
    orc_free(&rad);
    orc_free(&raid);
    return rc;
  }
  move_rcrs(d, &raid);
  return ORC_OK;
}

/* crs_matrix.hpp:149-169 (binary_search, RangeGroup) as called by
 * CRSMatrixBase::operator()(i, j) (:1197-1218): lower bound of j in the row's
 * flat entries; an odd position steps back to the range start; the answer is
 * whether that start is <= j.  Out-of-range (i, j) give 2 (at() throws,
 * crs_matrix.hpp:1222-1243). */
static inline uint8_t query_one(const orc_rcrs *m, uint32_t i, uint32_t j) {
  if (i >= m->nrows || j >= m->ncols) return 2;
  const uint32_t *begin = m->entries + m->rowmap[i];
  const uint32_t *end = m->entries + m->rowmap[i + 1];
  const uint32_t *first = begin;
  int64_t len = end - first;
  while (len > 0) {
    int64_t half = len / 2;
    const uint32_t *middle = first + half;
    if (*middle < j) {
      first = middle + 1;
      len = len - half - 1;
    } else
      len = half;
  }
  if ((first - begin) % 2) --first;
  return (uint8_t)(first != end && *first <= j);
}

int orc_query(const orc_rcrs *m, const uint32_t *qi, const uint32_t *qj,
              uint8_t *out, uint64_t nq) {
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (uint64_t q = 0; q < nq; ++q) {
    out[q] = query_one(m, qi[q], qj[q]);
    bad |= (out[q] == 2);
  }
  return bad ? ORC_EINVAL : ORC_OK;
}

/* Independent check, not a restatement: row i of the distance index is the
 * set of vertices j reachable from i by a walk whose length lies in
 * [dlo, dup] (manuscript Eq. (1)).  Forward BFS by exact walk length over A,
 * keeping the frontier as a sorted vertex set; the union over lengths
 * dlo..dup is emitted as maximal runs.  Rows are listed in `rows`; the
 * result is a (len(rows) x ncols) rCRS matrix. */
static int cmp_u32(const void *x, const void *y) {
  uint32_t a = *(const uint32_t *)x, b = *(const uint32_t *)y;
  return (a > b) - (a < b);
}

static uint64_t uniq_u32(uint32_t *v, uint64_t n) {
  if (n == 0) return 0;
  qsort(v, n, sizeof(uint32_t), cmp_u32);
  uint64_t m = 1;
  for (uint64_t i = 1; i < n; ++i)
    if (v[i] != v[m - 1]) v[m++] = v[i];
  return m;
}

int orc_walk_rows(const orc_rcrs *a, const uint32_t *rows, uint64_t nr,
                  uint32_t dlo, uint32_t dup, orc_rcrs *out) {
  if (a->nrows != a->ncols) return ORC_EINVAL;
  uint32_t **res = (uint32_t **)calloc(nr ? nr : 1, sizeof(uint32_t *));
  uint64_t *rlen = (uint64_t *)calloc(nr ? nr : 1, sizeof(uint64_t));
  if (!res || !rlen) {
    free(res);
    free(rlen);
    return ORC_ENOMEM;
  }
  int fail = 0;
#pragma omp parallel for schedule(dynamic, 1)
  for (uint64_t r = 0; r < nr; ++r) {
    uint64_t cap = 64, fl = 1;
    uint32_t *front = (uint32_t *)malloc(cap * sizeof(uint32_t));
    uint64_t acap = 64, al = 0;
    uint32_t *acc = (uint32_t *)malloc(acap * sizeof(uint32_t));
    if (!front || !acc || rows[r] >= a->nrows) {
      free(front);
      free(acc);
#pragma omp atomic write
      fail = 1;
      continue;
    }
    front[0] = rows[r];
    for (uint32_t len = 0;; ++len) {
      if (len >= dlo) {
        if (al + fl > acap) {
          while (al + fl > acap) acap *= 2;
          acc = (uint32_t *)realloc(acc, acap * sizeof(uint32_t));
        }
        memcpy(acc + al, front, fl * sizeof(uint32_t));
        al = uniq_u32(acc, al + fl);
      }
      if (len == dup || fl == 0) break;
      uint64_t ncap = 64, nl = 0;
      uint32_t *next = (uint32_t *)malloc(ncap * sizeof(uint32_t));
      for (uint64_t f = 0; f < fl; ++f) {
        uint32_t v = front[f];
        for (uint64_t e = a->rowmap[v]; e < a->rowmap[v + 1]; e += 2)
          for (uint64_t u = a->entries[e]; u <= a->entries[e + 1]; ++u) {
            if (nl == ncap) {
              ncap *= 2;
              next = (uint32_t *)realloc(next, ncap * sizeof(uint32_t));
            }
            next[nl++] = (uint32_t)u;
          }
      }
      free(front);
      front = next;
      fl = uniq_u32(next, nl);
    }
    free(front);
    res[r] = acc;
    rlen[r] = al;
  }
  if (fail) {
    for (uint64_t r = 0; r < nr; ++r) free(res[r]);
    free(res);
    free(rlen);
    return ORC_EINVAL;
  }
  /* maximal runs */
  uint64_t total = 0;
  for (uint64_t r = 0; r < nr; ++r) {
    uint64_t runs = 0;
    for (uint64_t k = 0; k < rlen[r]; ++k)
      if (k == 0 || res[r][k] != res[r][k - 1] + 1) ++runs;
    total += 2 * runs;
  }
  int rc = alloc_rcrs(out, nr, a->ncols, total);
  if (!rc) {
    uint64_t p = 0, nnz = 0;
    for (uint64_t r = 0; r < nr; ++r) {
      for (uint64_t k = 0; k < rlen[r]; ++k) {
        if (k == 0 || res[r][k] != res[r][k - 1] + 1) {
          if (k) out->entries[p++] = res[r][k - 1];
          out->entries[p++] = res[r][k];
        }
      }
      if (rlen[r]) out->entries[p++] = res[r][rlen[r] - 1];
      out->rowmap[r + 1] = p;
      nnz += rlen[r];
    }
    out->nnz = nnz;
  }
  for (uint64_t r = 0; r < nr; ++r) free(res[r]);
  free(res);
  free(rlen);
  return rc;
}

/* Character-level adjacency of nodes [lower, upper), restating
 * util::adjacency_matrix (adjacency.hpp:170-228): row l -> l + 1 inside a
 * node; the last locus of a node -> the first locus of each out-neighbour,
 * sorted (std::sort, duplicates kept); columns relative to the first locus of
 * node `lower`.  The Basic rows are then converted to ranges as
 * append(..., RangeGroup, BasicGroup) does (crs_matrix.hpp:197-242): a
 * column extends the open range only when it equals the range's end + 1.
 * nnz is the number of Basic entries.  Edges must stay inside the range. */
int orc_adjacency(uint64_t n_nodes, const uint32_t *node_len, const uint64_t *eoff,
                  const uint32_t *etgt, uint64_t lower, uint64_t upper, orc_rcrs *out) {
  if (upper == 0) upper = n_nodes;
  if (lower > upper || upper > n_nodes) return ORC_EINVAL;
  uint64_t *start = (uint64_t *)malloc(sizeof(uint64_t) * (upper - lower + 1));
  if (!start) return ORC_ENOMEM;
  start[0] = 0;
  for (uint64_t k = lower; k < upper; ++k) start[k - lower + 1] = start[k - lower] + node_len[k];
  const uint64_t n = start[upper - lower];
  uint64_t basic = 0, max_deg = 0;
  for (uint64_t k = lower; k < upper; ++k) {
    basic += node_len[k] - 1 + (eoff[k + 1] - eoff[k]);
    if (eoff[k + 1] - eoff[k] > max_deg) max_deg = eoff[k + 1] - eoff[k];
    for (uint64_t q = eoff[k]; q < eoff[k + 1]; ++q)
      if (etgt[q] < lower || etgt[q] >= upper) {
        free(start);
        return ORC_EINVAL;
      }
  }
  uint32_t *cols = (uint32_t *)malloc(sizeof(uint32_t) * (max_deg ? max_deg : 1));
  if (!cols || alloc_rcrs(out, n, n, 2 * basic)) {
    free(start);
    free(cols);
    return ORC_ENOMEM;
  }
  uint64_t w = 0, row = 0;
  for (uint64_t k = lower; k < upper; ++k) {
    const uint64_t s0 = start[k - lower];
    for (uint64_t j = 0; j + 1 < node_len[k]; ++j) {  /* inside the node */
      out->entries[w++] = (uint32_t)(s0 + j + 1);
      out->entries[w++] = (uint32_t)(s0 + j + 1);
      out->rowmap[++row] = w;
    }
    const uint64_t deg = eoff[k + 1] - eoff[k];
    for (uint64_t q = 0; q < deg; ++q) cols[q] = (uint32_t)start[etgt[eoff[k] + q] - lower];
    qsort(cols, deg, sizeof(uint32_t), cmp_u32);
    int pending = 0;
    uint32_t pend = 0;
    for (uint64_t q = 0; q < deg; ++q) {
      if (pending && cols[q] == pend + 1) {
        ++pend;
      } else {
        if (pending) out->entries[w++] = pend;
        out->entries[w++] = cols[q];
        pend = cols[q];
        pending = 1;
      }
    }
    if (pending) out->entries[w++] = pend;
    out->rowmap[++row] = w;
  }
  out->nnz = basic;
  free(start);
  free(cols);
  return ORC_OK;
}

/* The streaming merge of two Range-group distance indices, restating
 * merge_distance_index(..., RangeGroup, std::true_type) (dindex.hpp:290-420):
 * per row, while both have ranges, take the one with the smaller (start,
 * end), then absorb ranges from either side while they touch (u + 1 >= l);
 * when one side is exhausted the rest of the other is copied verbatim.  nnz
 * is the sum of the emitted range lengths. */
int orc_merge_stream(const orc_rcrs *a, const orc_rcrs *b, orc_rcrs *c) {
  if (a->nrows != b->nrows || a->ncols != b->ncols) return ORC_EINVAL;
  const uint64_t cap = a->rowmap[a->nrows] + b->rowmap[b->nrows];
  if (alloc_rcrs(c, a->nrows, a->ncols, cap)) return ORC_ENOMEM;
  const uint32_t *e1 = a->entries, *e2 = b->entries;
  uint64_t c1 = 0, c2 = 0, w = 0, nnz = 0;
  for (uint64_t r = 0; r < a->nrows; ++r) {
    const uint64_t end1 = a->rowmap[r + 1], end2 = b->rowmap[r + 1];
    while (c1 < end1) {
      if (c2 >= end2) {
        while (c1 < end1) {
          const uint32_t l = e1[c1++], u = e1[c1++];
          c->entries[w++] = l;
          c->entries[w++] = u;
          nnz += (uint64_t)u - l + 1;
        }
        break;
      }
      uint32_t l, u;
      if (e1[c1] < e2[c2] || (e1[c1] == e2[c2] && e1[c1 + 1] < e2[c2 + 1])) {
        l = e1[c1];
        u = e1[c1 + 1];
        c1 += 2;
      } else {
        l = e2[c2];
        u = e2[c2 + 1];
        c2 += 2;
      }
      for (;;) {
        const uint64_t l1 = c1 < end1 ? e1[c1] : UINT32_MAX, u1 = c1 < end1 ? e1[c1 + 1] : UINT32_MAX;
        const uint64_t l2 = c2 < end2 ? e2[c2] : UINT32_MAX, u2 = c2 < end2 ? e2[c2 + 1] : UINT32_MAX;
        if ((uint64_t)u + 1 >= l1) {
          if (l1 < l) l = (uint32_t)l1;
          if (u1 > u) u = (uint32_t)u1;
          c1 += 2;
        } else if ((uint64_t)u + 1 >= l2) {
          if (l2 < l) l = (uint32_t)l2;
          if (u2 > u) u = (uint32_t)u2;
          c2 += 2;
        } else {
          break;
        }
      }
      c->entries[w++] = l;
      c->entries[w++] = u;
      nnz += (uint64_t)u - l + 1;
    }
    while (c2 < end2) {
      const uint32_t l = e2[c2++], u = e2[c2++];
      c->entries[w++] = l;
      c->entries[w++] = u;
      nnz += (uint64_t)u - l + 1;
    }
    c->rowmap[r + 1] = w;
  }
  c->nnz = nnz;
  return ORC_OK;
}

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_threads(int t) { omp_set_num_threads(t); }
