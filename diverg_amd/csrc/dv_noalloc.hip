// Range SpAdd that allocates nothing: its scratch (per-row counts, scan tile
// sums, counters and the status word) lives in one file-scope device array
// of DV_NOALLOC_SCRATCH_BYTES, present on every GPU the library runs on for
// the life of the process.  For callers that cannot allocate device memory
// (code under stream capture, serving loops).  The array's device address is
// resolved once per device (dv_init, or the first call on a device) and only
// that address is handed to kernels and copies; the host-side symbol is not
// a device address.
//
// Every call on a GPU uses that GPU's one array: calls that overlap on one
// GPU race.  Counters and status are cleared by a kernel at the start of each
// call; the counts are written in full by the count pass before the scan
// reads them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "dv_runtime.hpp"

namespace {

__device__ __align__(256) unsigned char scratch[DV_NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void *scratch_ptr[MAX_DEVICES] = {};

constexpr uint64_t CTR_OFF = 0, STATUS_OFF = 64, TILES_OFF = 256;

uint64_t align256(uint64_t x) { return (x + 255) & ~255ull; }

// Bytes of scratch a sum of n-row matrices needs.
uint64_t scratch_bytes(uint64_t n) {
  return TILES_OFF + align256(8 * dv::scan_words(n)) + 4 * std::max<uint64_t>(1, n);
}

}  // namespace

namespace dv {

// Resolve the scratch array's address on `device` (made current for the
// call; the caller's current device is restored).
int noalloc_setup(int device) {
  if (device < 0 || device >= MAX_DEVICES) return DV_ERR_INVALID;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (scratch_ptr[device]) return DV_OK;
  int old = -1;
  (void)hipGetDevice(&old);
  void *p = nullptr;
  const bool ok = hipSetDevice(device) == hipSuccess &&
                  hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess && p;
  if (old >= 0) (void)hipSetDevice(old);
  if (!ok) {
    (void)hipGetLastError();
    return DV_ERR_DEVICE;
  }
  scratch_ptr[device] = p;
  return DV_OK;
}

}  // namespace dv

extern "C" {

uint64_t dv_range_spadd_noalloc_max_rows(void) {
  // the largest n with scratch_bytes(n) <= DV_NOALLOC_SCRATCH_BYTES
  uint64_t lo = 0, hi = DV_NOALLOC_SCRATCH_BYTES / 4;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo + 1) / 2;
    if (scratch_bytes(mid) <= DV_NOALLOC_SCRATCH_BYTES) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

int dv_range_spadd_noalloc(dv_context *ctx, const dv_rcrs *a, const dv_rcrs *b, void *c_rowmap,
                           uint32_t *c_entries, uint64_t c_capacity, uint64_t *c_n_entries,
                           uint64_t *c_nnz, void *stream) {
  if (!ctx || !c_n_entries || !c_nnz) return DV_ERR_INVALID;
  dv::Ctx *c = ctx;
  if (!a || !b) return c->fail("spadd_noalloc: a or b is NULL");
  if (a->location != DV_DEVICE || b->location != DV_DEVICE)
    return c->fail("spadd_noalloc: a and b must be device resident");
  if (a->nrows != b->nrows || a->ncols != b->ncols) return c->fail("spadd_noalloc: shape mismatch");
  if (a->size_bytes != b->size_bytes || (a->size_bytes != 4 && a->size_bytes != 8))
    return c->fail("spadd_noalloc: size_bytes must be 4 or 8, the same for a and b");
  if (a->ncols > 0xFFFFFFFEull || a->nrows > 0xFFFFFFFEull)
    return c->fail("spadd_noalloc: more than uint32_t max - 1 rows or columns", DV_ERR_OVERFLOW);
  if ((a->n_entries % 2) || (b->n_entries % 2)) return c->fail("spadd_noalloc: odd entry count");
  if (!a->rowmap || !b->rowmap || (a->n_entries && !a->entries) || (b->n_entries && !b->entries))
    return c->fail("spadd_noalloc: NULL rowmap or entries");
  if ((reinterpret_cast<uintptr_t>(a->entries) & 7) || (reinterpret_cast<uintptr_t>(b->entries) & 7) ||
      (reinterpret_cast<uintptr_t>(c_entries) & 7))
    return c->fail("spadd_noalloc: entries must be 8-byte aligned");
  if (!c_rowmap || (c_capacity && !c_entries))
    return c->fail("spadd_noalloc: NULL output rowmap or entries");
  const uint64_t n = a->nrows;
  if (scratch_bytes(n) > DV_NOALLOC_SCRATCH_BYTES)
    return c->fail("spadd_noalloc: " + std::to_string(n) + " rows exceed the scratch of " +
                       std::to_string(dv_range_spadd_noalloc_max_rows()) + " rows",
                   DV_ERR_BUDGET);
  int rc = dv::noalloc_setup(c->device);
  if (rc) return c->fail("spadd_noalloc: scratch address not resolved", rc);
  unsigned char *base = static_cast<unsigned char *>(scratch_ptr[c->device]);
  auto *ctr = reinterpret_cast<dv::OpCounters *>(base + CTR_OFF);
  auto *status = reinterpret_cast<unsigned int *>(base + STATUS_OFF);
  auto *tiles = reinterpret_cast<unsigned long long *>(base + TILES_OFF);
  auto *cnt = reinterpret_cast<uint32_t *>(base + TILES_OFF + align256(8 * dv::scan_words(n)));
  hipStream_t st = stream ? static_cast<hipStream_t>(stream) : c->stream;
  const uint32_t sb = a->size_bytes;
  const dv::DView va{a->rowmap, a->entries, a->nrows, a->ncols};
  const dv::DView vb{b->rowmap, b->entries, b->nrows, b->ncols};

  // clear counters and status, validate the operands (status 1 when malformed)
  dv::fill_u32(st, reinterpret_cast<uint32_t *>(base), TILES_OFF / 4, 0u);
  if (n) {
    dv::validate_matrix(st, sb, va, a->n_entries, status);
    dv::validate_matrix(st, sb, vb, b->n_entries, status);
  }
  unsigned int h_status = 0;
  if (!c->hip(hipMemcpyAsync(&h_status, status, sizeof(h_status), hipMemcpyDeviceToHost, st),
              "spadd_noalloc status") ||
      !c->hip(hipStreamSynchronize(st), "spadd_noalloc synchronize"))
    return DV_ERR_DEVICE;
  if (h_status)
    return c->fail("spadd_noalloc: malformed device matrix (rowmap not monotone/even, "
                   "rowmap[n] != n_entries, or a range outside [0, ncols))");

  // count, row pointers (status 1 when the entries exceed c_capacity, 2 when
  // they exceed a 32-bit rowmap), write, nnz: range_sparse.hpp:112-428
  dv::spadd_count(st, sb, va, vb, cnt, status);
  dv::scan_rowmap(st, sb, cnt, n, c_rowmap, tiles, ctr, c_capacity, status, nullptr, 0, nullptr);
  dv::spadd_write(st, sb, va, vb, c_rowmap, c_entries, status);
  dv::range_nnz(st, c_entries, c_capacity, ctr, status);
  dv::OpCounters h{};
  if (!c->hip(hipGetLastError(), "spadd_noalloc launch") ||
      !c->hip(hipMemcpyAsync(&h, ctr, sizeof(h), hipMemcpyDeviceToHost, st),
              "spadd_noalloc readback") ||
      !c->hip(hipMemcpyAsync(&h_status, status, sizeof(h_status), hipMemcpyDeviceToHost, st),
              "spadd_noalloc status") ||
      !c->hip(hipStreamSynchronize(st), "spadd_noalloc synchronize"))
    return DV_ERR_DEVICE;
  *c_n_entries = h.total;
  if (h_status & 2) return c->fail("spadd_noalloc: entry count exceeds a 32-bit rowmap", DV_ERR_OVERFLOW);
  if (h_status & 1)
    return c->fail("spadd_noalloc: the sum has " + std::to_string(h.total) +
                       " entries, more than c_capacity",
                   DV_ERR_OVERFLOW);
  *c_nnz = sb == 4 ? (uint64_t)(uint32_t)h.nnz : (uint64_t)h.nnz;
  return DV_OK;
}

}  // extern "C"
