// C-ABI implementation: device context, matrix movement, and the host
// orchestration of range SpAdd / SpGEMM / power / distance-index build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <unordered_map>

#include "dv_internal.hpp"
#include "dv_runtime.hpp"
#include "rcrs_kernels.hpp"

using dv::DevMat;

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
namespace dv {

bool Ctx::hip(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  (void)hipGetLastError();
  fail(std::string(what) + ": " + hipGetErrorString(e));
  return false;
}

int Ctx::fail(const std::string &msg, int code) {
  last_error = msg;
  return code;
}

void *Ctx::dalloc(uint64_t bytes) {
  if (bytes == 0) bytes = 8;
  void *p = nullptr;
  hipError_t e = hipMallocAsync(&p, (size_t)bytes, stream);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  sizes[p] = bytes;
  cur_bytes += bytes;
  peak_bytes = std::max(peak_bytes, cur_bytes);
  return p;
}

void Ctx::dfree(void *p) {
  if (!p) return;
  auto it = sizes.find(p);
  if (it != sizes.end()) {
    cur_bytes -= it->second;
    sizes.erase(it);
  }
  (void)hipFreeAsync(p, stream);
}

void Ctx::release(DevMat &m) {
  if (m.owned) {
    dfree(m.rowmap);
    dfree(m.entries);
  }
  m = DevMat{};
}

bool Ctx::read_counters(OpCounters &out) {
  if (!hip(hipMemcpyAsync(h_ctr, d_ctr, sizeof(OpCounters), hipMemcpyDeviceToHost, stream),
           "counter readback"))
    return false;
  if (!hip(hipStreamSynchronize(stream), "stream synchronize")) return false;
  out = *h_ctr;
  return true;
}

bool Ctx::scratch_carve(const uint64_t *bytes, void **out, int n) {
  uint64_t total = 0;
  for (int i = 0; i < n; ++i) total += (bytes[i] + 255) / 256 * 256;
  if (total > scratch_cap) {
    dfree(scratch);
    scratch = nullptr;
    scratch_cap = 0;
    const uint64_t want = total + total / 4;
    scratch = dalloc(want);
    if (!scratch) return false;
    scratch_cap = want;
  }
  auto *p = static_cast<char *>(scratch);
  for (int i = 0; i < n; ++i) {
    out[i] = p;
    p += (bytes[i] + 255) / 256 * 256;
  }
  return true;
}

uint64_t mat_bytes(const DevMat &m) { return (uint64_t)m.sb * (m.nrows + 1) + 4 * m.n_entries; }

DView view(const DevMat &m) { return DView{m.rowmap, m.entries, m.nrows, m.ncols}; }

}  // namespace dv

using dv::Ctx;

struct dv_context : public Ctx {};

namespace {

Ctx *C(dv_context *c) { return static_cast<Ctx *>(c); }

bool valid_sb(uint32_t sb) { return sb == 4 || sb == 8; }

uint64_t rm_at(const void *rm, uint32_t sb, uint64_t i) {
  return sb == 4 ? static_cast<const uint32_t *>(rm)[i] : static_cast<const uint64_t *>(rm)[i];
}

// Shape checks shared by every entry point.  Host matrices are also checked
// element by element; device matrices are checked by the validation kernel.
int check_desc(Ctx *c, const dv_rcrs *m, const char *name) {
  if (!m) return c->fail(std::string(name) + " is NULL");
  if (!valid_sb(m->size_bytes))
    return c->fail(std::string(name) + ": size_bytes must be 4 or 8");
  if (m->location != DV_HOST && m->location != DV_DEVICE)
    return c->fail(std::string(name) + ": bad location");
  if (m->n_entries % 2) return c->fail(std::string(name) + ": odd entry count");
  if (m->ncols > 0xFFFFFFFEull)
    return c->fail(std::string(name) + ": ncols exceeds uint32_t max - 1", DV_ERR_OVERFLOW);
  if (m->nrows > 0xFFFFFFFEull)
    return c->fail(std::string(name) + ": nrows exceeds uint32_t max - 1", DV_ERR_OVERFLOW);
  if (!m->rowmap) return c->fail(std::string(name) + ": rowmap is NULL");
  if (m->n_entries && !m->entries) return c->fail(std::string(name) + ": entries is NULL");
  if (m->location == DV_DEVICE && m->n_entries &&
      (reinterpret_cast<uintptr_t>(m->entries) & 7))
    return c->fail(std::string(name) + ": device entries must be 8-byte aligned");
  if (m->location == DV_HOST) {
    if (rm_at(m->rowmap, m->size_bytes, 0) != 0 ||
        rm_at(m->rowmap, m->size_bytes, m->nrows) != m->n_entries)
      return c->fail(std::string(name) + ": rowmap[0] != 0 or rowmap[n] != n_entries");
    for (uint64_t i = 0; i < m->nrows; ++i) {
      uint64_t b = rm_at(m->rowmap, m->size_bytes, i), e = rm_at(m->rowmap, m->size_bytes, i + 1);
      if (e < b || (e - b) % 2) return c->fail(std::string(name) + ": malformed rowmap");
    }
    for (uint64_t k = 0; k < m->n_entries; k += 2)
      if (m->entries[k] > m->entries[k + 1] || m->entries[k + 1] >= m->ncols)
        return c->fail(std::string(name) + ": range outside [0, ncols) or start > end");
  }
  return DV_OK;
}

// Bring a descriptor onto the device (a view when it already is there).
int to_dev(Ctx *c, const dv_rcrs *m, DevMat &out, double *h2d_ms) {
  out = DevMat{};
  out.nrows = m->nrows;
  out.ncols = m->ncols;
  out.n_entries = m->n_entries;
  out.nnz = m->nnz;
  out.sb = m->size_bytes;
  if (m->location == DV_DEVICE) {
    out.rowmap = m->rowmap;
    out.entries = m->entries;
    out.owned = false;
    return c->validate(out);
  }
  hipEvent_t e0 = c->ev[0], e1 = c->ev[1];
  (void)hipEventRecord(e0, c->stream);
  out.rowmap = c->dalloc(out.sb * (out.nrows + 1));
  out.entries = static_cast<uint32_t *>(c->dalloc(4 * out.n_entries));
  out.owned = true;
  if (!out.rowmap || !out.entries) {
    c->release(out);
    return c->fail("device allocation failed for an input", DV_ERR_NOMEM);
  }
  if (!c->hip(hipMemcpyAsync(out.rowmap, m->rowmap, out.sb * (out.nrows + 1),
                             hipMemcpyHostToDevice, c->stream),
              "H2D rowmap") ||
      (out.n_entries &&
       !c->hip(hipMemcpyAsync(out.entries, m->entries, 4 * out.n_entries,
                              hipMemcpyHostToDevice, c->stream),
               "H2D entries"))) {
    c->release(out);
    return DV_ERR_DEVICE;
  }
  (void)hipEventRecord(e1, c->stream);
  if (h2d_ms) {
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    *h2d_ms += ms;
  }
  return DV_OK;
}

// Hand a device matrix to the caller at `location` (D2H copy for host).
int export_mat(Ctx *c, DevMat &m, uint32_t location, dv_rcrs *out, double *d2h_ms) {
  std::memset(out, 0, sizeof(*out));
  out->nrows = m.nrows;
  out->ncols = m.ncols;
  out->n_entries = m.n_entries;
  out->nnz = m.sb == 4 ? (uint32_t)m.nnz : m.nnz;
  out->size_bytes = m.sb;
  out->location = location;
  if (location == DV_DEVICE) {
    if (!m.owned) {  // the result aliases an input: give the caller a copy
      DevMat cp;
      int rc = c->clone(m, cp);
      if (rc) return rc;
      m = cp;
    }
    out->rowmap = m.rowmap;
    out->entries = m.entries;
    out->owner = new dv::Owner{DV_DEVICE, m.rowmap, m.entries, m.sb * (m.nrows + 1),
                               4 * m.n_entries};
    // ownership moves to the descriptor; the pool accounting forgets it
    c->forget(m.rowmap);
    c->forget(m.entries);
    m = DevMat{};
    return DV_OK;
  }
  void *rm = std::malloc(m.sb * (m.nrows + 1));
  uint32_t *en = static_cast<uint32_t *>(std::malloc(4 * (m.n_entries ? m.n_entries : 1)));
  if (!rm || !en) {
    std::free(rm);
    std::free(en);
    return c->fail("host allocation failed for an output", DV_ERR_NOMEM);
  }
  (void)hipEventRecord(c->ev[0], c->stream);
  bool ok = c->hip(hipMemcpyAsync(rm, m.rowmap, m.sb * (m.nrows + 1), hipMemcpyDeviceToHost,
                                  c->stream),
                   "D2H rowmap") &&
            (m.n_entries == 0 ||
             c->hip(hipMemcpyAsync(en, m.entries, 4 * m.n_entries, hipMemcpyDeviceToHost,
                                   c->stream),
                    "D2H entries"));
  (void)hipEventRecord(c->ev[1], c->stream);
  ok = ok && c->hip(hipStreamSynchronize(c->stream), "D2H synchronize");
  if (!ok) {
    std::free(rm);
    std::free(en);
    return DV_ERR_DEVICE;
  }
  if (d2h_ms) {
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]);
    *d2h_ms += ms;
  }
  out->rowmap = rm;
  out->entries = en;
  out->owner = new dv::Owner{DV_HOST, rm, en, m.sb * (m.nrows + 1), 4 * m.n_entries};
  c->release(m);
  return DV_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// device-side building blocks (dv_runtime.hpp)
// ---------------------------------------------------------------------------
namespace dv {

int Ctx::validate(const DevMat &m) {
  if (m.nrows == 0 && m.n_entries == 0) return DV_OK;
  fill_u32(stream, d_flag, 1, 0u);
  validate_matrix(stream, m.sb, view(m), m.n_entries, d_flag);
  if (!hip(hipMemcpyAsync(h_flag, d_flag, sizeof(uint32_t), hipMemcpyDeviceToHost, stream),
           "validate readback") ||
      !hip(hipStreamSynchronize(stream), "validate synchronize"))
    return DV_ERR_DEVICE;
  if (*h_flag)
    return fail("malformed device matrix (rowmap not monotone/even, rowmap[n] != "
                "n_entries, or a range outside [0, ncols))");
  return DV_OK;
}

int Ctx::clone(const DevMat &src, DevMat &dst) {
  dst = src;
  dst.owned = true;
  dst.rowmap = dalloc(src.sb * (src.nrows + 1));
  dst.entries = static_cast<uint32_t *>(dalloc(4 * src.n_entries));
  if (!dst.rowmap || !dst.entries) {
    release(dst);
    return fail("device allocation failed", DV_ERR_NOMEM);
  }
  if (!hip(hipMemcpyAsync(dst.rowmap, src.rowmap, src.sb * (src.nrows + 1),
                          hipMemcpyDeviceToDevice, stream),
           "D2D rowmap") ||
      (src.n_entries && !hip(hipMemcpyAsync(dst.entries, src.entries, 4 * src.n_entries,
                                            hipMemcpyDeviceToDevice, stream),
                             "D2D entries"))) {
    release(dst);
    return DV_ERR_DEVICE;
  }
  return DV_OK;
}

void Ctx::forget(void *p) {
  auto it = sizes.find(p);
  if (it != sizes.end()) {
    cur_bytes -= it->second;
    sizes.erase(it);
  }
}

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
  return hip(hipGetLastError(), "identity launch") ? DV_OK : DV_ERR_DEVICE;
}

int Ctx::spadd_dev(const DevMat &a, const DevMat &b, DevMat &c) {
  if (a.nrows != b.nrows || a.ncols != b.ncols || a.sb != b.sb)
    return fail("spadd: operands differ in shape or size type");
  const uint64_t n = a.nrows;
  c = DevMat{};
  c.nrows = n;
  c.ncols = a.ncols;
  c.sb = a.sb;
  c.owned = true;
  void *buf[2];
  const uint64_t sz[2] = {4 * (n ? n : 1), 8 * scan_tmp_words(n)};
  const bool got = scratch_carve(sz, buf, 2);
  auto *cnt = static_cast<uint32_t *>(buf[0]);
  auto *tmp = static_cast<unsigned long long *>(buf[1]);
  c.rowmap = dalloc(c.sb * (n + 1));
  if (!got || !c.rowmap) {
    release(c);
    return fail("device allocation failed (spadd)", DV_ERR_NOMEM);
  }
  zero_counters(stream, d_ctr);
  prof_mark(0);
  spadd_count(stream, a.sb, view(a), view(b), cnt);
  scan_counts(stream, c.sb, cnt, n, c.rowmap, tmp, d_ctr);
  OpCounters h{};
  if (!read_counters(h)) {
    release(c);
    return DV_ERR_DEVICE;
  }
  if (c.sb == 4 && h.total > 0xFFFFFFFFull) {
    release(c);
    return fail("spadd: entry count exceeds a 32-bit rowmap", DV_ERR_OVERFLOW);
  }
  c.n_entries = h.total;
  c.entries = static_cast<uint32_t *>(dalloc(4 * c.n_entries));
  if (!c.entries) {
    release(c);
    return fail("device allocation failed (spadd entries)", DV_ERR_NOMEM);
  }
  spadd_write(stream, c.sb, view(a), view(b), c.rowmap, c.entries);
  zero_counters(stream, d_ctr);
  range_nnz(stream, c.entries, c.n_entries, d_ctr);
  prof_mark(1);
  if (!read_counters(h)) {
    release(c);
    return DV_ERR_DEVICE;
  }
  c.nnz = h.nnz;
  prof_add(3, 0, 1, mat_bytes(a) + mat_bytes(b) + mat_bytes(c), 0);
  stats_alg_bytes += mat_bytes(a) + mat_bytes(b) + mat_bytes(c);
  stats_n_spadd += 1;
  return DV_OK;
}

GroupLaunch Ctx::group_launch(uint32_t lanes, uint64_t nstage) const {
  GroupLaunch g{};
  g.lanes = lanes;
  g.lds_words = std::max<uint32_t>(1, tuning.lds_words * lanes / 64);
  g.stage = tuning.stage_entries;
  g.rows_per_group = std::max<uint32_t>(1, tuning.rows_per_group);
  g.grid = std::max<uint32_t>(8, (num_cu * 8 + 7) / 8 * 8);
  g.nstage = nstage;
  return g;
}

int Ctx::spgemm_dev(const DevMat &a, const DevMat &b, DevMat &c) {
  if (a.ncols != b.nrows) return fail("spgemm: a.ncols != b.nrows");
  if (a.sb != b.sb) return fail("spgemm: operands differ in size type");
  if (a.ncols > 0xFFFFFFFEull || b.ncols > 0xFFFFFFFEull)
    return fail("spgemm: ncols exceeds uint32_t max - 1", DV_ERR_OVERFLOW);
  const uint64_t n = a.nrows;
  c = DevMat{};
  c.nrows = n;
  c.ncols = b.ncols;
  c.sb = a.sb;
  c.owned = true;
  const uint32_t KS = tuning.stage_entries;
  const uint64_t n1 = n ? n : 1;
  void *buf[6];
  const uint64_t sz[6] = {4 * n1, 4ull * KS * n1, 4 * n1, 4 * n1, 4 * n1, 8 * scan_tmp_words(n)};
  const bool got = scratch_carve(sz, buf, 6);
  auto *cnt = static_cast<uint32_t *>(buf[0]);
  auto *stage = static_cast<uint32_t *>(buf[1]);
  auto *ovf = static_cast<uint32_t *>(buf[2]);
  auto *mid = static_cast<uint32_t *>(buf[3]);
  auto *big = static_cast<uint32_t *>(buf[4]);
  auto *tmp = static_cast<unsigned long long *>(buf[5]);
  c.rowmap = dalloc(c.sb * (n + 1));
  if (!got || !c.rowmap) {
    release(c);
    return fail("device allocation failed (spgemm)", DV_ERR_NOMEM);
  }
  zero_counters(stream, d_ctr);
  prof_mark(0);
  // persistent lane-tier grid: few blocks keep the per-block counter atomics
  // (nnz, slice ranges) uncontended
  const uint32_t lane_grid =
      (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, (uint64_t)num_cu * 8));
  spgemm_lane(stream, a.sb, view(a), view(b), KS, tuning.lane_slice_max, tuning.small_slice_max,
              lane_grid, cnt, stage, ovf, mid, big, d_ctr);
  if (n) {
    GroupLaunch gm = group_launch(tuning.lanes_small, n);
    gm.rows = mid;
    gm.n_dev = &d_ctr->n_mid;
    gm.cnt = cnt;
    gm.stage_buf = stage;
    gm.ovf_list = ovf;
    gm.ctr = d_ctr;
    spgemm_group(stream, a.sb, view(a), view(b), gm);
    GroupLaunch gb = gm;
    GroupLaunch g64 = group_launch(64, n);
    gb.lanes = 64;
    gb.lds_words = g64.lds_words;
    gb.rows = big;
    gb.n_dev = &d_ctr->n_big;
    spgemm_group(stream, a.sb, view(a), view(b), gb);
  }
  prof_mark(1);
  scan_counts(stream, c.sb, cnt, n, c.rowmap, tmp, d_ctr);
  prof_mark(2);
  OpCounters h{};
  if (!hip(hipGetLastError(), "spgemm count launch") || !read_counters(h)) {
    release(c);
    return DV_ERR_DEVICE;
  }
  if (c.sb == 4 && h.total > 0xFFFFFFFFull) {
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
  spgemm_emit_staged(stream, c.sb, n, c.rowmap, stage, KS, c.entries);
  if (h.overflow) {
    GroupLaunch gd = group_launch(64, n);
    gd.rows = ovf;
    gd.n_host = h.overflow;
    gd.c_rowmap = c.rowmap;
    gd.c_entries = c.entries;
    spgemm_group(stream, a.sb, view(a), view(b), gd);
  }
  prof_mark(4);
  if (!hip(hipGetLastError(), "spgemm write launch")) {
    release(c);
    return DV_ERR_DEVICE;
  }
  {
    const uint64_t alg = mat_bytes(a) + mat_bytes(b) + mat_bytes(c);
    prof_add(0, 0, 1, alg, 16ull * h.slice);
    prof_add(2, 1, 2, 0, 0);
    prof_add(1, 3, 4, 0, 0);
  }
  stats_alg_bytes += mat_bytes(a) + mat_bytes(b) + mat_bytes(c);
  stats_n_spgemm += 1;
  stats_overflow_rows += h.overflow;
  stats_mid_rows += h.n_mid;
  stats_big_rows += h.n_big;
  return DV_OK;
}

// range_sparse.hpp:2376-2431: C = I; loop { if (k & 1) C = C.A; k >>= 1;
// if (k) A = A.A; else break; }.  `a` is not modified.
int Ctx::power_dev(const DevMat &a_in, uint32_t k, DevMat &out) {
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

// dindex.hpp:46-189.
int Ctx::dindex_dev(const DevMat &a, uint32_t dlo, uint32_t dup, DevMat &d) {
  if (!(dlo <= dup && dup != 0)) return fail("create_distance_index: need dlo <= dup and dup != 0");
  if (a.nrows != a.ncols) return fail("create_distance_index: matrix is not square");
  DevMat ident, rai, rad, raid;
  int rc = identity_dev(a.nrows, a.sb, ident);
  if (rc) return rc;
  rc = spadd_dev(ident, a, rai);
  release(ident);
  if (rc) return rc;
  if (dlo != 0) {
    rc = power_dev(a, dlo, rad);
    if (rc) {
      release(rai);
      return rc;
    }
  }
  rc = power_dev(rai, dup - dlo, raid);
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

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int dv_abi_version(void) { return DV_ABI_VERSION; }

const char *dv_status_string(int s) {
  switch (s) {
    case DV_OK: return "ok";
    case DV_ERR_INVALID: return "invalid argument";
    case DV_ERR_OVERFLOW: return "overflow";
    case DV_ERR_NOMEM: return "out of memory";
    case DV_ERR_DEVICE: return "device error";
    case DV_ERR_BUDGET: return "HBM budget exceeded";
    case DV_ERR_RANGE: return "index out of range";
    case DV_ERR_BUSY: return "busy";
    case DV_ERR_NOT_FOUND: return "not found";
    default: return "unknown status";
  }
}

int dv_init(int device, dv_context **out) {
  if (!out) return DV_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return DV_ERR_DEVICE;
  }
  if (device < 0 || device >= ndev) return DV_ERR_INVALID;
  dv_context *c = new (std::nothrow) dv_context();
  if (!c) return DV_ERR_NOMEM;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return DV_ERR_DEVICE;
  }
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->num_cu = (uint32_t)prop.multiProcessorCount;
    c->lds_per_block = (uint32_t)prop.sharedMemPerBlock;
  }
  hipMemPool_t pool;
  if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
    uint64_t thr = UINT64_MAX;
    (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
  }
  (void)hipGetLastError();
  bool ok = hipEventCreate(&c->ev[0]) == hipSuccess && hipEventCreate(&c->ev[1]) == hipSuccess &&
            hipEventCreate(&c->ev[2]) == hipSuccess && hipEventCreate(&c->ev[3]) == hipSuccess &&
            hipHostMalloc((void **)&c->h_ctr, sizeof(dv::OpCounters), hipHostMallocDefault) == hipSuccess &&
            hipHostMalloc((void **)&c->h_flag, sizeof(uint32_t), hipHostMallocDefault) == hipSuccess &&
            hipMalloc((void **)&c->d_ctr, sizeof(dv::OpCounters)) == hipSuccess &&
            hipMalloc((void **)&c->d_flag, sizeof(uint32_t)) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    dv_finalize(c);
    return DV_ERR_DEVICE;
  }
  *out = c;
  return DV_OK;
}

int dv_finalize(dv_context *c) {
  if (!c) return DV_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &kv : c->sizes) (void)hipFreeAsync(kv.first, c->stream);
  c->sizes.clear();
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : c->pev)
    if (e) (void)hipEventDestroy(e);
  if (c->h_ctr) (void)hipHostFree(c->h_ctr);
  if (c->h_flag) (void)hipHostFree(c->h_flag);
  if (c->d_ctr) (void)hipFree(c->d_ctr);
  if (c->d_flag) (void)hipFree(c->d_flag);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  (void)hipGetLastError();
  delete c;
  return DV_OK;
}

const char *dv_last_error(const dv_context *c) { return c ? c->last_error.c_str() : ""; }

void *dv_stream(dv_context *c) { return c ? (void *)c->stream : nullptr; }

int dv_get_tuning(const dv_context *c, dv_tuning *t) {
  if (!c || !t) return DV_ERR_INVALID;
  *t = c->tuning;
  return DV_OK;
}

int dv_set_tuning(dv_context *c, const dv_tuning *t) {
  if (!c || !t) return DV_ERR_INVALID;
  const dv_tuning n = *t;
  const uint32_t ls = n.lanes_small;
  if (ls != 4 && ls != 8 && ls != 16 && ls != 32 && ls != 64)
    return C(c)->fail("lanes_small must be 4, 8, 16, 32 or 64");
  if (n.lds_words < 1 || n.lds_words * 8ull * 4 > c->lds_per_block)
    return C(c)->fail("lds_words must be 1 .. LDS per workgroup / 32");
  if (n.stage_entries < 2 || n.stage_entries % 2 || n.stage_entries > 64)
    return C(c)->fail("stage_entries must be even, 2 .. 64");
  if (n.rows_per_group < 1) return C(c)->fail("rows_per_group must be >= 1");
  if (n.lane_slice_max > (1u << 20)) return C(c)->fail("lane_slice_max must be 0 .. 2^20");
  c->tuning = n;
  return DV_OK;
}

int dv_profile(dv_context *ctx, int enable) {
  if (!ctx) return DV_ERR_INVALID;
  if (enable && !ctx->pev[0]) {
    for (auto &e : ctx->pev)
      if (hipEventCreate(&e) != hipSuccess) {
        (void)hipGetLastError();
        return C(ctx)->fail("hipEventCreate failed", DV_ERR_DEVICE);
      }
  }
  if (enable)
    for (auto &t : ctx->prof_t) t = dv_kernel_time{};
  ctx->prof = enable != 0;
  return DV_OK;
}

int dv_profile_read(const dv_context *ctx, dv_kernel_time out[DV_KERNEL_KINDS]) {
  if (!ctx || !out) return DV_ERR_INVALID;
  for (int k = 0; k < DV_KERNEL_KINDS; ++k) out[k] = ctx->prof_t[k];
  return DV_OK;
}

int dv_free(dv_context *ctx, dv_rcrs *m) {
  if (!m) return DV_OK;
  auto *o = static_cast<dv::Owner *>(m->owner);
  if (o) {
    if (o->location == DV_HOST) {
      std::free(o->rowmap);
      std::free(o->entries);
    } else {
      if (!ctx) return DV_ERR_INVALID;
      (void)hipFreeAsync(o->rowmap, ctx->stream);
      (void)hipFreeAsync(o->entries, ctx->stream);
    }
    delete o;
  }
  std::memset(m, 0, sizeof(*m));
  return DV_OK;
}

int dv_copy(dv_context *ctx, const dv_rcrs *src, uint32_t location, dv_rcrs *dst) {
  if (!ctx || !dst) return DV_ERR_INVALID;
  Ctx *c = C(ctx);
  int rc = check_desc(c, src, "src");
  if (rc) return rc;
  if (location != DV_HOST && location != DV_DEVICE) return c->fail("bad location");
  DevMat m;
  rc = to_dev(c, src, m, nullptr);
  if (rc) return rc;
  m.nnz = src->nnz;
  return export_mat(c, m, location, dst, nullptr);
}

int dv_range_identity(dv_context *ctx, uint64_t n, uint32_t sb, uint32_t loc, dv_rcrs *out) {
  if (!ctx || !out) return DV_ERR_INVALID;
  Ctx *c = C(ctx);
  if (!valid_sb(sb)) return c->fail("size_bytes must be 4 or 8");
  if (n > 0xFFFFFFFEull) return c->fail("n exceeds uint32_t max - 1", DV_ERR_OVERFLOW);
  if (sb == 4 && 2 * n > 0xFFFFFFFFull) return c->fail("entry count exceeds a 32-bit rowmap", DV_ERR_OVERFLOW);
  DevMat m;
  int rc = c->identity_dev(n, sb, m);
  if (rc) return rc;
  return export_mat(c, m, loc, out, nullptr);
}

int dv_range_spadd(dv_context *ctx, const dv_rcrs *a, const dv_rcrs *b, uint32_t loc,
                   dv_rcrs *out) {
  if (!ctx || !out) return DV_ERR_INVALID;
  Ctx *c = C(ctx);
  int rc;
  if ((rc = check_desc(c, a, "a")) || (rc = check_desc(c, b, "b"))) return rc;
  if (a->nrows != b->nrows || a->ncols != b->ncols) return c->fail("spadd: shape mismatch");
  if (a->size_bytes != b->size_bytes) return c->fail("spadd: size types differ");
  DevMat da, db, dc;
  if ((rc = to_dev(c, a, da, nullptr))) return rc;
  if ((rc = to_dev(c, b, db, nullptr))) {
    c->release(da);
    return rc;
  }
  c->reset_stats();
  rc = c->spadd_dev(da, db, dc);
  c->release(da);
  c->release(db);
  if (rc) return rc;
  return export_mat(c, dc, loc, out, nullptr);
}

int dv_range_spgemm(dv_context *ctx, const dv_rcrs *a, const dv_rcrs *b, uint32_t loc,
                    dv_rcrs *out) {
  if (!ctx || !out) return DV_ERR_INVALID;
  Ctx *c = C(ctx);
  int rc;
  if ((rc = check_desc(c, a, "a")) || (rc = check_desc(c, b, "b"))) return rc;
  if (a->ncols != b->nrows) return c->fail("spgemm: a.ncols != b.nrows");
  if (a->size_bytes != b->size_bytes) return c->fail("spgemm: size types differ");
  DevMat da, db, dc;
  if ((rc = to_dev(c, a, da, nullptr))) return rc;
  if ((rc = to_dev(c, b, db, nullptr))) {
    c->release(da);
    return rc;
  }
  c->reset_stats();
  rc = c->spgemm_dev(da, db, dc);
  c->release(da);
  c->release(db);
  if (rc) return rc;
  return export_mat(c, dc, loc, out, nullptr);
}

int dv_range_power(dv_context *ctx, const dv_rcrs *a, uint32_t k, uint32_t loc, dv_rcrs *out) {
  if (!ctx || !out) return DV_ERR_INVALID;
  Ctx *c = C(ctx);
  int rc;
  if ((rc = check_desc(c, a, "a"))) return rc;
  if (a->nrows != a->ncols) return c->fail("power: matrix is not square");
  DevMat da, dc;
  if ((rc = to_dev(c, a, da, nullptr))) return rc;
  c->reset_stats();
  rc = c->power_dev(da, k, dc);
  c->release(da);
  if (rc) return rc;
  return export_mat(c, dc, loc, out, nullptr);
}

int dv_create_distance_index(dv_context *ctx, const dv_rcrs *a, uint32_t dlo, uint32_t dup,
                             uint32_t loc, dv_rcrs *out, dv_build_stats *stats) {
  if (!ctx || !out) return DV_ERR_INVALID;
  Ctx *c = C(ctx);
  int rc;
  if ((rc = check_desc(c, a, "a"))) return rc;
  if (!(dlo <= dup && dup != 0)) return c->fail("create_distance_index: need dlo <= dup and dup != 0");
  if (a->nrows != a->ncols) return c->fail("create_distance_index: matrix is not square");
  dv_build_stats st{};
  DevMat da, dd;
  if ((rc = to_dev(c, a, da, &st.h2d_ms))) return rc;
  c->reset_stats();
  const uint64_t base_bytes = c->cur_bytes;
  c->peak_bytes = c->cur_bytes;
  (void)hipEventRecord(c->ev[2], c->stream);
  rc = c->dindex_dev(da, dlo, dup, dd);
  (void)hipEventRecord(c->ev[3], c->stream);
  c->release(da);
  if (rc) return rc;
  (void)hipEventSynchronize(c->ev[3]);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, c->ev[2], c->ev[3]);
  st.build_ms = ms;
  st.alg_bytes = c->stats_alg_bytes;
  st.n_spgemm = c->stats_n_spgemm;
  st.n_spadd = c->stats_n_spadd;
  st.overflow_rows = c->stats_overflow_rows;
  st.peak_device_bytes = c->peak_bytes - std::min(base_bytes, c->peak_bytes);
  rc = export_mat(c, dd, loc, out, &st.d2h_ms);
  if (stats) *stats = st;
  return rc;
}

int dv_query_batch(dv_context *ctx, const dv_rcrs *m, const uint32_t *qi, const uint32_t *qj,
                   uint8_t *out, uint64_t n, uint32_t qloc) {
  if (!ctx) return DV_ERR_INVALID;
  Ctx *c = C(ctx);
  int rc;
  if ((rc = check_desc(c, m, "m"))) return rc;
  if (n == 0) return DV_OK;
  if (!qi || !qj || !out) return c->fail("query: NULL buffer");
  DevMat dm;
  if ((rc = to_dev(c, m, dm, nullptr))) return rc;
  const uint32_t *di = qi, *dj = qj;
  uint8_t *dout = out;
  void *bi = nullptr, *bj = nullptr, *bo = nullptr;
  if (qloc == DV_HOST) {
    bi = c->dalloc(4 * n);
    bj = c->dalloc(4 * n);
    bo = c->dalloc(n);
    if (!bi || !bj || !bo) {
      c->dfree(bi);
      c->dfree(bj);
      c->dfree(bo);
      c->release(dm);
      return c->fail("device allocation failed (query)", DV_ERR_NOMEM);
    }
    (void)hipMemcpyAsync(bi, qi, 4 * n, hipMemcpyHostToDevice, c->stream);
    (void)hipMemcpyAsync(bj, qj, 4 * n, hipMemcpyHostToDevice, c->stream);
    di = static_cast<const uint32_t *>(bi);
    dj = static_cast<const uint32_t *>(bj);
    dout = static_cast<uint8_t *>(bo);
  }
  dv::fill_u32(c->stream, c->d_flag, 1, 0u);
  dv::query(c->stream, dm.sb, dv::view(dm), di, dj, dout, n, c->d_flag);
  if (qloc == DV_HOST) (void)hipMemcpyAsync(out, bo, n, hipMemcpyDeviceToHost, c->stream);
  (void)hipMemcpyAsync(c->h_flag, c->d_flag, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream);
  bool ok = c->hip(hipStreamSynchronize(c->stream), "query synchronize");
  c->dfree(bi);
  c->dfree(bj);
  c->dfree(bo);
  c->release(dm);
  if (!ok) return DV_ERR_DEVICE;
  if (*c->h_flag) return c->fail("query: (i, j) outside the matrix", DV_ERR_RANGE);
  return DV_OK;
}

}  // extern "C"
