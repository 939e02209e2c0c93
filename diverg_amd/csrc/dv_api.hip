// C-ABI implementation: device context, matrix movement, and the host
// orchestration of range SpAdd / SpGEMM / power / distance-index build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>
#include <new>
#include <string>
#include <unordered_map>

#include "diverg_amd_graph.h"
#include "dv_graph.hpp"
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

int set_error(dv_context *ctx, const std::string &msg, int code) { return ctx->fail(msg, code); }

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

bool Ctx::read_counters(OpCounters &out, uint32_t *status) {
  if (!hip(hipMemcpyAsync(h_ctr, d_ctr, sizeof(OpCounters), hipMemcpyDeviceToHost, stream),
           "counter readback"))
    return false;
  if (status && !hip(hipMemcpyAsync(h_status, d_status, sizeof(uint32_t),
                                    hipMemcpyDeviceToHost, stream),
                     "status readback"))
    return false;
  if (!hip(hipStreamSynchronize(stream), "stream synchronize")) return false;
  out = *h_ctr;
  if (status) *status = *h_status;
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
    // a device matrix this library produced (owner set) is valid by
    // construction; a caller's device buffers are checked on the device
    if (m->owner) return DV_OK;
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
    case DV_ERR_IO: return "I/O error";
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
  // the scratch array of dv_range_spadd_noalloc on this device (a failure
  // here is reported by that entry point, which retries it)
  (void)dv::noalloc_setup(device);
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
            hipMalloc((void **)&c->d_flag, sizeof(uint32_t)) == hipSuccess &&
            hipHostMalloc((void **)&c->h_status, sizeof(uint32_t), hipHostMallocDefault) == hipSuccess &&
            hipMalloc((void **)&c->d_status, sizeof(uint32_t)) == hipSuccess;
  if (!ok) {
    (void)hipGetLastError();
    dv_finalize(c);
    return DV_ERR_DEVICE;
  }
  *out = c;
  return DV_OK;
}

int dv_synchronize(dv_context *c) {
  if (!c) return DV_ERR_INVALID;
  return c->hip(hipStreamSynchronize(c->stream), "synchronize") ? DV_OK : DV_ERR_DEVICE;
}

int dv_finalize(dv_context *c) {
  if (!c) return DV_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &kv : c->sizes) (void)hipFreeAsync(kv.first, c->stream);
  c->sizes.clear();
  c->bufs.clear();
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : c->pev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : c->plan_ev) (void)hipEventDestroy(e);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->slot_graph_exec) (void)hipGraphExecDestroy(c->slot_graph_exec);
  if (c->h_ctr) (void)hipHostFree(c->h_ctr);
  if (c->h_status) (void)hipHostFree(c->h_status);
  if (c->d_status) (void)hipFree(c->d_status);
  if (c->h_plan_ctr) (void)hipHostFree(c->h_plan_ctr);
  if (c->d_plan_ctr) (void)hipFree(c->d_plan_ctr);
  if (c->slot_parts) (void)hipFree(c->slot_parts);
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

int dv_last_tuning(const dv_context *c, dv_tuning *t) {
  if (!c || !t) return DV_ERR_INVALID;
  *t = c->last_tuning;
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
  if (n.plan_mode > 2) return C(c)->fail("plan_mode must be 0, 1 or 2");
  if (n.graphs > 1) return C(c)->fail("graphs must be 0 or 1");
  if (n.autotune > 1) return C(c)->fail("autotune must be 0 or 1");
  if (n.schedule > 1) return C(c)->fail("schedule must be 0 or 1");
  if (n.layout > 2) return C(c)->fail("layout must be 0, 1 or 2");
  if (n.summary_span == 1) return C(c)->fail("summary_span must be 0 or >= 2");
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

int dv_adjacency_matrix(dv_context *ctx, const dv_graph *g, uint64_t lower, uint64_t upper,
                        uint32_t sb, uint32_t loc, dv_rcrs *out) {
  if (!ctx || !out) return DV_ERR_INVALID;
  Ctx *c = C(ctx);
  if (!g) return c->fail("adjacency_matrix: graph is NULL");
  if (!valid_sb(sb)) return c->fail("adjacency_matrix: size_bytes must be 4 or 8");
  if (loc != DV_HOST && loc != DV_DEVICE) return c->fail("adjacency_matrix: bad location");
  DevMat m;
  const int rc = c->adjacency_dev(*reinterpret_cast<const dv::Graph *>(g), lower, upper, sb, m);
  if (rc) return rc;
  return export_mat(c, m, loc, out, nullptr);
}

int dv_merge_distance_index(dv_context *ctx, const dv_rcrs *d1, const dv_rcrs *d2, uint32_t loc,
                            dv_rcrs *out) {
  if (!ctx || !out) return DV_ERR_INVALID;
  if (d1 && d2 && (d1->nrows != d2->nrows || d1->ncols != d2->ncols))
    return C(ctx)->fail("merge_distance_index: the indices differ in shape");
  return dv_range_spadd(ctx, d1, d2, loc, out);
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
  // an input resident in caller memory is not in the context's accounting
  const uint64_t input_bytes = a->location == DV_DEVICE ? dv::mat_bytes(da) : 0;
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
  st.deferred_rows = c->stats_mid_rows + c->stats_big_rows;
  st.planned = c->stats_planned;
  st.retried = c->stats_retried;
  st.layout = c->stats_layout;
  st.builds = 1;
  // everything the context holds at the high-water mark of the build: the
  // resident input, cached plan buffers, scratch and the output
  st.peak_device_bytes = c->peak_bytes + input_bytes;
  rc = export_mat(c, dd, loc, out, &st.d2h_ms);
  if (stats) *stats = st;
  return rc;
}

int dv_create_distance_index_batched(dv_context *ctx, const dv_rcrs *comps, uint32_t ncomp,
                                     uint32_t dlo, uint32_t dup, uint32_t loc, dv_rcrs *outs,
                                     dv_build_stats *stats) {
  if (!ctx || (ncomp && (!comps || !outs))) return DV_ERR_INVALID;
  Ctx *c = C(ctx);
  int rc;
  if (!(dlo <= dup && dup != 0)) return c->fail("create_distance_index: need dlo <= dup and dup != 0");
  for (uint32_t i = 0; i < ncomp; ++i) {
    if ((rc = check_desc(c, &comps[i], "component"))) return rc;
    if (comps[i].nrows != comps[i].ncols) return c->fail("batched: a component is not square");
  }
  dv_build_stats st{};
  std::vector<DevMat> dcomps(ncomp), douts;
  uint64_t input_bytes = 0;
  for (uint32_t i = 0; i < ncomp; ++i) {
    if ((rc = to_dev(c, &comps[i], dcomps[i], &st.h2d_ms))) {
      for (auto &m : dcomps) c->release(m);
      return rc;
    }
    if (comps[i].location == DV_DEVICE) input_bytes += dv::mat_bytes(dcomps[i]);
  }
  c->reset_stats();
  c->peak_bytes = c->cur_bytes;
  (void)hipEventRecord(c->ev[2], c->stream);
  rc = c->dindex_batched(dcomps, dlo, dup, douts);
  (void)hipEventRecord(c->ev[3], c->stream);
  for (auto &m : dcomps) c->release(m);
  if (rc) return rc;
  (void)hipEventSynchronize(c->ev[3]);
  float ms = 0;
  (void)hipEventElapsedTime(&ms, c->ev[2], c->ev[3]);
  st.build_ms = ms;
  st.alg_bytes = c->stats_alg_bytes;
  st.n_spgemm = c->stats_n_spgemm;
  st.n_spadd = c->stats_n_spadd;
  st.overflow_rows = c->stats_overflow_rows;
  st.deferred_rows = c->stats_mid_rows + c->stats_big_rows;
  st.planned = c->stats_planned;
  st.retried = c->stats_retried;
  st.layout = c->stats_layout;
  st.builds = c->stats_classes;
  st.peak_device_bytes = c->peak_bytes + input_bytes;
  for (uint32_t i = 0; i < ncomp; ++i) {
    if ((rc = export_mat(c, douts[i], loc, &outs[i], &st.d2h_ms))) {
      for (uint32_t j = 0; j < i; ++j) dv_free(ctx, &outs[j]);
      for (uint32_t j = i; j < ncomp; ++j) c->release(douts[j]);
      return rc;
    }
  }
  if (stats) *stats = st;
  return DV_OK;
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
