// On-disk form of a range-CRS matrix: the byte layout DiVerG's
// CRSMatrix<RangeDynamic,...>::serialize writes (crs_matrix.hpp:1392-1409
// for the base part, :1986-2000 for the trailing nnz; vectors as in
// utils.hpp:191-289: a uint64 element count, then the elements):
//
//   u64 E, E x u32 entries | u64 R, R x S rowmap | u64 ncols | u64 nnz
//
// with S the size type (4 or 8 bytes).  The file does not name S; dv_load
// takes it from the caller or, given 0, from the file length.  Files written
// by the reference for CRSMatrix<RangeDynamic,bool,uint32_t,S> load here and
// the other way round.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "dv_internal.hpp"

namespace {

struct File {
  std::FILE *f = nullptr;
  explicit File(std::FILE *p) : f(p) {}
  ~File() {
    if (f) std::fclose(f);
  }
};

bool put_u64(std::FILE *f, uint64_t v) { return std::fwrite(&v, 8, 1, f) == 1; }
bool get_u64(std::FILE *f, uint64_t &v) { return std::fread(&v, 8, 1, f) == 1; }

// Stream `bytes` in chunks (large indexes are several GB).
bool put_bytes(std::FILE *f, const void *p, uint64_t bytes) {
  const char *c = static_cast<const char *>(p);
  while (bytes) {
    const size_t k = (size_t)std::min<uint64_t>(bytes, 1ull << 30);
    if (std::fwrite(c, 1, k, f) != k) return false;
    c += k;
    bytes -= k;
  }
  return true;
}
bool get_bytes(std::FILE *f, void *p, uint64_t bytes) {
  char *c = static_cast<char *>(p);
  while (bytes) {
    const size_t k = (size_t)std::min<uint64_t>(bytes, 1ull << 30);
    if (std::fread(c, 1, k, f) != k) return false;
    c += k;
    bytes -= k;
  }
  return true;
}

// Record the message on the context (when there is one) and return `code`.
int io_fail(dv_context *ctx, const std::string &msg, int code) {
  return ctx ? dv::set_error(ctx, msg, code) : code;
}

}  // namespace

extern "C" {

int dv_save(dv_context *ctx, const dv_rcrs *m, const char *path) {
  if (!m || !path) return DV_ERR_INVALID;
  if (m->size_bytes != 4 && m->size_bytes != 8) return io_fail(ctx, "save: size_bytes must be 4 or 8", DV_ERR_INVALID);
  dv_rcrs host{};
  const dv_rcrs *src = m;
  if (m->location == DV_DEVICE) {
    if (!ctx) return DV_ERR_INVALID;
    const int rc = dv_copy(ctx, m, DV_HOST, &host);
    if (rc) return rc;
    src = &host;
  } else if (m->location != DV_HOST) {
    return io_fail(ctx, "save: bad location", DV_ERR_INVALID);
  }
  int rc = DV_OK;
  {
    File f(std::fopen(path, "wb"));
    if (!f.f) {
      rc = io_fail(ctx, std::string("save: cannot open ") + path, DV_ERR_IO);
    } else {
      const uint64_t rows = src->nrows + 1;
      const bool ok = put_u64(f.f, src->n_entries) &&
                      put_bytes(f.f, src->entries, 4 * src->n_entries) && put_u64(f.f, rows) &&
                      put_bytes(f.f, src->rowmap, (uint64_t)src->size_bytes * rows) &&
                      put_u64(f.f, src->ncols) && put_u64(f.f, src->nnz) && std::fflush(f.f) == 0;
      if (!ok) rc = io_fail(ctx, std::string("save: write failed: ") + path, DV_ERR_IO);
    }
  }
  if (src == &host) dv_free(ctx, &host);
  return rc;
}

int dv_load(dv_context *ctx, const char *path, uint32_t size_bytes, uint32_t location,
            dv_rcrs *out) {
  if (!path || !out) return DV_ERR_INVALID;
  if (size_bytes != 0 && size_bytes != 4 && size_bytes != 8)
    return io_fail(ctx, "load: size_bytes must be 0, 4 or 8", DV_ERR_INVALID);
  if (location != DV_HOST && location != DV_DEVICE)
    return io_fail(ctx, "load: bad location", DV_ERR_INVALID);
  if (location == DV_DEVICE && !ctx) return DV_ERR_INVALID;
  File f(std::fopen(path, "rb"));
  if (!f.f) return io_fail(ctx, std::string("load: cannot open ") + path, DV_ERR_IO);
  if (std::fseek(f.f, 0, SEEK_END) != 0) return io_fail(ctx, "load: cannot seek", DV_ERR_IO);
  const long flen = std::ftell(f.f);
  std::rewind(f.f);
  if (flen < 32) return io_fail(ctx, "load: file too short", DV_ERR_IO);
  const uint64_t len = (uint64_t)flen;
  uint64_t ne = 0, rows = 0, ncols = 0, nnz = 0;
  if (!get_u64(f.f, ne) || ne > (len - 32) / 4)
    return io_fail(ctx, "load: bad entry count", DV_ERR_IO);
  dv::HostMatrix hm;
  hm.entries = static_cast<uint32_t *>(std::malloc(4 * (ne ? ne : 1)));
  hm.n_entries = ne;
  if (!hm.entries) return io_fail(ctx, "load: out of memory", DV_ERR_NOMEM);
  if (!get_bytes(f.f, hm.entries, 4 * ne) || !get_u64(f.f, rows) || rows == 0)
    return io_fail(ctx, "load: truncated file", DV_ERR_IO);
  const uint64_t rest = len - 8 - 4 * ne - 8 - 16;  // bytes of the rowmap elements
  uint32_t sb = size_bytes;
  if (sb == 0) {
    if (rest == 4 * rows) sb = 4;
    else if (rest == 8 * rows) sb = 8;
    else return io_fail(ctx, "load: file length fits neither a 32- nor a 64-bit rowmap", DV_ERR_IO);
  } else if (rest != (uint64_t)sb * rows) {
    return io_fail(ctx, "load: file length does not match the size type", DV_ERR_IO);
  }
  if (!hm.allocate(rows - 1, 0, sb)) return io_fail(ctx, "load: out of memory", DV_ERR_NOMEM);
  if (!get_bytes(f.f, hm.rowmap, (uint64_t)sb * rows) || !get_u64(f.f, ncols) || !get_u64(f.f, nnz))
    return io_fail(ctx, "load: truncated file", DV_ERR_IO);
  // the structure must be a valid matrix before anything reads it
  uint64_t prev = 0;
  for (uint64_t i = 0; i < rows; ++i) {
    const uint64_t v = sb == 4 ? static_cast<const uint32_t *>(hm.rowmap)[i]
                               : static_cast<const uint64_t *>(hm.rowmap)[i];
    if ((i == 0 && v != 0) || v < prev || (v & 1)) return io_fail(ctx, "load: malformed rowmap", DV_ERR_IO);
    prev = v;
  }
  if (prev != ne) return io_fail(ctx, "load: rowmap does not end at the entry count", DV_ERR_IO);
  if (ncols > 0xFFFFFFFEull) return io_fail(ctx, "load: ncols exceeds UINT32_MAX - 1", DV_ERR_OVERFLOW);
  for (uint64_t k = 0; k < ne; ++k)
    if (hm.entries[k] >= ncols) return io_fail(ctx, "load: entry outside [0, ncols)", DV_ERR_IO);
  hm.ncols = ncols;
  dv_rcrs host{};
  hm.release_to(&host);
  host.nnz = nnz;
  if (location == DV_HOST) {
    *out = host;
    return DV_OK;
  }
  const int rc = dv_copy(ctx, &host, DV_DEVICE, out);
  dv_free(ctx, &host);
  return rc;
}

}  // extern "C"
