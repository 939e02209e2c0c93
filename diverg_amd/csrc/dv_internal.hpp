// Internal declarations shared by the library's host-side sources.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>

#include "diverg_amd.h"
#include "diverg_amd_gen.h"

namespace dv {

// Set the context's last error message; returns `code` (defined with the
// context in dv_api.hip).
int set_error(dv_context *ctx, const std::string &msg, int code);

// Allocation record behind dv_rcrs::owner for library-allocated matrices.
struct Owner {
  uint32_t location;  // dv_location
  void *rowmap;
  uint32_t *entries;
  uint64_t rowmap_bytes;
  uint64_t entries_bytes;
};

// A host matrix being assembled by library code (malloc'ed storage).
struct HostMatrix {
  uint64_t nrows = 0, ncols = 0;
  uint32_t sb = 8;
  void *rowmap = nullptr;
  uint32_t *entries = nullptr;
  uint64_t n_entries = 0;

  ~HostMatrix() {
    std::free(rowmap);
    std::free(entries);
  }
  bool allocate(uint64_t nr, uint64_t nc, uint32_t size_bytes) {
    nrows = nr;
    ncols = nc;
    sb = size_bytes;
    rowmap = std::malloc((size_t)(sb * (nr + 1)));
    return rowmap != nullptr;
  }
  void set_rowmap(uint64_t i, uint64_t v) {
    if (sb == 4)
      static_cast<uint32_t *>(rowmap)[i] = (uint32_t)v;
    else
      static_cast<uint64_t *>(rowmap)[i] = v;
  }
  bool set_entries(const uint32_t *src, uint64_t n) {
    entries = (uint32_t *)std::malloc(sizeof(uint32_t) * (n ? n : 1));
    if (!entries) return false;
    if (n) std::memcpy(entries, src, sizeof(uint32_t) * n);
    n_entries = n;
    return true;
  }
  // Hand the storage to a dv_rcrs (freed later by dv_free).
  void release_to(dv_rcrs *out) {
    Owner *o = new Owner{DV_HOST, rowmap, entries, sb * (nrows + 1),
                         4 * (n_entries ? n_entries : 1)};
    out->nrows = nrows;
    out->ncols = ncols;
    out->rowmap = rowmap;
    out->entries = entries;
    out->n_entries = n_entries;
    out->nnz = 0;
    out->size_bytes = sb;
    out->location = DV_HOST;
    out->owner = o;
    rowmap = nullptr;
    entries = nullptr;
  }
};

}  // namespace dv
