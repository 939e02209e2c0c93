// Device context and device-resident matrices of the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <unordered_map>

#include "dv_internal.hpp"
#include "rcrs_kernels.hpp"

namespace dv {

// A device-resident rCRS matrix.  `owned` matrices were allocated from the
// context's stream-ordered pool; others alias caller memory.
struct DevMat {
  uint64_t nrows = 0, ncols = 0, n_entries = 0;
  uint64_t nnz = 0;  // exact, 64-bit
  uint32_t sb = 8;
  void *rowmap = nullptr;
  uint32_t *entries = nullptr;
  bool owned = false;
};

uint64_t mat_bytes(const DevMat &m);
DView view(const DevMat &m);

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  uint32_t num_cu = 256;
  uint32_t lds_per_block = 65536;
  dv_tuning tuning{512, 8, 4, 16, 64, 4096, 0};
  std::string last_error;

  // pool accounting
  std::unordered_map<void *, uint64_t> sizes;
  uint64_t cur_bytes = 0, peak_bytes = 0;

  // scratch arena reused by consecutive operations on the stream
  void *scratch = nullptr;
  uint64_t scratch_cap = 0;
  // Carve `n` buffers of the given byte sizes (256-byte aligned) from the
  // arena, growing it when needed; nullptr on allocation failure.
  bool scratch_carve(const uint64_t *bytes, void **out, int n);

  // one-word readbacks
  OpCounters *d_ctr = nullptr, *h_ctr = nullptr;
  uint32_t *d_flag = nullptr, *h_flag = nullptr;

  // per-kernel timing (dv_profile)
  bool prof = false;
  dv_kernel_time prof_t[DV_KERNEL_KINDS] = {};
  hipEvent_t pev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  void prof_mark(int i) {
    if (prof) (void)hipEventRecord(pev[i], stream);
  }
  void prof_add(int kind, int e0, int e1, uint64_t alg, uint64_t slice) {
    if (!prof) return;
    float ms = 0;
    (void)hipEventSynchronize(pev[e1]);
    (void)hipEventElapsedTime(&ms, pev[e0], pev[e1]);
    prof_t[kind].launches += 1;
    prof_t[kind].total_ms += ms;
    prof_t[kind].alg_bytes += alg;
    prof_t[kind].slice_bytes += slice;
  }

  // per-operation statistics
  uint64_t stats_alg_bytes = 0, stats_overflow_rows = 0, stats_mid_rows = 0, stats_big_rows = 0;
  uint32_t stats_n_spgemm = 0, stats_n_spadd = 0;
  void reset_stats() {
    stats_alg_bytes = stats_overflow_rows = stats_mid_rows = stats_big_rows = 0;
    stats_n_spgemm = stats_n_spadd = 0;
  }

  bool hip(hipError_t e, const char *what);
  int fail(const std::string &msg, int code = DV_ERR_INVALID);
  void *dalloc(uint64_t bytes);
  void dfree(void *p);
  void forget(void *p);
  void release(DevMat &m);
  bool read_counters(OpCounters &out);

  int validate(const DevMat &m);
  int clone(const DevMat &src, DevMat &dst);
  int identity_dev(uint64_t n, uint32_t sb, DevMat &out);
  int spadd_dev(const DevMat &a, const DevMat &b, DevMat &c);
  GroupLaunch group_launch(uint32_t lanes, uint64_t nstage) const;
  int spgemm_dev(const DevMat &a, const DevMat &b, DevMat &c);
  int power_dev(const DevMat &a, uint32_t k, DevMat &out);
  int dindex_dev(const DevMat &a, uint32_t dlo, uint32_t dup, DevMat &d);
};

}  // namespace dv
