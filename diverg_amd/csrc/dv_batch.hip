// Batched distance-index build: many independent graph components in one
// planned build.
//
// The components are assembled block-diagonally into one matrix in HBM (rows
// and columns of component c shifted by the rows before it, 64-bit rowmap),
// the planned build runs once over the assembly -- each power step is one
// launch sequence that covers every component, and the sequence can be
// replayed from a hipGraph -- and the index is split back.  No edge crosses a
// component, rows of every product depend only on rows of their own
// component, and the identity, SpAdd and SpGEMM are row-wise, so each
// component's index is the corresponding diagonal block shifted by its
// offset: bit-identical to a plain build of that component (the block
// assembly the reference itself checks in test_dindex.cpp:111-173).
// The launch shape of a build is the reference's SparseConfig / ExecGrid
// choice (range_sparse_base.hpp:99-498); with dv_tuning.autotune the
// components are grouped in size classes and each class is one build with
// its own autotuned shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "dv_runtime.hpp"

namespace dv {

// Size class of a component for the per-class tuned batched build: by rows,
// below 2^20, below 2^24, and larger.
static int size_class(uint64_t nrows) { return nrows < (1ull << 20) ? 0 : nrows < (1ull << 24) ? 1 : 2; }

int Ctx::dindex_batched(const std::vector<DevMat> &comps, uint32_t dlo, uint32_t dup,
                        std::vector<DevMat> &outs) {
  if (!(dlo <= dup && dup != 0)) return fail("create_distance_index: need dlo <= dup and dup != 0");
  // A component with a non-canonical row is built on its own: the reference
  // clips such rows to an accumulator window placed by row and column index
  // (range_sparse.hpp:1192-1242), which the assembly would shift.
  std::vector<size_t> odd, even;
  for (size_t c = 0; c < comps.size(); ++c) {
    const uint8_t k = comps[c].nrows ? canonical(comps[c]) : 1;
    if (k == 0) return DV_ERR_DEVICE;
    (k == 2 ? odd : even).push_back(c);
  }
  if (!odd.empty()) {
    outs.assign(comps.size(), DevMat{});
    std::vector<DevMat> sub, sub_out;
    for (size_t c : even) sub.push_back(comps[c]);
    int rc = dindex_batched_classes(sub, dlo, dup, sub_out);
    const uint32_t n_spgemm = stats_n_spgemm, n_spadd = stats_n_spadd,
                   builds = stats_classes;
    const uint64_t alg = stats_alg_bytes;
    if (!rc)
      for (size_t k = 0; k < even.size(); ++k) outs[even[k]] = sub_out[k];
    for (size_t k = 0; !rc && k < odd.size(); ++k) rc = dindex_dev(comps[odd[k]], dlo, dup, outs[odd[k]]);
    if (rc) {
      for (DevMat &o : outs) release(o);
      return rc;
    }
    stats_n_spgemm += n_spgemm;
    stats_n_spadd += n_spadd;
    stats_alg_bytes += alg;
    stats_classes = builds + (uint32_t)odd.size();
    stats_retried = 1;
    return DV_OK;
  }
  return dindex_batched_classes(comps, dlo, dup, outs);
}

int Ctx::dindex_batched_classes(const std::vector<DevMat> &comps, uint32_t dlo, uint32_t dup,
                                std::vector<DevMat> &outs) {
  // With the autotuner on, the components are built per size class: one
  // block-diagonal build per class, each with the launch shape the
  // autotuner picks for that class's build shape (and caches).
  if (!tuning.autotune || tuning.plan_mode != 0 || comps.size() < 2)
    return dindex_batched_one(comps, dlo, dup, outs);
  std::vector<std::vector<size_t>> cls(3);
  for (size_t c = 0; c < comps.size(); ++c) cls[size_class(comps[c].nrows)].push_back(c);
  outs.assign(comps.size(), DevMat{});
  uint32_t n_spgemm = 0, n_spadd = 0, retried = 0, classes = 0;
  uint64_t alg = 0;
  for (const auto &members : cls) {
    if (members.empty()) continue;
    std::vector<DevMat> sub, sub_out;
    for (size_t c : members) sub.push_back(comps[c]);
    const int rc = dindex_batched_one(sub, dlo, dup, sub_out);
    if (rc) {
      for (DevMat &o : outs) release(o);
      return rc;
    }
    for (size_t k = 0; k < members.size(); ++k) outs[members[k]] = sub_out[k];
    n_spgemm += stats_n_spgemm;
    n_spadd += stats_n_spadd;
    alg += stats_alg_bytes;
    retried |= stats_retried;
    classes += 1;
  }
  stats_n_spgemm = n_spgemm;
  stats_n_spadd = n_spadd;
  stats_alg_bytes = alg;
  stats_retried = retried;
  stats_classes = classes;
  return DV_OK;
}

int Ctx::dindex_batched_one(const std::vector<DevMat> &comps, uint32_t dlo, uint32_t dup,
                            std::vector<DevMat> &outs) {
  const size_t m = comps.size();
  outs.assign(m, DevMat{});
  stats_classes = 1;
  if (m == 0) return DV_OK;
  std::vector<uint64_t> row0(m + 1, 0), ent0(m + 1, 0);
  for (size_t c = 0; c < m; ++c) {
    if (comps[c].nrows != comps[c].ncols) return fail("batched: a component is not square");
    row0[c + 1] = row0[c] + comps[c].nrows;
    ent0[c + 1] = ent0[c] + comps[c].n_entries;
  }
  if (row0[m] > 0xFFFFFFFEull)
    return fail("batched: the components hold more than uint32_t max - 1 loci", DV_ERR_OVERFLOW);

  // 1. block-diagonal assembly (uint64 rowmap)
  DevMat big;
  big.nrows = big.ncols = row0[m];
  big.n_entries = ent0[m];
  big.sb = 8;
  big.owned = true;
  big.rowmap = dalloc(8 * (big.nrows + 1));
  big.entries = static_cast<uint32_t *>(dalloc(4 * big.n_entries));
  if (!big.rowmap || !big.entries) {
    release(big);
    return fail("device allocation failed (batched assembly)", DV_ERR_NOMEM);
  }
  for (size_t c = 0; c < m; ++c)
    concat_component(stream, comps[c].sb, comps[c].rowmap, comps[c].entries, comps[c].nrows,
                     comps[c].n_entries, row0[c], ent0[c], (uint32_t)row0[c],
                     static_cast<uint64_t *>(big.rowmap), big.entries, c + 1 == m);
  for (const DevMat &c : comps) big.nnz += c.nnz;

  // 2. one planned build over the assembly
  DevMat d;
  int rc = dindex_dev(big, dlo, dup, d);
  const uint32_t n_spgemm = stats_n_spgemm, n_spadd = stats_n_spadd;
  const uint64_t alg = stats_alg_bytes;
  release(big);
  if (rc) return rc;

  // 3. split: component entry offsets and nnz, then one extract per component
  uint64_t *bounds = static_cast<uint64_t *>(dalloc(8 * (m + 1)));
  auto *eoff = static_cast<unsigned long long *>(dalloc(8 * (m + 1)));
  auto *nnz = static_cast<unsigned long long *>(dalloc(8 * m));
  std::vector<unsigned long long> h_eoff(m + 1), h_nnz(m);
  bool ok = bounds && eoff && nnz &&
            hip(hipMemcpyAsync(bounds, row0.data(), 8 * (m + 1), hipMemcpyHostToDevice, stream),
                "bounds upload");
  if (ok) {
    gather_bounds(stream, static_cast<const uint64_t *>(d.rowmap), bounds, (uint32_t)(m + 1), eoff);
    fill_u32(stream, reinterpret_cast<uint32_t *>(nnz), 2 * m, 0u);
    ok = hip(hipMemcpyAsync(h_eoff.data(), eoff, 8 * (m + 1), hipMemcpyDeviceToHost, stream),
             "bounds readback") &&
         hip(hipStreamSynchronize(stream), "bounds synchronize");
  }
  if (ok) {
    for (size_t c = 0; c < m; ++c) {
      DevMat &o = outs[c];
      o.nrows = o.ncols = comps[c].nrows;
      o.sb = comps[c].sb;
      o.n_entries = h_eoff[c + 1] - h_eoff[c];
      o.owned = true;
      if (o.sb == 4 && o.n_entries > 0xFFFFFFFFull) {
        ok = false;
        rc = fail("batched: a component's index exceeds a 32-bit rowmap", DV_ERR_OVERFLOW);
        break;
      }
      o.rowmap = dalloc((uint64_t)o.sb * (o.nrows + 1));
      o.entries = static_cast<uint32_t *>(dalloc(4 * o.n_entries));
      if (!o.rowmap || !o.entries) {
        ok = false;
        rc = fail("device allocation failed (batched outputs)", DV_ERR_NOMEM);
        break;
      }
      extract_component(stream, o.sb, static_cast<const uint64_t *>(d.rowmap), d.entries, row0[c],
                        o.nrows, o.n_entries, (uint32_t)row0[c], o.rowmap, o.entries);
      slice_nnz(stream, o.entries, o.n_entries, nnz + c);
    }
  }
  if (ok)
    ok = hip(hipMemcpyAsync(h_nnz.data(), nnz, 8 * m, hipMemcpyDeviceToHost, stream),
             "nnz readback") &&
         hip(hipStreamSynchronize(stream), "split synchronize");
  dfree(bounds);
  dfree(eoff);
  dfree(nnz);
  release(d);
  if (!ok) {
    for (DevMat &o : outs) release(o);
    return rc ? rc : DV_ERR_DEVICE;
  }
  for (size_t c = 0; c < m; ++c) outs[c].nnz = h_nnz[c];
  stats_n_spgemm = n_spgemm;
  stats_n_spadd = n_spadd;
  stats_alg_bytes = alg;
  return DV_OK;
}

}  // namespace dv
