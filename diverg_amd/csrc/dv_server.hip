// Serving loop: continuous batching of locus-pair queries over HBM-resident
// indexes, with index builds admitted by HBM budget on a second stream.
//
// Threads: the caller's threads submit; one loop thread packs whatever queries
// are queued into the next batch whenever a batch slot is free (so queries
// that arrive while earlier batches are on the GPU ride the next launch) and
// completes batches as their events fire; one build thread admits and runs
// builds.  Queries run on their own stream (their own hardware queue), builds
// on another.  All streams keep the default priority.
//
// Queries are queued as segments (a request's queries, or the replayed
// queries that fell due together), so packing a batch and handing answers
// back are contiguous copies, not work per query.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "dv_runtime.hpp"

namespace {

using steady = std::chrono::steady_clock;

double now_us() {
  return std::chrono::duration<double, std::micro>(steady::now().time_since_epoch()).count();
}

struct Request {
  uint32_t kind = 0;  // 0 queries, 1 build
  std::atomic<int32_t> status{DV_ERR_BUSY};
  uint32_t index_id = 0;
  uint64_t n = 0;
  double submit_us = 0, start_us = 0, done_us = 0;
  std::vector<uint8_t> answers;
  std::vector<uint32_t> qi, qj;  // the request's pairs (a private copy)
  std::atomic<uint64_t> remaining{0};
  // build
  std::vector<uint8_t> rowmap;
  std::vector<uint32_t> entries;
  dv_rcrs a{};
  uint32_t dlo = 0, dup = 0;
  uint64_t est_bytes = 0;
};

// Queries [first, first + count) of a request or of a replayed stream.
struct Seg {
  Request *req;              // ticketed request, or nullptr for replayed queries
  uint32_t idx;              // the index of every query (requests)
  const uint32_t *qidx;      // replay: per-query index ids
  const uint32_t *qi, *qj;   // the pairs (request copies or the replay arrays)
  uint64_t first, count;
  uint8_t *ans;              // replay: answers array
  double *lat;               // replay: latency array
  const double *arrival;     // replay: arrival times, microseconds after t0_us
  double t0_us;
};

// One batch in flight: pinned host staging (index ids, rows, columns:
// max_batch words each; answers) and their device copies.
struct Slot {
  hipEvent_t ev = nullptr;
  uint32_t *h_in = nullptr;
  uint8_t *h_out = nullptr;
  uint32_t *d_in = nullptr;
  uint8_t *d_out = nullptr;
  std::vector<std::pair<Seg, uint64_t>> items;  // pieces and their batch offsets
  uint64_t n = 0;
  bool busy = false;
};

struct Index {
  dv_rcrs m{};
  dv_context *owner = nullptr;  // context whose dv_free releases it
  uint64_t bytes = 0;
};

constexpr uint32_t TABLE_CAP = 4096;

}  // namespace

struct dv_server {
  int device = 0;
  dv_server_opts opts{};
  dv_context *qctx = nullptr;  // query stream owner
  dv_context *bctx = nullptr;  // builds
  hipStream_t qstream = nullptr;
  // serialises the callers' copies through qctx (its allocation map and
  // counters are not thread-safe); the loop thread enqueues on qstream too,
  // which HIP allows from any thread
  std::mutex copy_mu;

  // index table
  std::mutex idx_mu;
  std::vector<Index> idxs;
  dv::IndexDesc *h_table = nullptr, *d_table = nullptr;
  std::atomic<uint32_t> ntable{0};
  std::atomic<uint64_t> index_bytes{0};

  // queries
  std::mutex q_mu;
  std::deque<Seg> pending;
  std::vector<Slot> slots;
  std::atomic<uint64_t> replay_done{0};

  // requests
  std::mutex r_mu;
  std::condition_variable r_cv;
  std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
  uint64_t next_ticket = 1;

  // builds
  std::mutex b_mu;
  std::condition_variable b_cv;
  std::deque<Request *> b_pending;

  std::atomic<bool> stopping{false};
  std::thread loop_thread, build_thread;

  void finish(Request *r, int32_t status) {
    {
      std::lock_guard<std::mutex> lk(r_mu);
      r->done_us = now_us();
      r->status.store(status);
    }
    r_cv.notify_all();
  }

  // register a device-resident index (takes ownership of m)
  int publish(const dv_rcrs &m, dv_context *owner, uint32_t *id) {
    std::lock_guard<std::mutex> lk(idx_mu);
    const uint32_t k = (uint32_t)idxs.size();
    if (k >= TABLE_CAP) return DV_ERR_BUDGET;
    Index ix;
    ix.m = m;
    ix.owner = owner;
    ix.bytes = (uint64_t)m.size_bytes * (m.nrows + 1) + 4 * m.n_entries;
    idxs.push_back(ix);
    index_bytes += ix.bytes;
    h_table[k] = dv::IndexDesc{m.rowmap, m.entries, m.nrows, m.ncols, m.size_bytes, 0};
    // the slot of h_table is written once, so the asynchronous copy reads a
    // stable source; batches launched after ntable is published follow it
    if (hipMemcpyAsync(d_table + k, h_table + k, sizeof(dv::IndexDesc), hipMemcpyHostToDevice,
                       qstream) != hipSuccess ||
        hipStreamSynchronize(qstream) != hipSuccess) {
      (void)hipGetLastError();
      return DV_ERR_DEVICE;
    }
    ntable.store(k + 1, std::memory_order_release);
    *id = k;
    return DV_OK;
  }

  void complete_slot(Slot &s) {
    const double t = now_us();
    uint64_t replayed = 0;
    for (const auto &piece : s.items) {
      const Seg &g = piece.first;
      const uint8_t *src = s.h_out + piece.second;
      if (g.req) {
        std::memcpy(g.req->answers.data() + g.first, src, g.count);
        if (g.req->remaining.fetch_sub(g.count) == g.count) finish(g.req, DV_OK);
      } else {
        std::memcpy(g.ans + g.first, src, g.count);
        for (uint64_t q = 0; q < g.count; ++q)
          g.lat[g.first + q] = t - (g.t0_us + g.arrival[g.first + q]);
        replayed += g.count;
      }
    }
    if (replayed) replay_done.fetch_add(replayed, std::memory_order_release);
    s.items.clear();
    s.n = 0;
    s.busy = false;
  }

  bool launch_slot(Slot &s) {
    const uint64_t cap = opts.max_batch;
    {
      std::lock_guard<std::mutex> lk(q_mu);
      if (pending.empty()) return false;
      uint64_t n = 0;
      while (n < cap && !pending.empty()) {
        Seg &g = pending.front();
        const uint64_t take = std::min<uint64_t>(g.count, cap - n);
        Seg piece = g;
        piece.count = take;
        s.items.emplace_back(piece, n);
        n += take;
        if (take == g.count) {
          pending.pop_front();
        } else {
          g.first += take;
          g.count -= take;
        }
      }
      s.n = n;
    }
    const double t = now_us();
    uint32_t *h_idx = s.h_in, *h_i = s.h_in + cap, *h_j = s.h_in + 2 * cap;
    for (const auto &piece : s.items) {
      const Seg &g = piece.first;
      const uint64_t o = piece.second;
      if (g.qidx) std::memcpy(h_idx + o, g.qidx + g.first, 4 * g.count);
      else std::fill(h_idx + o, h_idx + o + g.count, g.idx);
      std::memcpy(h_i + o, g.qi + g.first, 4 * g.count);
      std::memcpy(h_j + o, g.qj + g.first, 4 * g.count);
      if (g.req && g.req->start_us == 0) g.req->start_us = t;
    }
    const uint32_t n = (uint32_t)s.n;
    const uint32_t nt = ntable.load(std::memory_order_acquire);
    // one copy when the batch is full, else the three used prefixes
    if (s.n == cap) {
      (void)hipMemcpyAsync(s.d_in, s.h_in, 12ull * cap, hipMemcpyHostToDevice, qstream);
    } else {
      for (int k = 0; k < 3; ++k)
        (void)hipMemcpyAsync(s.d_in + k * cap, s.h_in + k * cap, 4ull * n, hipMemcpyHostToDevice,
                             qstream);
    }
    dv::query_multi(qstream, d_table, nt, s.d_in, s.d_in + cap, s.d_in + 2 * cap, n, s.d_out);
    (void)hipMemcpyAsync(s.h_out, s.d_out, n, hipMemcpyDeviceToHost, qstream);
    (void)hipEventRecord(s.ev, qstream);
    s.busy = true;
    return true;
  }

  void loop() {
    for (;;) {
      bool did = false, any_busy = false;
      for (Slot &s : slots)
        if (s.busy) {
          const hipError_t e = hipEventQuery(s.ev);
          if (e == hipSuccess) {
            complete_slot(s);
            did = true;
          } else if (e == hipErrorNotReady) {
            any_busy = true;
          } else {  // the batch failed: answer 2 (no answer) and move on
            (void)hipGetLastError();
            std::memset(s.h_out, 2, s.n);
            complete_slot(s);
            did = true;
          }
        }
      for (Slot &s : slots)
        if (!s.busy && launch_slot(s)) did = true;
      if (!did) {
        if (stopping.load() && !any_busy) {
          std::lock_guard<std::mutex> lk(q_mu);
          if (pending.empty()) break;
        }
        std::this_thread::yield();
      }
    }
  }

  uint64_t budget_left() {
    const uint64_t held = index_bytes.load() + bctx->cur_bytes;
    return held >= opts.hbm_budget_bytes ? 0 : opts.hbm_budget_bytes - held;
  }

  void builds() {
    for (;;) {
      Request *r = nullptr;
      {
        std::unique_lock<std::mutex> lk(b_mu);
        b_cv.wait(lk, [&] { return stopping.load() || !b_pending.empty(); });
        if (b_pending.empty()) return;
        r = b_pending.front();
        b_pending.pop_front();
      }
      // admission: release cached build memory if the head does not fit
      if (r->est_bytes > budget_left()) {
        bctx->buf_trim();
        bctx->dfree(bctx->scratch);
        bctx->scratch = nullptr;
        bctx->scratch_cap = 0;
        (void)hipStreamSynchronize(bctx->stream);
      }
      if (r->est_bytes > budget_left()) {
        finish(r, DV_ERR_BUDGET);
        continue;
      }
      r->start_us = now_us();
      dv_rcrs out{};
      int rc = dv_create_distance_index(bctx, &r->a, r->dlo, r->dup, DV_DEVICE, &out, nullptr);
      r->rowmap.clear();
      r->rowmap.shrink_to_fit();
      r->entries.clear();
      r->entries.shrink_to_fit();
      if (rc == DV_OK) rc = publish(out, bctx, &r->index_id);
      finish(r, rc);
    }
  }

  uint64_t add_request(std::unique_ptr<Request> r) {
    std::lock_guard<std::mutex> lk(r_mu);
    const uint64_t t = next_ticket++;
    reqs[t] = std::move(r);
    return t;
  }

  Request *find(uint64_t t) {
    std::lock_guard<std::mutex> lk(r_mu);
    auto it = reqs.find(t);
    return it == reqs.end() ? nullptr : it->second.get();
  }

  // copy: take a private copy of `a` (the caller's buffers may go away when
  // the call returns); without it the caller keeps them alive until the
  // build has finished (the replay load generator does)
  int submit_build(const dv_rcrs *a, uint32_t dlo, uint32_t dup, uint64_t *ticket,
                   bool copy = true) {
    if (!a || !ticket || !(dlo <= dup && dup != 0) || a->nrows != a->ncols ||
        (a->size_bytes != 4 && a->size_bytes != 8))
      return DV_ERR_INVALID;
    auto r = std::make_unique<Request>();
    r->kind = 1;
    r->dlo = dlo;
    r->dup = dup;
    r->submit_us = now_us();
    r->est_bytes = 256ull * (a->nrows + 1) + (uint64_t)a->size_bytes * (a->nrows + 1) +
                   4 * a->n_entries;
    if (r->est_bytes > opts.hbm_budget_bytes) return DV_ERR_BUDGET;
    r->a = *a;
    r->a.owner = nullptr;
    if (copy) {
      // private host copy, so the caller's buffers may go away
      const uint64_t rmb = (uint64_t)a->size_bytes * (a->nrows + 1);
      r->rowmap.resize(rmb);
      r->entries.resize(a->n_entries);
      const hipMemcpyKind kind = a->location == DV_DEVICE ? hipMemcpyDeviceToHost
                                                          : hipMemcpyHostToHost;
      if (hipMemcpy(r->rowmap.data(), a->rowmap, rmb, kind) != hipSuccess ||
          (a->n_entries && hipMemcpy(r->entries.data(), a->entries, 4 * a->n_entries, kind) !=
                               hipSuccess)) {
        (void)hipGetLastError();
        return DV_ERR_DEVICE;
      }
      r->a.rowmap = r->rowmap.data();
      r->a.entries = r->entries.data();
      r->a.location = DV_HOST;
    }
    Request *raw = r.get();
    *ticket = add_request(std::move(r));
    {
      std::lock_guard<std::mutex> lk(b_mu);
      b_pending.push_back(raw);
    }
    b_cv.notify_all();
    return DV_OK;
  }
};

extern "C" {

int dv_server_start(int device, const dv_server_opts *opts, dv_server **out) {
  if (!out) return DV_ERR_INVALID;
  *out = nullptr;
  // CU-masked query streams are switched off (query_cus must be 0)
  if (opts && opts->query_cus) return DV_ERR_INVALID;
  auto *s = new (std::nothrow) dv_server();
  if (!s) return DV_ERR_NOMEM;
  if (opts) s->opts = *opts;
  if (!s->opts.max_batch) s->opts.max_batch = 65536;
  if (!s->opts.max_inflight) s->opts.max_inflight = 2;
  s->opts.max_inflight = std::min<uint32_t>(s->opts.max_inflight, 8);
  s->device = device;
  int rc = dv_init(device, &s->qctx);
  if (!rc) rc = dv_init(device, &s->bctx);
  if (rc) {
    if (s->qctx) dv_finalize(s->qctx);
    delete s;
    return rc;
  }
  s->qstream = s->qctx->stream;
  size_t free_b = 0, total_b = 0;
  (void)hipSetDevice(device);
  (void)hipMemGetInfo(&free_b, &total_b);
  if (!s->opts.hbm_budget_bytes) s->opts.hbm_budget_bytes = free_b / 4 * 3;
  bool ok = hipHostMalloc((void **)&s->h_table, TABLE_CAP * sizeof(dv::IndexDesc),
                          hipHostMallocDefault) == hipSuccess &&
            hipMalloc((void **)&s->d_table, TABLE_CAP * sizeof(dv::IndexDesc)) == hipSuccess;
  s->slots.resize(s->opts.max_inflight);
  for (Slot &sl : s->slots) {
    ok = ok && hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) == hipSuccess &&
         hipHostMalloc((void **)&sl.h_in, 12ull * s->opts.max_batch, hipHostMallocDefault) ==
             hipSuccess &&
         hipHostMalloc((void **)&sl.h_out, s->opts.max_batch, hipHostMallocDefault) ==
             hipSuccess &&
         hipMalloc((void **)&sl.d_in, 12ull * s->opts.max_batch) == hipSuccess &&
         hipMalloc((void **)&sl.d_out, s->opts.max_batch) == hipSuccess;
  }
  if (!ok) {
    (void)hipGetLastError();
    s->stopping = true;
    dv_server_stop(s);
    return DV_ERR_DEVICE;
  }
  s->loop_thread = std::thread([s] { s->loop(); });
  s->build_thread = std::thread([s] { s->builds(); });
  *out = s;
  return DV_OK;
}

int dv_server_stop(dv_server *s) {
  if (!s) return DV_OK;
  s->stopping = true;
  s->b_cv.notify_all();
  if (s->loop_thread.joinable()) s->loop_thread.join();
  if (s->build_thread.joinable()) s->build_thread.join();
  if (s->qstream) (void)hipStreamSynchronize(s->qstream);
  for (Index &ix : s->idxs) dv_free(ix.owner, &ix.m);
  for (Slot &sl : s->slots) {
    if (sl.ev) (void)hipEventDestroy(sl.ev);
    if (sl.h_in) (void)hipHostFree(sl.h_in);
    if (sl.h_out) (void)hipHostFree(sl.h_out);
    if (sl.d_in) (void)hipFree(sl.d_in);
    if (sl.d_out) (void)hipFree(sl.d_out);
  }
  if (s->h_table) (void)hipHostFree(s->h_table);
  if (s->d_table) (void)hipFree(s->d_table);
  if (s->qctx && s->qstream && s->qstream != s->qctx->stream) (void)hipStreamDestroy(s->qstream);
  if (s->bctx) dv_finalize(s->bctx);
  if (s->qctx) dv_finalize(s->qctx);
  (void)hipGetLastError();
  delete s;
  return DV_OK;
}

int dv_server_add_index(dv_server *s, const dv_rcrs *index, uint32_t *id) {
  if (!s || !index || !id) return DV_ERR_INVALID;
  dv_rcrs dev{};
  int rc;
  {
    std::lock_guard<std::mutex> lk(s->copy_mu);
    const uint64_t bytes = (uint64_t)index->size_bytes * (index->nrows + 1) + 4 * index->n_entries;
    if (s->index_bytes.load() + bytes > s->opts.hbm_budget_bytes) return DV_ERR_BUDGET;
    rc = dv_copy(s->qctx, index, DV_DEVICE, &dev);
  }
  if (rc) return rc;
  return s->publish(dev, s->qctx, id);
}

int dv_server_submit_queries(dv_server *s, uint32_t index_id, const uint32_t *qi,
                             const uint32_t *qj, uint64_t n, uint64_t *ticket) {
  if (!s || !ticket || (n && (!qi || !qj))) return DV_ERR_INVALID;
  if (index_id >= s->ntable.load()) return DV_ERR_NOT_FOUND;
  auto r = std::make_unique<Request>();
  r->kind = 0;
  r->index_id = index_id;
  r->n = n;
  r->submit_us = now_us();
  r->answers.assign(n, 0);
  r->qi.assign(qi, qi + n);
  r->qj.assign(qj, qj + n);
  r->remaining = n;
  Request *raw = r.get();
  *ticket = s->add_request(std::move(r));
  if (n == 0) {
    s->finish(raw, DV_OK);
    return DV_OK;
  }
  std::lock_guard<std::mutex> lk(s->q_mu);
  s->pending.push_back(Seg{raw, index_id, nullptr, raw->qi.data(), raw->qj.data(), 0, n, nullptr,
                           nullptr, nullptr, 0});
  return DV_OK;
}

int dv_server_submit_build(dv_server *s, const dv_rcrs *a, uint32_t dlo, uint32_t dup,
                           uint64_t *ticket) {
  if (!s) return DV_ERR_INVALID;
  return s->submit_build(a, dlo, dup, ticket);
}

static void fill_info(const Request *r, dv_request_info *info) {
  if (!info) return;
  info->kind = r->kind;
  info->status = r->status.load();
  info->index_id = r->index_id;
  info->n = r->n;
  info->submit_us = r->submit_us;
  info->start_us = r->start_us;
  info->done_us = r->done_us;
}

int dv_server_poll(dv_server *s, uint64_t ticket, dv_request_info *info) {
  if (!s) return DV_ERR_INVALID;
  Request *r = s->find(ticket);
  if (!r) return DV_ERR_NOT_FOUND;
  std::lock_guard<std::mutex> lk(s->r_mu);
  fill_info(r, info);
  return r->status.load() == DV_ERR_BUSY ? DV_ERR_BUSY : DV_OK;
}

int dv_server_wait(dv_server *s, uint64_t ticket, dv_request_info *info) {
  if (!s) return DV_ERR_INVALID;
  Request *r = s->find(ticket);
  if (!r) return DV_ERR_NOT_FOUND;
  std::unique_lock<std::mutex> lk(s->r_mu);
  s->r_cv.wait(lk, [&] { return r->status.load() != DV_ERR_BUSY; });
  fill_info(r, info);
  return DV_OK;
}

int dv_server_results(dv_server *s, uint64_t ticket, uint8_t *out, uint64_t n) {
  if (!s || (n && !out)) return DV_ERR_INVALID;
  Request *r = s->find(ticket);
  if (!r || r->kind != 0) return DV_ERR_NOT_FOUND;
  if (r->status.load() == DV_ERR_BUSY) return DV_ERR_BUSY;
  if (n > r->answers.size()) return DV_ERR_INVALID;
  std::memcpy(out, r->answers.data(), n);
  return r->status.load();
}

int dv_server_get_index(dv_server *s, uint32_t id, dv_rcrs *out) {
  if (!s || !out) return DV_ERR_INVALID;
  dv_rcrs m{};
  {
    std::lock_guard<std::mutex> lk(s->idx_mu);
    if (id >= s->idxs.size()) return DV_ERR_NOT_FOUND;
    m = s->idxs[id].m;
  }
  m.owner = nullptr;  // a view: dv_copy makes the host copy
  std::lock_guard<std::mutex> lk(s->copy_mu);
  return dv_copy(s->qctx, &m, DV_HOST, out);
}

int dv_server_replay(dv_server *s, uint64_t n, const uint32_t *qidx, const uint32_t *qi,
                     const uint32_t *qj, const double *arrival_us, uint8_t *answers,
                     double *latency_us, uint32_t nbuilds, const dv_rcrs *builds,
                     const uint32_t *build_dlo, const uint32_t *build_dup,
                     const double *build_us, double *build_done_us, uint32_t *build_index,
                     double *elapsed_us) {
  if (!s || (n && (!qidx || !qi || !qj || !arrival_us || !answers || !latency_us)) ||
      (nbuilds && (!builds || !build_dlo || !build_dup || !build_us || !build_done_us ||
                   !build_index)))
    return DV_ERR_INVALID;
  for (uint64_t q = 1; q < n; ++q)
    if (arrival_us[q] < arrival_us[q - 1]) return DV_ERR_INVALID;
  std::vector<uint64_t> tickets(nbuilds, 0);
  std::vector<int> build_rc(nbuilds, DV_OK);
  const uint64_t done0 = s->replay_done.load();
  const double t0 = now_us();
  uint64_t qpos = 0;
  uint32_t bpos = 0;
  // builds in order of their arrival times
  std::vector<uint32_t> border(nbuilds);
  for (uint32_t b = 0; b < nbuilds; ++b) border[b] = b;
  std::sort(border.begin(), border.end(),
            [&](uint32_t x, uint32_t y) { return build_us[x] < build_us[y]; });
  while (qpos < n || bpos < nbuilds) {
    const double t = now_us() - t0;
    while (bpos < nbuilds && build_us[border[bpos]] <= t) {
      const uint32_t b = border[bpos++];
      // the caller's build inputs outlive the replay: no copy on this thread
      build_rc[b] = s->submit_build(&builds[b], build_dlo[b], build_dup[b], &tickets[b], false);
    }
    if (qpos < n && arrival_us[qpos] <= t) {
      // every query due by now, as one segment
      const uint64_t end =
          (uint64_t)(std::upper_bound(arrival_us + qpos, arrival_us + n, t) - arrival_us);
      std::lock_guard<std::mutex> lk(s->q_mu);
      s->pending.push_back(Seg{nullptr, 0, qidx, qi, qj, qpos, end - qpos, answers, latency_us,
                               arrival_us, t0});
      qpos = end;
    }
  }
  while (s->replay_done.load(std::memory_order_acquire) - done0 < n) std::this_thread::yield();
  int rc = DV_OK;
  for (uint32_t b = 0; b < nbuilds; ++b) {
    if (build_rc[b]) {
      rc = build_rc[b];
      build_done_us[b] = -1;
      build_index[b] = 0xFFFFFFFFu;
      continue;
    }
    dv_request_info info{};
    dv_server_wait(s, tickets[b], &info);
    build_done_us[b] = info.done_us - (t0 + build_us[b]);
    build_index[b] = info.status == DV_OK ? info.index_id : 0xFFFFFFFFu;
    if (info.status != DV_OK) rc = info.status;
  }
  if (elapsed_us) *elapsed_us = now_us() - t0;
  return rc;
}

}  // extern "C"
