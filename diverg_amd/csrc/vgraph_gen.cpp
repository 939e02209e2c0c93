// Seeded synthetic variation-graph generator (host code).
//
// The adjacency semantics follow util::adjacency_matrix (adjacency.hpp:
// 170-228): base k of a node points to base k+1; the last base of a node
// points to the first base of every out-neighbour node; rows are sorted.
// Duplicate edges are dropped so that every row is canonical range-CRS.
//
// Topology (all positions are backbone bases, 0-based):
//   SNP at p        1-base alternate vertex x: (p-1) -> x -> (p+1)
//   deletion p,len  (p-1) -> (p+len)
//   insertion p,len new chain x0..x(len-1): (p-1) -> x0 ... x(len-1) -> p
//   inversion p,len new chain carrying the reverse-complement copy of
//                   [p, p+len): (p-1) -> x0 ... x(len-1) -> (p+len); link
//                   orientation is ignored as in adjacency.hpp:212-217
// Node boundaries are cut at uniform lengths and forced at every event edge,
// so each bubble leaves from the last base of a node and enters the first
// base of a node.  Vertices are numbered in backbone order; the vertices of
// a bubble's path follow the last base of the node the bubble leaves from.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "dv_internal.hpp"

namespace {

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Distributions written out explicitly: the std:: ones are implementation
// defined, mt19937_64's raw output is not.
struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t seed) : g(seed) {}
  double u01() { return (double)(g() >> 11) * 0x1.0p-53; }  // [0, 1)
  double u01_open0() { return 1.0 - u01(); }                 // (0, 1]
  uint64_t uniform(uint64_t a, uint64_t b) {                 // [a, b]
    return a + (uint64_t)(u01() * (double)(b - a + 1));
  }
  // trials up to and including the first success, P(success) = p
  uint64_t geometric(double p) {
    if (p >= 1.0) return 1;
    double v = std::floor(std::log(u01_open0()) / std::log1p(-p));
    if (v > 1e18) v = 1e18;
    return 1 + (uint64_t)v;
  }
  uint64_t log_uniform(uint64_t lo, uint64_t hi) {
    double a = std::log((double)lo), b = std::log((double)hi + 1.0);
    uint64_t v = (uint64_t)std::floor(std::exp(a + u01() * (b - a)));
    return v < lo ? lo : (v > hi ? hi : v);
  }
  double normal() {  // Box-Muller
    double u1 = u01_open0(), u2 = u01();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

enum Kind : uint8_t { SNP = 0, DEL = 1, INS = 2, INV = 3 };

struct Event {
  uint64_t p;
  uint32_t len;
  uint8_t kind;
  uint32_t order;  // generation order, breaks ties after (p, kind)
};

uint32_t alt_len(const Event &e) {
  switch (e.kind) {
    case SNP: return 1;
    case DEL: return 0;
    default: return e.len;
  }
}

}  // namespace

extern "C" int dv_graph_preset(uint32_t preset, uint32_t component,
                               dv_graph_params *p) {
  if (!p) return DV_ERR_INVALID;
  std::memset(p, 0, sizeof(*p));
  p->node_min = 1;
  p->node_max = 32;
  p->snp_rate = 0.01;
  p->sv_min = 50;
  p->sv_max = 5000;
  p->sv_inversion_frac = 0.30;
  p->sv_deletion_frac = 0.35;
  switch (preset) {
    case 1:
      p->backbone_bp = 100000;
      p->indel_rate = 0.002;
      p->indel_geometric = 0;
      p->indel_min = 1;
      p->indel_max = 10;
      p->seed = 0xD1E50000ull + 1;
      break;
    case 2:
    case 3:
    case 4:
    case 5:
      p->backbone_bp = preset == 2 ? 5000000ull : 50000000ull;
      p->indel_rate = 0.001;
      p->indel_geometric = 1;
      p->indel_min = 1;
      p->indel_max = 50;
      p->indel_mean = 4.0;
      p->sv_rate = (preset == 3 || preset == 5) ? 1e-5 : 0.0;
      p->seed = 0xD1E50000ull + preset;
      if (preset == 4) {
        p->seed = 0xD1E54000ull + component;
        Rng r(splitmix64(p->seed ^ 0x4C454E475448ull));  // "LENGTH"
        double len = std::exp(std::log(5e6) + 1.0 * r.normal());
        len = std::min(std::max(len, 1e5), 1.2e8);
        p->backbone_bp = (uint64_t)len;
      }
      break;
    default:
      return DV_ERR_INVALID;
  }
  return DV_OK;
}

extern "C" int dv_generate_graph(const dv_graph_params *prm, uint32_t size_bytes,
                                 dv_rcrs *out, uint32_t **vpos_out) {
  if (!prm || !out || (size_bytes != 4 && size_bytes != 8)) return DV_ERR_INVALID;
  const uint64_t L = prm->backbone_bp;
  if (L < 2 || prm->node_min < 1 || prm->node_max < prm->node_min)
    return DV_ERR_INVALID;
  try {
    // 1. events, one stream per kind
    std::vector<Event> ev;
    uint32_t order = 0;
    auto scan = [&](double rate, uint64_t salt, auto &&draw) {
      if (rate <= 0.0) return;
      Rng r(splitmix64(prm->seed ^ salt));
      int64_t p = -1;
      for (;;) {
        p += (int64_t)r.geometric(rate);
        if ((uint64_t)p >= L) break;
        draw(r, (uint64_t)p);
      }
    };
    scan(prm->snp_rate, 0x534E50ull, [&](Rng &, uint64_t p) {
      if (p >= 1 && p + 1 <= L - 1) ev.push_back({p, 1, SNP, order++});
    });
    scan(prm->indel_rate, 0x494E44454Cull, [&](Rng &r, uint64_t p) {
      uint64_t len = prm->indel_geometric
                         ? std::min<uint64_t>(r.geometric(1.0 / prm->indel_mean),
                                              prm->indel_max)
                         : r.uniform(prm->indel_min, prm->indel_max);
      uint8_t kind = r.u01() < 0.5 ? DEL : INS;
      if (kind == DEL && p >= 1 && p + len <= L - 1)
        ev.push_back({p, (uint32_t)len, DEL, order++});
      if (kind == INS && p >= 1 && p <= L - 1)
        ev.push_back({p, (uint32_t)len, INS, order++});
    });
    scan(prm->sv_rate, 0x5356ull, [&](Rng &r, uint64_t p) {
      uint64_t len = r.log_uniform(prm->sv_min, prm->sv_max);
      double u = r.u01();
      uint8_t kind = u < prm->sv_inversion_frac ? INV
                     : u < prm->sv_inversion_frac + prm->sv_deletion_frac ? DEL
                                                                         : INS;
      bool ok = (kind == INS) ? (p >= 1 && p <= L - 1)
                              : (p >= 1 && p + len <= L - 1);
      if (ok) ev.push_back({p, (uint32_t)len, kind, order++});
    });
    std::sort(ev.begin(), ev.end(), [](const Event &a, const Event &b) {
      if (a.p != b.p) return a.p < b.p;
      if (a.kind != b.kind) return a.kind < b.kind;
      return a.order < b.order;
    });

    // 2. node boundaries: bit b set <=> a node starts at backbone position b
    std::vector<uint64_t> bnd((L + 64) / 64 + 1, 0);
    auto mark = [&](uint64_t b) {
      if (b > 0 && b < L) bnd[b >> 6] |= 1ull << (b & 63);
    };
    {
      Rng r(splitmix64(prm->seed ^ 0x4E4F4445ull));  // "NODE"
      uint64_t pos = 0;
      while (pos < L) {
        pos += r.uniform(prm->node_min, prm->node_max);
        mark(pos);
      }
    }
    for (const Event &e : ev) {
      mark(e.p);
      if (e.kind == SNP) mark(e.p + 1);
      if (e.kind == DEL || e.kind == INV) mark(e.p + e.len);
    }
    auto node_end = [&](uint64_t pos) {  // is `pos` the last base of its node
      return pos + 1 == L || ((bnd[(pos + 1) >> 6] >> ((pos + 1) & 63)) & 1);
    };

    // 3. vertex ids
    uint64_t n = L;
    for (const Event &e : ev) n += alt_len(e);
    if (n > 0xFFFFFFFEull) return DV_ERR_OVERFLOW;
    std::vector<uint32_t> vid(L);
    std::vector<uint32_t> alt0(ev.size());
    {
      uint64_t next = 0;
      size_t k = 0;  // events are sorted by p; anchor = p - 1
      for (uint64_t pos = 0; pos < L; ++pos) {
        vid[pos] = (uint32_t)next++;
        if (node_end(pos)) {
          while (k < ev.size() && ev[k].p - 1 < pos) ++k;  // (none: p >= 1)
          size_t k2 = k;
          while (k2 < ev.size() && ev[k2].p - 1 == pos) {
            alt0[k2] = (uint32_t)next;
            next += alt_len(ev[k2]);
            ++k2;
          }
          k = k2;
        }
      }
      if (next != n) return DV_ERR_INVALID;
    }

    // 4. rows in vertex order
    dv::HostMatrix hm;
    if (!hm.allocate(n, n, size_bytes)) return DV_ERR_NOMEM;
    std::vector<uint32_t> ent;
    ent.reserve((size_t)(n * 2 + n / 8));
    uint32_t *vpos = nullptr;
    if (vpos_out) {
      vpos = (uint32_t *)std::malloc(sizeof(uint32_t) * (n ? n : 1));
      if (!vpos) return DV_ERR_NOMEM;
    }
    uint64_t row = 0;
    std::vector<uint32_t> cols;
    auto emit_row = [&](void) {
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
      for (size_t i = 0; i < cols.size();) {
        size_t j = i;
        while (j + 1 < cols.size() && cols[j + 1] == cols[j] + 1) ++j;
        ent.push_back(cols[i]);
        ent.push_back(cols[j]);
        i = j + 1;
      }
      cols.clear();
      ++row;
      hm.set_rowmap(row, ent.size());
    };
    auto target = [&](const Event &e) -> uint32_t {
      switch (e.kind) {
        case SNP: return vid[e.p + 1];
        case INS: return vid[e.p];
        default: return vid[e.p + e.len];
      }
    };
    hm.set_rowmap(0, 0);
    size_t k = 0;
    for (uint64_t pos = 0; pos < L; ++pos) {
      if (vpos) vpos[row] = (uint32_t)pos;
      if (pos + 1 < L) cols.push_back(vid[pos + 1]);
      size_t k2 = k;
      if (node_end(pos)) {
        while (k2 < ev.size() && ev[k2].p - 1 == pos) {
          const Event &e = ev[k2];
          cols.push_back(e.kind == DEL ? target(e) : alt0[k2]);
          ++k2;
        }
      }
      emit_row();
      for (size_t q = k; q < k2; ++q) {  // the bubble paths leaving `pos`
        const Event &e = ev[q];
        uint32_t m = alt_len(e);
        for (uint32_t t = 0; t < m; ++t) {
          if (vpos) vpos[row] = (uint32_t)pos;
          cols.push_back(t + 1 < m ? alt0[q] + t + 1 : target(e));
          emit_row();
        }
      }
      k = k2;
    }
    if (row != n) {
      std::free(vpos);
      return DV_ERR_INVALID;
    }
    if (size_bytes == 4 && ent.size() > 0xFFFFFFFFull) {
      std::free(vpos);
      return DV_ERR_OVERFLOW;
    }
    if (!hm.set_entries(ent.data(), ent.size())) {
      std::free(vpos);
      return DV_ERR_NOMEM;
    }
    hm.release_to(out);
    out->nnz = 0;
    {
      uint64_t nnz = 0;
      for (size_t i = 0; i < ent.size(); i += 2) nnz += (uint64_t)ent[i + 1] - ent[i] + 1;
      out->nnz = size_bytes == 4 ? (uint32_t)nnz : nnz;
    }
    if (vpos_out) *vpos_out = vpos;
    return DV_OK;
  } catch (const std::bad_alloc &) {
    return DV_ERR_NOMEM;
  }
}

extern "C" void dv_free_host(void *ptr) { std::free(ptr); }
