// Sequence graphs on the host: GFA 1/2 loading, node orderings, component
// blocks.  The device-side adjacency build is in dv_graph.hip.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <queue>
#include <random>
#include <string>
#include <unordered_map>
#include <vector>

#include "diverg_amd_graph.h"
#include "dv_graph.hpp"

namespace dv {

void Graph::finish() {
  const uint64_t n = len.size();
  charorder.assign(n + 1, 0);
  for (uint64_t k = 0; k < n; ++k) charorder[k + 1] = charorder[k] + len[k];
  // out-neighbours ascending and distinct
  for (uint64_t k = 0; k < n; ++k) {
    auto b = etgt.begin() + eoff[k], e = etgt.begin() + eoff[k + 1];
    std::sort(b, e);
  }
  std::vector<uint64_t> off(n + 1, 0);
  uint64_t w = 0;
  for (uint64_t k = 0; k < n; ++k) {
    off[k] = w;
    for (uint64_t p = eoff[k]; p < eoff[k + 1]; ++p)
      if (p == eoff[k] || etgt[p] != etgt[p - 1]) etgt[w++] = etgt[p];
  }
  off[n] = w;
  etgt.resize(w);
  eoff.swap(off);
}

// Build from (from, to) pairs; node lengths already set.
static void build_csr(Graph &g, const std::vector<std::pair<uint32_t, uint32_t>> &edges) {
  const uint64_t n = g.len.size();
  g.eoff.assign(n + 1, 0);
  for (const auto &e : edges) g.eoff[e.first + 1] += 1;
  for (uint64_t k = 0; k < n; ++k) g.eoff[k + 1] += g.eoff[k];
  g.etgt.assign(edges.size(), 0);
  std::vector<uint64_t> fill(g.eoff.begin(), g.eoff.end() - 1);
  for (const auto &e : edges) g.etgt[fill[e.first]++] = e.second;
  g.finish();
}

// Renumber: node `perm[i]` (old number) becomes node i.
void Graph::permute(const std::vector<uint32_t> &perm) {
  const uint64_t n = len.size();
  std::vector<uint32_t> rank(n);
  for (uint64_t i = 0; i < n; ++i) rank[perm[i]] = (uint32_t)i;
  std::vector<uint32_t> nlen(n);
  std::vector<std::string> nnames(names.empty() ? 0 : n);
  std::vector<std::pair<uint32_t, uint32_t>> edges;
  edges.reserve(etgt.size());
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t o = perm[i];
    nlen[i] = len[o];
    if (!names.empty()) nnames[i] = std::move(names[o]);
    for (uint64_t p = eoff[o]; p < eoff[o + 1]; ++p) edges.emplace_back((uint32_t)i, rank[etgt[p]]);
  }
  len.swap(nlen);
  names.swap(nnames);
  build_csr(*this, edges);
}

// Weakly connected component of every node, numbered by lowest node.
std::vector<uint32_t> Graph::weak_components() const {
  const uint64_t n = len.size();
  std::vector<uint32_t> parent(n);
  std::iota(parent.begin(), parent.end(), 0u);
  auto find = [&](uint32_t x) {
    while (parent[x] != x) {
      parent[x] = parent[parent[x]];
      x = parent[x];
    }
    return x;
  };
  for (uint64_t u = 0; u < n; ++u)
    for (uint64_t p = eoff[u]; p < eoff[u + 1]; ++p) {
      uint32_t a = find((uint32_t)u), b = find(etgt[p]);
      if (a != b) parent[std::max(a, b)] = std::min(a, b);
    }
  std::vector<uint32_t> comp(n);
  for (uint64_t u = 0; u < n; ++u) comp[u] = find((uint32_t)u);  // root = lowest node
  return comp;
}

}  // namespace dv

using dv::Graph;

namespace {

// Split a line on tabs (GFA fields).
void split_tabs(const std::string &line, std::vector<std::string> &f) {
  f.clear();
  size_t s = 0;
  while (true) {
    const size_t t = line.find('\t', s);
    f.emplace_back(line.substr(s, t == std::string::npos ? std::string::npos : t - s));
    if (t == std::string::npos) break;
    s = t + 1;
  }
}

bool all_digits(const std::string &s) {
  if (s.empty()) return false;
  for (char c : s)
    if (c < '0' || c > '9') return false;
  return true;
}

// A GFA 2 reference "name+" / "name-"; GFA 1 names carry no sign.
std::string strip_sign(const std::string &s) {
  if (!s.empty() && (s.back() == '+' || s.back() == '-')) return s.substr(0, s.size() - 1);
  return s;
}

}  // namespace

extern "C" {

int dv_graph_load_gfa(const char *path, uint32_t version, dv_graph **out) {
  if (!path || !out || version > 2) return DV_ERR_INVALID;
  *out = nullptr;
  std::FILE *fp = std::fopen(path, "rb");
  if (!fp) return DV_ERR_IO;
  auto *g = new Graph();
  std::unordered_map<std::string, uint32_t> id;
  std::vector<std::pair<std::string, std::string>> links;
  std::vector<std::string> f;
  std::string line;
  std::vector<char> buf(1 << 16);
  uint32_t ver = version;
  int rc = DV_OK;
  bool eof = false;
  while (!eof && rc == DV_OK) {
    // read one line of any length
    line.clear();
    while (true) {
      if (!std::fgets(buf.data(), (int)buf.size(), fp)) {
        eof = true;
        break;
      }
      const size_t k = std::strlen(buf.data());
      line.append(buf.data(), k);
      if (k && buf[k - 1] == '\n') break;
    }
    while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
    if (line.empty() || line[0] == '#') continue;
    split_tabs(line, f);
    const std::string &t = f[0];
    if (t == "H") {
      for (size_t i = 1; i < f.size(); ++i)
        if (f[i].rfind("VN:Z:", 0) == 0 && version == DV_GFA_AUTO)
          ver = f[i].compare(5, 1, "2") == 0 ? DV_GFA_2 : DV_GFA_1;
    } else if (t == "S") {
      if (f.size() < 3) {
        rc = DV_ERR_IO;
        break;
      }
      uint32_t v = ver;
      if (v == DV_GFA_AUTO) v = (f.size() >= 4 && all_digits(f[2])) ? DV_GFA_2 : DV_GFA_1;
      uint64_t L = 0;
      if (v == DV_GFA_2) {
        if (f.size() < 4 || !all_digits(f[2])) {
          rc = DV_ERR_IO;
          break;
        }
        L = std::strtoull(f[2].c_str(), nullptr, 10);
      } else {
        L = f[2] == "*" ? 0 : f[2].size();
        for (size_t i = 3; i < f.size(); ++i)
          if (f[i].rfind("LN:i:", 0) == 0) L = std::strtoull(f[i].c_str() + 5, nullptr, 10);
      }
      if (L == 0 || L > 0xFFFFFFFEull || id.count(f[1])) {
        rc = DV_ERR_IO;
        break;
      }
      id.emplace(f[1], (uint32_t)g->len.size());
      g->names.push_back(f[1]);
      g->len.push_back((uint32_t)L);
    } else if (t == "L") {  // GFA 1: L from fo to to overlap
      if (f.size() < 5) {
        rc = DV_ERR_IO;
        break;
      }
      links.emplace_back(f[1], f[3]);
    } else if (t == "E") {  // GFA 2: E eid sid1+ sid2+ beg1 end1 beg2 end2 alignment
      if (f.size() < 4) {
        rc = DV_ERR_IO;
        break;
      }
      links.emplace_back(strip_sign(f[2]), strip_sign(f[3]));
    }
  }
  std::fclose(fp);
  if (rc == DV_OK) {
    std::vector<std::pair<uint32_t, uint32_t>> edges;
    edges.reserve(links.size());
    for (const auto &l : links) {
      auto a = id.find(l.first), b = id.find(l.second);
      if (a == id.end() || b == id.end()) {
        rc = DV_ERR_IO;
        break;
      }
      edges.emplace_back(a->second, b->second);
    }
    if (rc == DV_OK) {
      build_csr(*g, edges);
      if (g->charorder.back() > 0xFFFFFFFEull) rc = DV_ERR_OVERFLOW;
    }
  }
  if (rc != DV_OK) {
    delete g;
    return rc;
  }
  *out = reinterpret_cast<dv_graph *>(g);
  return DV_OK;
}

int dv_graph_from_arrays(uint64_t n_nodes, const uint32_t *node_len, uint64_t n_edges,
                         const uint32_t *from, const uint32_t *to, dv_graph **out) {
  if (!out || (n_nodes && !node_len) || (n_edges && (!from || !to))) return DV_ERR_INVALID;
  if (n_nodes > 0xFFFFFFFEull) return DV_ERR_OVERFLOW;
  *out = nullptr;
  auto *g = new Graph();
  g->len.assign(node_len, node_len + n_nodes);
  uint64_t loci = 0;
  for (uint32_t L : g->len) {
    if (L == 0) {
      delete g;
      return DV_ERR_INVALID;
    }
    loci += L;
  }
  if (loci > 0xFFFFFFFEull) {
    delete g;
    return DV_ERR_OVERFLOW;
  }
  std::vector<std::pair<uint32_t, uint32_t>> edges(n_edges);
  for (uint64_t e = 0; e < n_edges; ++e) {
    if (from[e] >= n_nodes || to[e] >= n_nodes) {
      delete g;
      return DV_ERR_INVALID;
    }
    edges[e] = {from[e], to[e]};
  }
  build_csr(*g, edges);
  *out = reinterpret_cast<dv_graph *>(g);
  return DV_OK;
}

void dv_graph_free(dv_graph *g) { delete reinterpret_cast<Graph *>(g); }

int dv_graph_info_get(const dv_graph *gp, dv_graph_info *out) {
  if (!gp || !out) return DV_ERR_INVALID;
  const Graph &g = *reinterpret_cast<const Graph *>(gp);
  const uint64_t n = g.len.size();
  *out = dv_graph_info{};
  out->nodes = n;
  out->edges = g.etgt.size();
  out->loci = g.charorder.back();
  uint64_t c = 0;
  dv_graph_components(gp, nullptr, 0, &c);
  out->components = c;
  for (uint64_t u = 0; u < n; ++u)
    for (uint64_t p = g.eoff[u]; p < g.eoff[u + 1]; ++p) {
      const uint64_t v = g.etgt[p];
      if (v != u + 1) out->breaks += 1;
      const uint64_t src = g.charorder[u + 1] - 1, dst = g.charorder[v];
      out->bandwidth = std::max(out->bandwidth, src > dst ? src - dst : dst - src);
    }
  return DV_OK;
}

int dv_graph_arrays(const dv_graph *gp, uint32_t *node_len, uint64_t *charorder,
                    uint64_t *edge_off, uint32_t *edge_to) {
  if (!gp) return DV_ERR_INVALID;
  const Graph &g = *reinterpret_cast<const Graph *>(gp);
  if (node_len) std::copy(g.len.begin(), g.len.end(), node_len);
  if (charorder) std::copy(g.charorder.begin(), g.charorder.end(), charorder);
  if (edge_off) std::copy(g.eoff.begin(), g.eoff.end(), edge_off);
  if (edge_to) std::copy(g.etgt.begin(), g.etgt.end(), edge_to);
  return DV_OK;
}

const char *dv_graph_node_name(const dv_graph *gp, uint64_t k) {
  if (!gp) return nullptr;
  const Graph &g = *reinterpret_cast<const Graph *>(gp);
  if (k >= g.len.size()) return nullptr;
  return k < g.names.size() ? g.names[k].c_str() : "";
}

int dv_graph_sort(dv_graph *gp, uint32_t order, uint64_t seed) {
  if (!gp || order > DV_ORDER_RANDOM) return DV_ERR_INVALID;
  Graph &g = *reinterpret_cast<Graph *>(gp);
  const uint64_t n = g.len.size();
  if (order == DV_ORDER_NATIVE || n == 0) return DV_OK;
  std::vector<uint32_t> perm;
  perm.reserve(n);
  if (order == DV_ORDER_RANDOM) {
    // Fisher-Yates with an unbiased bounded draw (rejection on the top bits)
    std::mt19937_64 rng(seed);
    perm.resize(n);
    std::iota(perm.begin(), perm.end(), 0u);
    for (uint64_t i = n - 1; i > 0; --i) {
      const uint64_t bound = i + 1;
      const uint64_t lim = UINT64_MAX - UINT64_MAX % bound;
      uint64_t r;
      do r = rng();
      while (r >= lim);
      std::swap(perm[i], perm[r % bound]);
    }
    g.permute(perm);
    return DV_OK;
  }
  // in-degrees and in-lists (for the undirected / topological passes)
  std::vector<uint64_t> ioff(n + 1, 0);
  for (uint32_t v : g.etgt) ioff[v + 1] += 1;
  for (uint64_t k = 0; k < n; ++k) ioff[k + 1] += ioff[k];
  std::vector<uint32_t> isrc(g.etgt.size());
  {
    std::vector<uint64_t> fill(ioff.begin(), ioff.end() - 1);
    for (uint64_t u = 0; u < n; ++u)
      for (uint64_t p = g.eoff[u]; p < g.eoff[u + 1]; ++p) isrc[fill[g.etgt[p]]++] = (uint32_t)u;
  }
  const std::vector<uint32_t> comp = g.weak_components();
  // nodes of each component, ascending; components by lowest node
  std::vector<uint64_t> coff(n + 1, 0);
  for (uint64_t u = 0; u < n; ++u) coff[comp[u] + 1] += 1;
  for (uint64_t k = 0; k < n; ++k) coff[k + 1] += coff[k];
  std::vector<uint32_t> cnodes(n);
  {
    std::vector<uint64_t> fill(coff.begin(), coff.end() - 1);
    for (uint64_t u = 0; u < n; ++u) cnodes[fill[comp[u]]++] = (uint32_t)u;
  }
  std::vector<uint8_t> placed(n, 0);
  // in-degree without self-loops (a loop never blocks its own node)
  std::vector<uint64_t> indeg(n, 0);
  for (uint64_t u = 0; u < n; ++u)
    for (uint64_t p = g.eoff[u]; p < g.eoff[u + 1]; ++p)
      if (g.etgt[p] != u) indeg[g.etgt[p]] += 1;
  for (uint64_t c = 0; c < n; ++c) {
    if (coff[c + 1] == coff[c]) continue;
    const uint32_t *nodes = cnodes.data() + coff[c];
    const uint64_t m = coff[c + 1] - coff[c];
    uint64_t next_forced = 0;  // scan position for the lowest unplaced node
    auto force = [&]() -> uint32_t {
      while (placed[nodes[next_forced]]) ++next_forced;
      return nodes[next_forced];
    };
    if (order == DV_ORDER_TOPOLOGICAL) {
      std::priority_queue<uint32_t, std::vector<uint32_t>, std::greater<uint32_t>> ready;
      for (uint64_t i = 0; i < m; ++i)
        if (indeg[nodes[i]] == 0) ready.push(nodes[i]);
      for (uint64_t done = 0; done < m; ++done) {
        uint32_t u;
        do {
          if (ready.empty()) {
            u = force();
            break;
          }
          u = ready.top();
          ready.pop();
        } while (placed[u]);
        placed[u] = 1;
        perm.push_back(u);
        for (uint64_t p = g.eoff[u]; p < g.eoff[u + 1]; ++p) {
          const uint32_t v = g.etgt[p];
          if (!placed[v] && indeg[v] > 0 && --indeg[v] == 0) ready.push(v);
        }
      }
    } else if (order == DV_ORDER_MIN_BREAKS) {
      // depth-first flavour of Kahn: successors that become ready are taken
      // next (lowest first), so a chain u -> u+1 -> ... stays consecutive
      std::vector<uint32_t> stack;
      std::priority_queue<uint32_t, std::vector<uint32_t>, std::greater<uint32_t>> roots;
      for (uint64_t i = 0; i < m; ++i)
        if (indeg[nodes[i]] == 0) roots.push(nodes[i]);
      std::vector<uint32_t> fresh;
      for (uint64_t done = 0; done < m; ++done) {
        uint32_t u = UINT32_MAX;
        while (!stack.empty() && u == UINT32_MAX) {
          const uint32_t x = stack.back();
          stack.pop_back();
          if (!placed[x]) u = x;
        }
        while (u == UINT32_MAX && !roots.empty()) {
          const uint32_t x = roots.top();
          roots.pop();
          if (!placed[x]) u = x;
        }
        if (u == UINT32_MAX) u = force();
        placed[u] = 1;
        perm.push_back(u);
        fresh.clear();
        for (uint64_t p = g.eoff[u]; p < g.eoff[u + 1]; ++p) {
          const uint32_t v = g.etgt[p];
          if (!placed[v] && indeg[v] > 0 && --indeg[v] == 0) fresh.push_back(v);
        }
        // push highest first so the lowest is popped next
        std::sort(fresh.begin(), fresh.end(), std::greater<uint32_t>());
        for (uint32_t v : fresh) stack.push_back(v);
      }
    } else {  // Cuthill-McKee over the undirected graph
      auto degree = [&](uint32_t u) {
        return (g.eoff[u + 1] - g.eoff[u]) + (ioff[u + 1] - ioff[u]);
      };
      std::vector<uint32_t> nb;
      uint64_t head = perm.size();
      for (uint64_t done = 0; done < m;) {
        // a new BFS tree from the lowest-degree unplaced node (lowest number on ties)
        uint32_t start = UINT32_MAX;
        for (uint64_t i = 0; i < m; ++i) {
          const uint32_t x = nodes[i];
          if (!placed[x] && (start == UINT32_MAX || degree(x) < degree(start))) start = x;
        }
        placed[start] = 1;
        perm.push_back(start);
        ++done;
        while (head < perm.size()) {
          const uint32_t u = perm[head++];
          nb.clear();
          for (uint64_t p = g.eoff[u]; p < g.eoff[u + 1]; ++p)
            if (!placed[g.etgt[p]]) nb.push_back(g.etgt[p]);
          for (uint64_t p = ioff[u]; p < ioff[u + 1]; ++p)
            if (!placed[isrc[p]]) nb.push_back(isrc[p]);
          std::sort(nb.begin(), nb.end(), [&](uint32_t a, uint32_t b) {
            const uint64_t da = degree(a), db = degree(b);
            return da != db ? da < db : a < b;
          });
          for (uint32_t v : nb)
            if (!placed[v]) {
              placed[v] = 1;
              perm.push_back(v);
              ++done;
            }
        }
      }
    }
  }
  g.permute(perm);
  return DV_OK;
}

int dv_graph_components(const dv_graph *gp, uint64_t *lower, uint64_t cap, uint64_t *count) {
  if (!gp || !count) return DV_ERR_INVALID;
  const Graph &g = *reinterpret_cast<const Graph *>(gp);
  const uint64_t n = g.len.size();
  // reach[r]: the highest node an edge touching a node < r reaches; a cut
  // before node r is valid when every edge with an end below r stays below r
  std::vector<uint64_t> hi(n, 0);
  for (uint64_t u = 0; u < n; ++u) {
    hi[u] = std::max<uint64_t>(hi[u], u);
    for (uint64_t p = g.eoff[u]; p < g.eoff[u + 1]; ++p) {
      const uint64_t v = g.etgt[p], a = std::min(u, v), b = std::max(u, v);
      hi[a] = std::max(hi[a], b);
    }
  }
  uint64_t c = 0, reach = 0;
  std::vector<uint64_t> cuts;
  cuts.push_back(0);
  for (uint64_t r = 0; r < n; ++r) {
    reach = std::max(reach, hi[r]);
    if (reach == r) cuts.push_back(r + 1);  // nothing below r + 1 reaches r + 1 or above
  }
  if (n == 0) cuts.assign(1, 0);
  c = cuts.size() - 1;
  *count = c;
  if (lower && cap >= c + 1) std::copy(cuts.begin(), cuts.end(), lower);
  return DV_OK;
}

}  // extern "C"
