// Host-side sequence graph shared by the GFA loader (vgraph_gfa.cpp) and the
// device adjacency build (dv_graph.hip).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace dv {

struct Graph {
  std::vector<std::string> names;   // segment names in node order (may be empty)
  std::vector<uint32_t> len;        // bases per node
  std::vector<uint64_t> charorder;  // first locus of each node, n + 1 entries
  std::vector<uint64_t> eoff;       // out-edge offsets, n + 1 entries
  std::vector<uint32_t> etgt;       // out-neighbours of each node, ascending, distinct

  // sort and deduplicate the out-lists, recompute charorder
  void finish();
  // node perm[i] becomes node i
  void permute(const std::vector<uint32_t> &perm);
  // weakly connected component of each node (its lowest node)
  std::vector<uint32_t> weak_components() const;
};

}  // namespace dv
