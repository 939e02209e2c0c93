/*
 * diverg_amd_graph.h -- sequence graphs for the range-CRS engine: GFA
 * loading, node orderings, component ranges and the character-level
 * adjacency matrix (the step before util::create_distance_index).
 *
 * A graph here is a list of nodes (segments) with their lengths in bases and
 * directed edges between nodes.  Nodes are numbered 0 .. n-1 in their current
 * order (the reference's gum ranks are the same numbers plus one).  Loci are
 * the bases of the nodes in that order: node k covers loci
 * charorder[k] .. charorder[k+1] - 1.
 */
#ifndef DIVERG_AMD_GRAPH_H
#define DIVERG_AMD_GRAPH_H

#include <stdint.h>

#include "diverg_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dv_graph dv_graph;

typedef enum dv_gfa_version { DV_GFA_AUTO = 0, DV_GFA_1 = 1, DV_GFA_2 = 2 } dv_gfa_version;

/* Load a GFA file (version 1: S/L lines; version 2: S/E lines; other record
 * types are skipped).  Replaces gum::util::load(graph, path, GFAFormat{}, ...)
 * as the reference's tests and benchmark call it (test_dindex.cpp:65-67,
 * rcrs_benchmark.hpp:315-333); gum and GFAKluge are not part of this library.
 *  - Nodes are numbered in the order of their S lines (the reference's
 *    "native" order with sort=false, rcrs_benchmark.hpp:315).
 *  - A node's length is its sequence length, or the LN:i: tag (GFA 1) /
 *    the length field (GFA 2) when given; a node of length 0 is an error.
 *  - An L or E record gives the edge from its first segment to its second;
 *    orientations are not used, as the reference's adjacency builder does
 *    not use the link type (adjacency.hpp:212-217).  An edge given more than
 *    once is one edge.
 * Errors: DV_ERR_IO (unreadable, unknown segment in an edge, malformed
 * record), DV_ERR_OVERFLOW (more than UINT32_MAX - 1 loci). */
int dv_graph_load_gfa(const char *path, uint32_t version, dv_graph **out);

/* Build a graph from arrays: node lengths and edges (from[e] -> to[e]). */
int dv_graph_from_arrays(uint64_t n_nodes, const uint32_t *node_len, uint64_t n_edges,
                         const uint32_t *from, const uint32_t *to, dv_graph **out);

void dv_graph_free(dv_graph *g);

typedef struct dv_graph_info {
  uint64_t nodes;
  uint64_t edges;        /* distinct directed edges */
  uint64_t loci;         /* gum::util::total_nof_loci */
  uint64_t components;   /* blocks of dv_graph_components */
  uint64_t breaks;       /* edges u -> v with v != u + 1 */
  uint64_t bandwidth;    /* max |charorder(v) - (charorder(u + 1) - 1)| over edges */
} dv_graph_info;
int dv_graph_info_get(const dv_graph *g, dv_graph_info *out);

/* Copy out the graph: node_len[n], charorder[n + 1] (prefix sums of the
 * lengths), edge_off[n + 1] and edge_to[edges] (out-neighbours of each node,
 * ascending).  Any pointer may be NULL to skip that array. */
int dv_graph_arrays(const dv_graph *g, uint32_t *node_len, uint64_t *charorder,
                    uint64_t *edge_off, uint32_t *edge_to);

/* Node name (the GFA segment name) of node k; NULL when out of range. */
const char *dv_graph_node_name(const dv_graph *g, uint64_t k);

typedef enum dv_graph_order {
  DV_ORDER_NATIVE = 0,        /* keep the current order */
  DV_ORDER_TOPOLOGICAL = 1,   /* per component, Kahn's order taking the lowest
                                 numbered ready node; on a cycle the lowest
                                 numbered unplaced node is forced next */
  DV_ORDER_MIN_BREAKS = 2,    /* per component, a topological order that
                                 continues from the node just placed whenever
                                 one of its successors becomes ready, so
                                 chains stay consecutive (few breaks) */
  DV_ORDER_CUTHILL_MCKEE = 3, /* per component, breadth-first from a lowest
                                 degree node over the undirected graph,
                                 neighbours by ascending degree */
  DV_ORDER_RANDOM = 4         /* uniform permutation from `seed` (mt19937_64) */
} dv_graph_order;

/* Renumber the nodes.  The orderings of load_graph in the reference's
 * benchmark (rcrs_benchmark.hpp:335-369); gum's own sorts are not vendored,
 * so the orders are this library's (an index is a function of the order, and
 * every order gives a valid index).  All but NATIVE and RANDOM keep each
 * weakly connected component contiguous, components by their lowest current
 * number. */
int dv_graph_sort(dv_graph *g, uint32_t order, uint64_t seed);

/* Block boundaries of the node order: lower[0] = 0 < lower[1] < ... <
 * lower[c] = n with no edge between two blocks (the finest such cut, which
 * equals the weakly connected components when each is contiguous).  With
 * cap < c + 1 only *count is set.  Replaces the component walk over
 * upper_node_rank_in_component (adjacency.hpp:83-105). */
int dv_graph_components(const dv_graph *g, uint64_t *lower, uint64_t cap, uint64_t *count);

/* Character-level adjacency of nodes [lower, upper) (upper 0: all nodes) as a
 * range-CRS matrix, built on the device: locus l -> l + 1 inside a node; the
 * last locus of a node -> the first locus of each out-neighbour; columns are
 * relative to the first locus of node `lower`.  Replaces
 * util::adjacency_matrix + range_adjacency_matrix (adjacency.hpp:170-228,
 * 249-262): the same rows, converted to ranges as the reference's Basic ->
 * Range conversion does (crs_matrix.hpp:197-242).  nnz = loci - nodes + edges
 * (adjacency.hpp:192-194).  An edge leaving [lower, upper) is DV_ERR_INVALID
 * (the reference requires the range to cover whole components). */
int dv_adjacency_matrix(dv_context *ctx, const dv_graph *g, uint64_t lower, uint64_t upper,
                        uint32_t size_bytes, uint32_t out_location, dv_rcrs *out);

#ifdef __cplusplus
}
#endif

#endif /* DIVERG_AMD_GRAPH_H */
