/*
 * diverg_amd_gen.h -- seeded synthetic variation-graph generator.
 *
 * No counterpart in the reference: DiVerG loads GFA files through gum
 * (rcrs_benchmark.cpp:344-505) and its random-matrix helpers are seeded from
 * std::random_device (range_sparse_utils.hpp:298), so they are not
 * reproducible.  This generator emits the character-level adjacency matrix A
 * of a linear backbone with SNP, indel and structural-variant bubbles,
 * following the semantics of util::adjacency_matrix (adjacency.hpp:170-228):
 * one vertex per base, an edge from each base to the next base of its node,
 * and an edge from the last base of a node to the first base of each
 * out-neighbour.  Rows are sorted and duplicate-free, so A is canonical
 * range-CRS (the Basic -> Range conversion, crs_matrix.hpp:197-242).
 *
 * Randomness: std::mt19937_64 (its output is fixed by the C++ standard) with
 * distributions written here, so graphs are identical on every platform
 * that has the same libm log/exp.
 */
#ifndef DIVERG_AMD_GEN_H
#define DIVERG_AMD_GEN_H

#include <stdint.h>

#include "diverg_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dv_graph_params {
  uint64_t backbone_bp;     /* backbone length in bases */
  uint64_t seed;
  uint32_t node_min, node_max;   /* backbone node lengths, uniform */
  double snp_rate;               /* per base: a 1-base alternate node */
  double indel_rate;             /* per base */
  uint32_t indel_geometric;      /* 0: uniform [indel_min, indel_max]; 1: geometric */
  uint32_t indel_min, indel_max; /* bounds (the cap for the geometric law) */
  double indel_mean;             /* mean of the geometric law (support 1, 2, ...) */
  double sv_rate;                /* per base */
  uint32_t sv_min, sv_max;       /* log-uniform length */
  double sv_inversion_frac;      /* share of SVs that are inversions */
  double sv_deletion_frac;       /* share of SVs that are deletions (rest: insertions) */
} dv_graph_params;

/* Presets for the synthetic workloads:
 *   1: 1e5 bp, SNP 0.01, indel 0.002 (uniform 1-10)
 *   2: 5e6 bp, SNP 0.01, indel 0.001 (geometric mean 4, cap 50)
 *   3: 5e7 bp, as 2 plus SVs at 1e-5 (log-uniform 50-5000; 30% inversions,
 *      35% deletions, 35% insertions)
 *   4: one component of the batched workload: the rates of preset 2 and a
 *      log-normal length (median 5e6, sigma 1.0, clipped to 1e5-1.2e8)
 *      drawn from the component's seed
 *   5: as 3 (the serving workload's indexes)
 * Seeds: preset k uses 0xD1E50000 + k; component c of preset 4 uses
 * 0xD1E54000 + c. */
int dv_graph_preset(uint32_t preset, uint32_t component, dv_graph_params *p);

/* Generate A on the host (out->location == DV_HOST, freed with dv_free).
 * `vpos`, if not NULL, receives a malloc'ed array of out->nrows uint32_t:
 * the backbone position of each vertex (for a bubble vertex, the position of
 * the base its path leaves from); release it with dv_free_host(). */
int dv_generate_graph(const dv_graph_params *p, uint32_t size_bytes,
                      dv_rcrs *out, uint32_t **vpos);
void dv_free_host(void *ptr);

#ifdef __cplusplus
}
#endif

#endif /* DIVERG_AMD_GEN_H */
