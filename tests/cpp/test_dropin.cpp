// Drop-in check of the C++ wrapper headers: the reference's call sites
// (diverg::util::create_distance_index, diverg::range_spgemm / range_spadd /
// range_power, CRSMatrix<RangeDynamic,...>::operator(), serialize/load,
// util::range_adjacency_matrix, upper_node_rank_in_component,
// util::merge_distance_index) compiled
// unchanged against diverg_amd/include and compared with the CPU oracle, in the
// manner of check_identical (test_dindex.cpp:34-53).  Needs a GPU.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>
#include <stdexcept>
#include <vector>

#include <diverg/adjacency.hpp>
#include <diverg/crs_matrix.hpp>
#include <diverg/dindex.hpp>
#include <diverg/range_sparse.hpp>
#include <diverg/runtime.hpp>

#include "diverg_amd_gen.h"

extern "C" {
typedef struct {
  uint64_t nrows, ncols;
  uint64_t *rowmap;
  uint32_t *entries;
  uint64_t nnz;
} orc_rcrs;
int orc_dindex(const orc_rcrs *a, uint32_t dlo, uint32_t dup, orc_rcrs *d);
int orc_spgemm(const orc_rcrs *a, const orc_rcrs *b, orc_rcrs *c);
int orc_spadd(const orc_rcrs *a, const orc_rcrs *b, orc_rcrs *c);
int orc_power(const orc_rcrs *a, uint32_t k, orc_rcrs *c);
int orc_identity(uint64_t n, orc_rcrs *out);
int orc_adjacency(uint64_t n_nodes, const uint32_t *node_len, const uint64_t *eoff,
                  const uint32_t *etgt, uint64_t lower, uint64_t upper, orc_rcrs *out);
void orc_free(orc_rcrs *m);
}

static int failures = 0;
#define REQUIRE(cond)                                                       \
  do {                                                                      \
    if (!(cond)) {                                                          \
      std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);       \
      ++failures;                                                           \
    }                                                                       \
  } while (0)

template <typename TSize>
using rcrs_t = diverg::CRSMatrix<diverg::crs_matrix::RangeDynamic, bool, uint32_t, TSize>;

template <typename TSize>
static rcrs_t<TSize> graph(uint32_t preset, uint64_t bp) {
  dv_graph_params p;
  dv_graph_preset(preset, 0, &p);
  p.backbone_bp = bp;
  dv_rcrs d{};
  if (dv_generate_graph(&p, sizeof(TSize), &d, nullptr) != DV_OK) std::abort();
  return diverg::amd::adopt<rcrs_t<TSize>>(d);
}

template <typename TSize>
struct OracleView {
  std::vector<uint64_t> rm;
  orc_rcrs o;
  explicit OracleView(rcrs_t<TSize> const &m) {
    rm.assign(m.get_rowmap().begin(), m.get_rowmap().end());
    o.nrows = m.numRows();
    o.ncols = m.numCols();
    o.rowmap = rm.data();
    o.entries = const_cast<uint32_t *>(m.get_entries().data());
    o.nnz = m.nnz();
  }
};

// check_identical (test_dindex.cpp:34-53)
template <typename TSize>
static void check_identical(rcrs_t<TSize> const &actual, orc_rcrs const &expected) {
  REQUIRE(actual.numRows() == expected.nrows);
  REQUIRE(actual.numCols() == expected.ncols);
  REQUIRE(actual.nnz() == static_cast<TSize>(expected.nnz));
  REQUIRE(actual.get_entries().size() == expected.rowmap[expected.nrows]);
  bool same = actual.get_entries().size() == expected.rowmap[expected.nrows];
  for (std::size_t i = 0; same && i < actual.get_entries().size(); ++i)
    same = actual.get_entries()[i] == expected.entries[i];
  REQUIRE(same);
  same = actual.get_rowmap().size() == expected.nrows + 1;
  for (std::size_t i = 0; same && i <= expected.nrows; ++i)
    same = actual.get_rowmap()[i] == expected.rowmap[i];
  REQUIRE(same);
}

template <typename TSize>
static void scenario(const char *name) {
  std::printf("scenario: distance index, size type %s\n", name);
  auto ra = graph<TSize>(1, 20000);
  OracleView<TSize> ov(ra);
  for (auto bounds : {std::pair<unsigned, unsigned>{15, 45}, {0, 30}, {6, 6}}) {
    auto d = diverg::util::create_distance_index(ra, bounds.first, bounds.second);
    orc_rcrs e;
    REQUIRE(orc_dindex(&ov.o, bounds.first, bounds.second, &e) == 0);
    check_identical(d, e);
    // queries: operator() and at() agree with a scan of the row
    for (uint32_t i = 0; i < d.numRows(); i += 997)
      for (uint32_t j = i; j < i + 60 && j < d.numCols(); ++j) {
        bool hit = false;
        for (auto k = d.get_rowmap()[i]; k < d.get_rowmap()[i + 1]; k += 2)
          hit = hit || (d.get_entries()[k] <= j && j <= d.get_entries()[k + 1]);
        REQUIRE(d(i, j) == hit);
        REQUIRE(d.at(i, j) == hit);
      }
    bool threw = false;
    try {
      d.at(d.numRows(), 0);
    } catch (std::runtime_error const &) {
      threw = true;
    }
    REQUIRE(threw);
    // serialize / load round trip
    std::stringstream ss;
    d.serialize(ss);
    rcrs_t<TSize> back;
    back.load(ss);
    check_identical(back, e);
    orc_free(&e);
  }
  std::printf("scenario: range_spgemm / range_spadd / range_power\n");
  auto sq = diverg::range_spgemm(ra, ra);
  orc_rcrs e;
  REQUIRE(orc_spgemm(&ov.o, &ov.o, &e) == 0);
  check_identical(sq, e);
  orc_free(&e);
  auto id = diverg::create_range_identity_matrix<rcrs_t<TSize>>(ra.numRows());
  OracleView<TSize> oi(id);
  auto ai = diverg::range_spadd(id, ra);
  REQUIRE(orc_spadd(&oi.o, &ov.o, &e) == 0);
  check_identical(ai, e);
  orc_free(&e);
  auto p = diverg::range_power(ra, 21);
  REQUIRE(orc_power(&ov.o, 21, &e) == 0);
  check_identical(p, e);
  orc_free(&e);
}

// test_dindex.cpp:55-240 on the reference's multi.gfa: the adjacency of the
// whole graph and of each component (walked with upper_node_rank_in_component)
// and distance indices of it, against the oracle.
template <typename TSize>
static void graph_scenario(const std::string &gfa, unsigned dmin, unsigned dmax) {
  std::printf("scenario: adjacency and distance index of %s, size type %zu bytes\n",
              gfa.c_str(), sizeof(TSize));
  diverg::amd::SeqGraph graph;
  diverg::amd::load_gfa(graph, gfa);
  const uint64_t n = graph.get_node_count();
  std::vector<uint32_t> len(n);
  std::vector<uint64_t> eoff(n + 1);
  dv_graph_info info = graph.info();
  std::vector<uint32_t> etgt(info.edges ? info.edges : 1);
  dv_graph_arrays(graph.handle(), len.data(), nullptr, eoff.data(), etgt.data());
  auto ra = diverg::util::range_adjacency_matrix<rcrs_t<TSize>>(graph);
  orc_rcrs e;
  REQUIRE(orc_adjacency(n, len.data(), eoff.data(), etgt.data(), 0, 0, &e) == 0);
  check_identical(ra, e);
  orc_free(&e);
  REQUIRE(ra.numRows() == diverg::util::total_nof_loci(graph));
  std::size_t checked = 0;
  for (uint64_t lower = 1; lower <= n;) {
    const uint64_t upper = std::get<0>(diverg::util::upper_node_rank_in_component(graph, lower));
    auto block = diverg::util::range_adjacency_matrix<rcrs_t<TSize>>(graph, lower, upper);
    REQUIRE(orc_adjacency(n, len.data(), eoff.data(), etgt.data(), lower - 1, upper - 1, &e) == 0);
    check_identical(block, e);
    orc_free(&e);
    ++checked;
    lower = upper;
  }
  REQUIRE(checked == std::get<2>(diverg::util::upper_node_rank_in_component(graph, 1)));
  REQUIRE(checked >= 2);
  OracleView<TSize> ov(ra);
  auto d = diverg::util::create_distance_index(ra, dmin, dmax);
  REQUIRE(orc_dindex(&ov.o, dmin, dmax, &e) == 0);
  check_identical(d, e);
  // merging the windows [dmin, mid] and [mid + 1, dmax] gives [dmin, dmax]
  const unsigned mid = (dmin + dmax) / 2;
  auto d1 = diverg::util::create_distance_index(ra, dmin, mid);
  auto d2 = diverg::util::create_distance_index(ra, mid + 1, dmax);
  auto merged = diverg::util::merge_distance_index<rcrs_t<TSize>>(d1, d2);
  check_identical(merged, e);
  orc_free(&e);
}

int main(int argc, char **argv) {
  diverg::ScopedRuntime runtime;
  scenario<uint64_t>("uint64_t");
  scenario<uint32_t>("uint32_t");
  const std::string gfa = argc > 1 ? argv[1] : "tests/golden/gfa/multi.gfa";
  graph_scenario<uint64_t>(gfa, 1, 8);   // test_dindex.cpp:180-206
  graph_scenario<uint32_t>(gfa, 1, 3);   // test_dindex.cpp:208-240
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "passed", failures);
  return failures ? 1 : 0;
}
