// Drop-in check of the C++ wrapper headers: the reference's call sites
// (diverg::util::create_distance_index with and without its config and
// timers, diverg::range_spgemm / range_spadd / range_power and their
// RangeSpGEMM / RangeSpAdd / RangePower classes, SparseConfig with every
// accumulator/partition pair, CRSMatrix<RangeDynamic,...>::operator(), serialize/load,
// util::range_adjacency_matrix, upper_node_rank_in_component,
// util::merge_distance_index) compiled
// unchanged against diverg_amd/include and compared with the CPU oracle, in the
// manner of check_identical (test_dindex.cpp:34-53).  Needs a GPU.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <memory>
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

// A seeded random Boolean matrix of order n with about nnz set cells (rows
// sorted and duplicate-free, so the ranges are canonical), as the
// reference's create_random_binary_matrix gives test_range_sparse.cpp:512-545.
template <typename TMatrix>
static TMatrix random_matrix(uint32_t n, uint64_t nnz, uint64_t seed) {
  typedef typename TMatrix::ordinal_type O;
  typedef typename TMatrix::size_type S;
  std::vector<std::vector<uint32_t>> rows(n);
  uint64_t x = seed;
  auto next = [&]() {  // splitmix64
    x += 0x9E3779B97F4A7C15ull;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  for (uint64_t k = 0; k < nnz; ++k) rows[next() % n].push_back((uint32_t)(next() % n));
  typename TMatrix::entries_type en;
  typename TMatrix::rowmap_type rm(1, 0);
  for (auto &r : rows) {
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    for (std::size_t k = 0; k < r.size();) {
      std::size_t e = k;
      while (e + 1 < r.size() && r[e + 1] == r[e] + 1) ++e;
      en.push_back((O)r[k]);
      en.push_back((O)r[e]);
      k = e + 1;
    }
    rm.push_back((S)en.size());
  }
  return TMatrix((O)n, std::move(en), std::move(rm));
}

// test_range_sparse.cpp:512-545: SparseConfig< grid::Suggested, TAccumulator,
// TPartition > for every accumulator/partition pair the reference tests,
// range_spgemm( m, m, config_type{} ) on an int32_t-ordinal matrix, and the
// config's execution space fenced.
template <typename TAccumulator, typename TPartition>
static void config_scenario(const char *name) {
  using namespace diverg;
  std::printf("scenario: range_spgemm with SparseConfig< grid::Suggested, %s >\n", name);
  typedef SparseConfig<grid::Suggested, TAccumulator, TPartition> config_type;
  typedef typename config_type::execution_space execution_space;
  typedef CRSMatrix<crs_matrix::RangeDynamic, bool, int32_t, uint64_t> rcrsmatrix_t;
  rcrsmatrix_t rrand_mat = random_matrix<rcrsmatrix_t>(6521, 200000, 0xD1E5);
  auto rc = range_spgemm(rrand_mat, rrand_mat, config_type{});
  execution_space space;
  space.fence();
  // oracle on the same matrix
  std::vector<uint64_t> rm(rrand_mat.get_rowmap().begin(), rrand_mat.get_rowmap().end());
  std::vector<uint32_t> en(rrand_mat.get_entries().begin(), rrand_mat.get_entries().end());
  orc_rcrs o{(uint64_t)rrand_mat.numRows(), (uint64_t)rrand_mat.numCols(), rm.data(), en.data(),
             rrand_mat.nnz()};
  orc_rcrs e;
  REQUIRE(orc_spgemm(&o, &o, &e) == 0);
  REQUIRE((uint64_t)rc.numRows() == e.nrows && (uint64_t)rc.numCols() == e.ncols);
  REQUIRE(rc.nnz() == e.nnz);
  bool same = rc.get_entries().size() == e.rowmap[e.nrows];
  for (std::size_t i = 0; same && i < rc.get_entries().size(); ++i)
    same = (uint32_t)rc.get_entries()[i] == e.entries[i];
  for (std::size_t i = 0; same && i <= e.nrows; ++i) same = rc.get_rowmap()[i] == e.rowmap[i];
  REQUIRE(same);
  orc_free(&e);
}

// rcrs_benchmark.cpp:449-450 and :685-720: a RunTime grid whose sizes the
// caller sets, create_distance_index( ra, dlo, dup, config, &timer,
// inner_timer.get() ), and the ETI operator classes RangeSpGEMM::spgemm,
// RangeSpAdd::spadd, RangePower::power (range_sparse.hpp:430-468,
// 2283-2368) with create_range_identity_matrix( n, space ).
template <typename TSize>
static void benchmark_scenario() {
  using namespace diverg;
  std::printf("scenario: benchmark call shapes, size type %zu bytes\n", sizeof(TSize));
  typedef rcrs_t<TSize> range_crsmatrix_t;
  typedef SparseConfig<grid::RunTime, HBitVectorAccumulator<>, TeamSequentialPartition> config_type;
  typedef typename config_type::execution_space execution_space;
  config_type config;
  config.grid.m_team_work_size = 16;
  config.grid.m_team_size = 4;
  config.grid.m_vector_size = 8;
  REQUIRE(config.grid.team_size() == 4 && config.grid.vector_size() == 8);
  auto ra = graph<TSize>(1, 20000);
  OracleView<TSize> ov(ra);
  amd::Timer timer;
  std::shared_ptr<amd::Timer> inner_timer = std::make_shared<amd::Timer>();
  const unsigned dlo = 15, dup = 45;
  auto rc = util::create_distance_index(ra, dlo, dup, config, &timer, inner_timer.get());
  REQUIRE(timer.seconds() > 0.0);
  orc_rcrs e;
  REQUIRE(orc_dindex(&ov.o, dlo, dup, &e) == 0);
  check_identical(rc, e);
  orc_free(&e);
  // the ETI operator classes
  execution_space space;
  auto i = create_range_identity_matrix<range_crsmatrix_t>(ra.numRows(), space);
  OracleView<TSize> oi(i);
  auto ai = RangeSpAdd<range_crsmatrix_t, execution_space>::spadd(i, ra, space);
  REQUIRE(orc_spadd(&oi.o, &ov.o, &e) == 0);
  check_identical(ai, e);
  orc_free(&e);
  auto sq = RangeSpGEMM<range_crsmatrix_t, config_type>::spgemm(ra, ra, config);
  REQUIRE(orc_spgemm(&ov.o, &ov.o, &e) == 0);
  check_identical(sq, e);
  orc_free(&e);
  auto p = RangePower<range_crsmatrix_t, config_type>::power(ra, 13, config);
  REQUIRE(orc_power(&ov.o, 13, &e) == 0);
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
  config_scenario<diverg::BTreeAccumulator, diverg::ThreadRangePolicyPartition>(
      "BTreeAccumulator, ThreadRangePolicyPartition");
  config_scenario<diverg::HBitVectorAccumulator<>, diverg::ThreadSequentialPartition>(
      "HBitVectorAccumulator<>, ThreadSequentialPartition");
  config_scenario<diverg::HBitVectorAccumulator<>, diverg::TeamSequentialPartition>(
      "HBitVectorAccumulator<>, TeamSequentialPartition");
  config_scenario<diverg::HBitVectorAccumulator<>, diverg::ThreadRangeParallelPartition>(
      "HBitVectorAccumulator<>, ThreadRangeParallelPartition");
  config_scenario<diverg::HBitVectorAccumulator<>, diverg::ThreadParallelPartition>(
      "HBitVectorAccumulator<>, ThreadParallelPartition");
  benchmark_scenario<uint64_t>();
  benchmark_scenario<uint32_t>();
  const std::string gfa = argc > 1 ? argv[1] : "tests/golden/gfa/multi.gfa";
  graph_scenario<uint64_t>(gfa, 1, 8);   // test_dindex.cpp:180-206
  graph_scenario<uint32_t>(gfa, 1, 3);   // test_dindex.cpp:208-240
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "passed", failures);
  return failures ? 1 : 0;
}
