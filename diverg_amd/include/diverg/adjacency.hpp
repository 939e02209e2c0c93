/**
 *  @file  diverg/adjacency.hpp
 *  @brief Sequence graphs and their character-level adjacency on the MI355X
 *         engine, call-compatible with DiVerG's adjacency.hpp (83-262) for
 *         the calls its tests and benchmark make.
 *
 *  gum is not part of this library: `diverg::amd::SeqGraph` stands in for
 *  gum::SeqGraph< gum::Succinct > with the members the reference's adjacency
 *  code uses (node ranks from 1, get_node_count, node_length, rank_to_id).
 *  Load a graph with `diverg::amd::load_gfa( graph, path )` in place of
 *  gum::util::load( graph, path, gum::util::GFAFormat{}, ... ).
 */
#ifndef DIVERG_AMD_ADJACENCY_HPP__
#define DIVERG_AMD_ADJACENCY_HPP__

#include <cstdint>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "diverg_amd_graph.h"
#include "range_sparse.hpp"

namespace diverg {
  namespace amd {
    /** A sequence graph (dv_graph); node ranks are 1 .. get_node_count(). */
    class SeqGraph {
    public:
      using rank_type = uint64_t;
      using id_type = uint64_t;
      using offset_type = uint64_t;

      SeqGraph( ) = default;
      explicit SeqGraph( dv_graph* g ) : g_( g ) { }
      SeqGraph( SeqGraph const& ) = delete;
      SeqGraph& operator=( SeqGraph const& ) = delete;
      SeqGraph( SeqGraph&& o ) noexcept : g_( o.g_ ) { o.g_ = nullptr; }
      SeqGraph& operator=( SeqGraph&& o ) noexcept
      {
        std::swap( g_, o.g_ );
        return *this;
      }
      ~SeqGraph( ) { dv_graph_free( g_ ); }

      rank_type get_node_count( ) const { return info().nodes; }
      /** Node ids equal ranks here (gum keeps separate ids). */
      id_type rank_to_id( rank_type rank ) const { return rank; }
      offset_type node_length( id_type id ) const
      {
        load_arrays();
        return len_.at( id - 1 );
      }
      dv_graph_info info( ) const
      {
        dv_graph_info i{};
        dv_graph_info_get( g_, &i );
        return i;
      }
      /** Renumber the nodes (the orderings of load_graph in the reference's
       *  benchmark, rcrs_benchmark.hpp:335-369). */
      void sort( dv_graph_order order, uint64_t seed = 0 )
      {
        if ( dv_graph_sort( g_, order, seed ) != DV_OK )
          throw std::runtime_error( "dv_graph_sort failed" );
        len_.clear();
      }
      dv_graph* handle( ) const { return g_; }

    private:
      void load_arrays( ) const
      {
        if ( !len_.empty() ) return;
        len_.resize( get_node_count() );
        dv_graph_arrays( g_, len_.data(), nullptr, nullptr, nullptr );
      }
      dv_graph* g_ = nullptr;
      mutable std::vector< uint32_t > len_;
    };

    /** Load a GFA file (version 0: detect from the header or the S lines). */
    inline void
    load_gfa( SeqGraph& graph, std::string const& path, uint32_t version = 0 )
    {
      dv_graph* g = nullptr;
      const int rc = dv_graph_load_gfa( path.c_str(), version, &g );
      if ( rc != DV_OK )
        throw std::runtime_error( "cannot load GFA '" + path + "': " + dv_status_string( rc ) );
      graph = SeqGraph( g );
    }

    /** Component starts as 1-based ranks, then get_node_count() + 1. */
    inline std::vector< uint64_t >
    component_bounds( SeqGraph const& graph )
    {
      uint64_t c = 0;
      dv_graph_components( graph.handle(), nullptr, 0, &c );
      std::vector< uint64_t > b( c + 1 );
      dv_graph_components( graph.handle(), b.data(), b.size(), &c );
      for ( auto& x : b ) x += 1;
      return b;
    }
  }  // namespace amd

  namespace util {
    /** gum::util::total_nof_loci for nodes [lower, upper) (1-based ranks). */
    inline uint64_t
    total_nof_loci( amd::SeqGraph const& graph, uint64_t lower = 1, uint64_t upper = 0 )
    {
      const uint64_t n = graph.get_node_count();
      if ( upper == 0 ) upper = n + 1;
      std::vector< uint64_t > co( n + 1 );
      dv_graph_arrays( graph.handle(), nullptr, co.data(), nullptr, nullptr );
      return co[ upper - 1 ] - co[ lower - 1 ];
    }

    /**
     *  upper_node_rank_in_component (adjacency.hpp:83-105): the upper rank
     *  (exclusive) of the component holding `node_rank`, the component's
     *  1-based index and the number of components.
     */
    inline std::tuple< uint64_t, std::size_t, std::size_t >
    upper_node_rank_in_component( amd::SeqGraph const& graph, uint64_t node_rank )
    {
      const auto b = amd::component_bounds( graph );
      std::size_t comp_i = 0;
      while ( comp_i + 1 < b.size() && b[ comp_i + 1 ] <= node_rank ) ++comp_i;
      return std::make_tuple( b[ comp_i + 1 ], comp_i + 1, b.size() - 1 );
    }

    /**
     *  range_adjacency_matrix< TRangeCRSMatrix >( graph, lower, upper )
     *  (adjacency.hpp:249-262; rows as adjacency_matrix, :170-228): the
     *  character-level adjacency of nodes [lower, upper), 1-based ranks,
     *  upper 0 for the whole graph, built on the device.
     */
    template< typename TRangeCRSMatrix >
    inline TRangeCRSMatrix
    range_adjacency_matrix( amd::SeqGraph const& graph, uint64_t lower = 1, uint64_t upper = 0 )
    {
      using size_type = typename TRangeCRSMatrix::size_type;
      dv_rcrs out{};
      amd::check( dv_adjacency_matrix( amd::context(), graph.handle(), lower - 1,
                                       upper == 0 ? 0 : upper - 1,
                                       (uint32_t)sizeof( size_type ), DV_HOST, &out ),
                  "range_adjacency_matrix" );
      return amd::adopt< TRangeCRSMatrix >( out );
    }
  }  // namespace util
}  // namespace diverg

#endif  // DIVERG_AMD_ADJACENCY_HPP__
