/**
 *  @file  diverg/range_sparse.hpp
 *  @brief Range SpAdd / SpGEMM / power on the MI355X engine, call-compatible
 *         with DiVerG's range_sparse.hpp: the operator classes RangeSpAdd
 *         (:430-468), RangeSpGEMM (:2283-2324) and RangePower (:2334-2360),
 *         the free functions range_spadd (:470-477), range_spgemm
 *         (:2326-2332), range_power (:2362-2368) that forward to them, and
 *         create_range_identity_matrix (:91-110).
 *
 *  The SparseConfig / execution-space arguments are accepted and carried
 *  (range_sparse_base.hpp): launch shapes are the library's tuning
 *  (dv_set_tuning).  Operands are copied to HBM and the result is copied back
 *  into a host matrix, as the reference's matrix-level calls do with Kokkos
 *  views.
 */
#ifndef DIVERG_AMD_RANGE_SPARSE_HPP__
#define DIVERG_AMD_RANGE_SPARSE_HPP__

#include <cassert>
#include <cstdint>
#include <cstdlib>
#include <limits>
#include <type_traits>
#include <utility>
#include <vector>

#include "crs_matrix.hpp"
#include "range_sparse_base.hpp"
#include "runtime.hpp"

namespace diverg {
  namespace amd {
    template< typename TMatrix >
    inline dv_rcrs
    describe( TMatrix const& m )
    {
      dv_rcrs d{};
      auto const& rm = m.get_rowmap();
      auto const& en = m.get_entries();
      d.nrows = rm.empty() ? 0 : rm.size() - 1;
      d.ncols = m.numCols();
      d.rowmap = const_cast< typename TMatrix::size_type* >( rm.data() );
      // int32_t ordinals hold non-negative columns: the same bits as uint32_t
      d.entries = reinterpret_cast< uint32_t* >(
          const_cast< typename TMatrix::ordinal_type* >( en.data() ) );
      d.n_entries = en.size();
      d.nnz = m.nnz();
      d.size_bytes = sizeof( typename TMatrix::size_type );
      d.location = DV_HOST;
      d.owner = nullptr;
      return d;
    }

    /** Move a library-allocated host result into a matrix and free it. */
    template< typename TMatrix >
    inline TMatrix
    adopt( dv_rcrs& d )
    {
      typedef typename TMatrix::size_type size_type;
      auto const* rm = static_cast< const size_type* >( d.rowmap );
      typename TMatrix::rowmap_type rowmap( rm, rm + d.nrows + 1 );
      typename TMatrix::entries_type entries( d.entries, d.entries + d.n_entries );
      const size_type nnz = static_cast< size_type >( d.nnz );
      const auto ncols = static_cast< typename TMatrix::ordinal_type >( d.ncols );
      dv_free( context(), &d );
      TMatrix out( ncols, std::move( entries ), std::move( rowmap ), nnz );
      return out;
    }
  }  // namespace amd

  /**
   *  The identity of order n in range CRS (range_sparse.hpp:91-110): row i
   *  holds the one range [i, i].
   */
  template< typename TRCRSMatrix, typename TExecSpace = amd::DeviceSpace >
  inline TRCRSMatrix
  create_range_identity_matrix( typename TRCRSMatrix::ordinal_type n, TExecSpace = {} )
  {
    dv_rcrs out{};
    amd::check( dv_range_identity( amd::context(), n, sizeof( typename TRCRSMatrix::size_type ),
                                   DV_HOST, &out ),
                "create_range_identity_matrix" );
    return amd::adopt< TRCRSMatrix >( out );
  }

  /* === range_spadd: c = a + b (range_sparse.hpp:430-477) === */
  template< typename TRCRSMatrix, typename TExecSpace >
  class RangeSpAdd {
  public:
    static TRCRSMatrix spadd( TRCRSMatrix const& a, TRCRSMatrix const& b, TExecSpace space );
  };

  template< typename TRCRSMatrix, typename TExecSpace >
  inline TRCRSMatrix
  RangeSpAdd< TRCRSMatrix, TExecSpace >::spadd( TRCRSMatrix const& a, TRCRSMatrix const& b,
                                                TExecSpace )
  {
    assert( a.numCols() == b.numCols() );
    assert( a.numRows() == b.numRows() );
    dv_rcrs da = amd::describe( a ), db = amd::describe( b ), out{};
    amd::check( dv_range_spadd( amd::context(), &da, &db, DV_HOST, &out ), "range_spadd" );
    return amd::adopt< TRCRSMatrix >( out );
  }

  template< typename TRCRSMatrix, typename TExecSpace = amd::DeviceSpace >
  inline TRCRSMatrix
  range_spadd( TRCRSMatrix const& a, TRCRSMatrix const& b, TExecSpace space = {} )
  {
    return RangeSpAdd< TRCRSMatrix, TExecSpace >::spadd( a, b, space );
  }

  /* === range_spgemm: c = a . b (range_sparse.hpp:2283-2332) === */
  template< typename TRCRSMatrix, typename TSparseConfig >
  class RangeSpGEMM {
  public:
    static TRCRSMatrix spgemm( TRCRSMatrix const& a, TRCRSMatrix const& b,
                               TSparseConfig config );
  };

  template< typename TRCRSMatrix, typename TSparseConfig >
  inline TRCRSMatrix
  RangeSpGEMM< TRCRSMatrix, TSparseConfig >::spgemm( TRCRSMatrix const& a, TRCRSMatrix const& b,
                                                     TSparseConfig )
  {
    typedef typename TRCRSMatrix::ordinal_type ordinal_type;
    assert( a.numCols() == b.numRows() );
    // range_sparse.hpp:2299-2300
    DIVERG_AMD_ASSERT( a.numCols() <= std::numeric_limits< ordinal_type >::max() - 1 );
    DIVERG_AMD_ASSERT( b.numCols() <= std::numeric_limits< ordinal_type >::max() - 1 );
    dv_rcrs da = amd::describe( a ), db = amd::describe( b ), out{};
    amd::check( dv_range_spgemm( amd::context(), &da, &db, DV_HOST, &out ), "range_spgemm" );
    return amd::adopt< TRCRSMatrix >( out );
  }

  template< typename TRCRSMatrix, typename TSparseConfig = DefaultSparseConfiguration >
  inline TRCRSMatrix
  range_spgemm( TRCRSMatrix const& a, TRCRSMatrix const& b, TSparseConfig config = {} )
  {
    return RangeSpGEMM< TRCRSMatrix, TSparseConfig >::spgemm( a, b, config );
  }

  /* === range_power: c = a^k (range_sparse.hpp:2334-2368) === */
  template< typename TRCRSMatrix, typename TSparseConfig >
  class RangePower {
  public:
    static TRCRSMatrix power( TRCRSMatrix const& a, unsigned int k, TSparseConfig config );
  };

  template< typename TRCRSMatrix, typename TSparseConfig >
  inline TRCRSMatrix
  RangePower< TRCRSMatrix, TSparseConfig >::power( TRCRSMatrix const& a, unsigned int k,
                                                   TSparseConfig )
  {
    dv_rcrs da = amd::describe( a ), out{};
    amd::check( dv_range_power( amd::context(), &da, k, DV_HOST, &out ), "range_power" );
    return amd::adopt< TRCRSMatrix >( out );
  }

  template< typename TRCRSMatrix, typename TSparseConfig = DefaultSparseConfiguration >
  inline TRCRSMatrix
  range_power( TRCRSMatrix const& a, unsigned int k, TSparseConfig config = {} )
  {
    return RangePower< TRCRSMatrix, TSparseConfig >::power( a, k, config );
  }
}  // namespace diverg

#endif  // DIVERG_AMD_RANGE_SPARSE_HPP__
