/**
 *  @file  diverg/dindex.hpp
 *  @brief Distance-index construction on the MI355X engine, call-compatible
 *         with util::create_distance_index (dindex.hpp:46-189 of DiVerG; the
 *         compiled overloads of dindex_eti.hpp:44-53).
 *
 *  D = A^dlo . (A + I)^(dup - dlo), or (A + I)^dup when dlo == 0.
 *  D(i, j) is true iff a walk of length in [dlo, dup] leads from locus i to
 *  locus j.
 */
#ifndef DIVERG_AMD_DINDEX_HPP__
#define DIVERG_AMD_DINDEX_HPP__

#include <cassert>

#include "range_sparse.hpp"

namespace diverg {
  namespace util {
    template< typename TRCRSMatrix, typename TSparseConfig = DefaultSparseConfiguration >
    inline TRCRSMatrix
    create_distance_index( TRCRSMatrix const& ra, unsigned int dlo, unsigned int dup,
                           TSparseConfig = {} )
    {
      DIVERG_AMD_ASSERT( dlo <= dup && dup != 0 );  // dindex.hpp:66
      dv_rcrs da = amd::describe( ra ), out{};
      amd::check( dv_create_distance_index( amd::context(), &da, dlo, dup, DV_HOST, &out,
                                            nullptr ),
                  "create_distance_index" );
      return amd::adopt< TRCRSMatrix >( out );
    }

    /**
     *  util::merge_distance_index< TMutableCRSMatrix >( d1, d2 ) for the Range
     *  group (dindex.hpp:556-579): the union of two indices of one shape.  The
     *  reference takes range_spadd for in-core RangeDynamic operands
     *  (dindex.hpp:435-456); so does this (on the device).
     */
    template< typename TMutableCRSMatrix, typename TCRSMatrix >
    inline TMutableCRSMatrix
    merge_distance_index( TCRSMatrix& dindex1, TCRSMatrix& dindex2 )
    {
      assert( dindex1.numRows() == dindex2.numRows() );  // dindex.hpp:307-308
      assert( dindex1.numCols() == dindex2.numCols() );
      dv_rcrs d1 = amd::describe( dindex1 ), d2 = amd::describe( dindex2 ), out{};
      amd::check( dv_merge_distance_index( amd::context(), &d1, &d2, DV_HOST, &out ),
                  "merge_distance_index" );
      return amd::adopt< TMutableCRSMatrix >( out );
    }
  }  // namespace util
}  // namespace diverg

#endif  // DIVERG_AMD_DINDEX_HPP__
