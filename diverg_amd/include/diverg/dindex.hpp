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
#include <iostream>

#include "range_sparse.hpp"

namespace diverg {
  namespace util {
    /**
     *  create_distance_index( ra, dlo, dup, config, timer1_ptr, timer2_ptr )
     *  (dindex.hpp:46-189).  timer1 is reset when the input is resident on
     *  the device and the index is built ("whole body minus host<->device
     *  copies", dindex.hpp:52-53); timer2 is reset before the input's copy to
     *  the device.  With DIVERG_STATS defined the reference's timing lines
     *  are printed from them.  The config's tags do not change the launch
     *  (range_sparse_base.hpp).
     */
    template< typename TRCRSMatrix,
              typename TSparseConfig = DefaultSparseConfiguration,
              typename TTimer = amd::Timer >
    inline TRCRSMatrix
    create_distance_index( TRCRSMatrix const& ra, unsigned int dlo, unsigned int dup,
                           TSparseConfig = {},
                           TTimer* timer1_ptr = nullptr,
                           TTimer* timer2_ptr = nullptr )
    {
      DIVERG_AMD_ASSERT( dlo <= dup && dup != 0 );  // dindex.hpp:66
      assert( ra.numCols() == ra.numRows() );
      dv_context* ctx = amd::context();
      dv_rcrs da = amd::describe( ra ), out{};
      if ( timer1_ptr == nullptr && timer2_ptr == nullptr ) {
        amd::check( dv_create_distance_index( ctx, &da, dlo, dup, DV_HOST, &out, nullptr ),
                    "create_distance_index" );
        return amd::adopt< TRCRSMatrix >( out );
      }
      // timed: input to the device, build there, then the index to the host
      if ( timer2_ptr ) timer2_ptr->reset();
      dv_rcrs dev{}, dd{};
      amd::check( dv_copy( ctx, &da, DV_DEVICE, &dev ), "create_distance_index: copy to device" );
      amd::check( dv_synchronize( ctx ), "create_distance_index: copy to device" );
#ifdef DIVERG_STATS
      if ( timer2_ptr )
        std::cout << "diverg::create_distance_index::copy_to_device time: "
                  << timer2_ptr->seconds() * 1000 << "ms" << std::endl;
#endif
      if ( timer1_ptr ) timer1_ptr->reset();
      int rc = dv_create_distance_index( ctx, &dev, dlo, dup, DV_DEVICE, &dd, nullptr );
      if ( rc == DV_OK ) rc = dv_synchronize( ctx );
#ifdef DIVERG_STATS
      if ( timer1_ptr && rc == DV_OK )
        std::cout << "diverg::create_distance_index time: " << timer1_ptr->seconds() * 1000
                  << "ms" << std::endl;
#endif
      dv_free( ctx, &dev );
      if ( rc != DV_OK ) dv_free( ctx, &dd );
      amd::check( rc, "create_distance_index" );
      rc = dv_copy( ctx, &dd, DV_HOST, &out );
      dv_free( ctx, &dd );
      amd::check( rc, "create_distance_index: copy to host" );
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
