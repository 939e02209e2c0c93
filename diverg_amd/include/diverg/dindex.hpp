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
  }  // namespace util
}  // namespace diverg

#endif  // DIVERG_AMD_DINDEX_HPP__
