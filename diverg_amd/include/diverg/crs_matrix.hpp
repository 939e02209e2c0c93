/**
 *  @file  diverg/crs_matrix.hpp
 *  @brief Range-CRS matrix type of the MI355X engine, call-compatible with
 *         diverg::CRSMatrix< crs_matrix::RangeDynamic, bool, TOrdinal, TSize >
 *         (crs_matrix.hpp:1873-2004 of DiVerG).
 *
 *  Rows of closed column intervals [s, e] stored flat as entries =
 *  s0,e0,s1,e1,... with rowmap[i]..rowmap[i+1] delimiting row i.  Only the
 *  RangeDynamic specification exists here: the distance-index path produces
 *  and consumes nothing else.
 */
#ifndef DIVERG_AMD_CRS_MATRIX_HPP__
#define DIVERG_AMD_CRS_MATRIX_HPP__

#include <cstdint>
#include <istream>
#include <ostream>
#include <stdexcept>
#include <type_traits>
#include <utility>
#include <vector>

namespace diverg {
  namespace crs_matrix {
    struct RangeGroup { };
    struct RangeDynamic { };  // crs_matrix.hpp:54

    // binary_search(begin, end, key, RangeGroup) (crs_matrix.hpp:149-169):
    // lower bound of key in the flat entries; an odd position steps back to
    // the start of its range; true when that start is <= key.
    template< typename TIter, typename TOrdinal >
    inline bool
    binary_search( TIter begin, TIter end, TOrdinal key, RangeGroup )
    {
      auto first = begin;
      auto len = end - first;
      while ( len > 0 ) {
        auto half = len / 2;
        auto middle = first + half;
        if ( *middle < key ) {
          first = middle + 1;
          len = len - half - 1;
        }
        else len = half;
      }
      if ( ( first - begin ) % 2 ) --first;
      return first != end && *first <= key;
    }
  }  // namespace crs_matrix

  template< typename TSpec, typename TValue = bool, typename TOrdinal = uint32_t,
            typename TSize = uint64_t >
  class CRSMatrix;

  template< typename TOrdinal, typename TSize >
  class CRSMatrix< crs_matrix::RangeDynamic, bool, TOrdinal, TSize > {
    // the engine stores 32-bit column ranges; int32_t ordinals (the
    // reference's tests use them) hold the same non-negative values
    static_assert( std::is_same< TOrdinal, uint32_t >::value
                       || std::is_same< TOrdinal, int32_t >::value,
                   "the engine stores 32-bit column ranges" );
    static_assert( std::is_same< TSize, uint32_t >::value || std::is_same< TSize, uint64_t >::value,
                   "the size type is uint32_t or uint64_t" );
  public:
    typedef crs_matrix::RangeDynamic spec_type;
    typedef bool value_type;
    typedef TOrdinal ordinal_type;
    typedef TSize size_type;
    typedef std::vector< TOrdinal > entries_type;
    typedef std::vector< TSize > rowmap_type;

    CRSMatrix( ) : num_cols( 0 ), m_nnz( 0 ) { }

    /**
     *  From entries/rowmap arrays; nnz = 0 recomputes sum(e - s + 1) as
     *  crs_matrix.hpp:1895-1902 does.
     */
    CRSMatrix( ordinal_type ncols, entries_type e_entries, rowmap_type e_rowmap,
               size_type e_nnz = 0 )
      : entries( std::move( e_entries ) ), rowmap( std::move( e_rowmap ) ),
        num_cols( ncols ), m_nnz( e_nnz )
    {
      if ( this->m_nnz == 0 ) this->m_nnz = this->count_nnz();
    }

    inline ordinal_type numCols( ) const { return this->num_cols; }
    inline ordinal_type numRows( ) const
    {
      return static_cast< ordinal_type >( this->rowmap.size() - 1 );
    }
    inline size_type nnz( ) const { return this->m_nnz; }

    inline entries_type const& get_entries( ) const { return this->entries; }
    inline entries_type& get_entries( ) { return this->entries; }
    inline rowmap_type const& get_rowmap( ) const { return this->rowmap; }
    inline rowmap_type& get_rowmap( ) { return this->rowmap; }

    /** crs_matrix.hpp:1197-1218 */
    inline value_type
    operator()( ordinal_type i, ordinal_type j ) const
    {
      return crs_matrix::binary_search( this->entries.begin() + this->rowmap[ i ],
                                        this->entries.begin() + this->rowmap[ i + 1 ], j,
                                        crs_matrix::RangeGroup{} );
    }

    /** crs_matrix.hpp:1222-1243 */
    inline value_type
    at( ordinal_type i, ordinal_type j ) const
    {
      if ( i >= this->numRows() || j >= this->numCols() )
        throw std::runtime_error( "index out of range" );
      return ( *this )( i, j );
    }

    inline void
    clear( )
    {
      this->entries.clear();
      this->rowmap.clear();
      this->num_cols = 0;
      this->m_nnz = 0;
    }

    /**
     *  Binary layout of crs_matrix.hpp:1392-1409 and :1986-2000 (with
     *  utils.hpp:191-289): u64 length, entries; u64 length, rowmap;
     *  u64 ncols; u64 nnz.
     */
    inline void
    serialize( std::ostream& out ) const
    {
      write_vector( out, this->entries );
      write_vector( out, this->rowmap );
      write_u64( out, this->num_cols );
      write_u64( out, this->m_nnz );
    }

    inline void
    load( std::istream& in )
    {
      this->clear();
      read_vector( in, this->entries );
      read_vector( in, this->rowmap );
      this->num_cols = static_cast< ordinal_type >( read_u64( in ) );
      this->m_nnz = static_cast< size_type >( read_u64( in ) );
    }

  private:
    entries_type entries;
    rowmap_type rowmap;
    ordinal_type num_cols;
    size_type m_nnz;

    inline size_type
    count_nnz( ) const  // crs_matrix.hpp:175-189
    {
      size_type n = 0;
      for ( std::size_t k = 0; k + 1 < this->entries.size(); k += 2 )
        n += this->entries[ k + 1 ] - this->entries[ k ] + 1;
      return n;
    }

    static void write_u64( std::ostream& out, uint64_t v )
    {
      out.write( reinterpret_cast< const char* >( &v ), sizeof( v ) );
    }
    static uint64_t read_u64( std::istream& in )
    {
      uint64_t v = 0;
      in.read( reinterpret_cast< char* >( &v ), sizeof( v ) );
      return v;
    }
    template< typename T >
    static void write_vector( std::ostream& out, std::vector< T > const& v )
    {
      write_u64( out, v.size() );
      if ( !v.empty() )
        out.write( reinterpret_cast< const char* >( v.data() ), sizeof( T ) * v.size() );
    }
    template< typename T >
    static void read_vector( std::istream& in, std::vector< T >& v )
    {
      v.resize( read_u64( in ) );
      if ( !v.empty() ) in.read( reinterpret_cast< char* >( v.data() ), sizeof( T ) * v.size() );
    }
  };
}  // namespace diverg

#endif  // DIVERG_AMD_CRS_MATRIX_HPP__
