/**
 *  @file  diverg/range_sparse_base.hpp
 *  @brief Configuration tags of DiVerG's range-sparse operators
 *         (range_sparse_base.hpp:28-498 of DiVerG), call-compatible.
 *
 *  In the reference these tags pick a Kokkos accumulator (BTree or
 *  HBitVector), a row partition (thread or team, sequential or parallel), a
 *  launch grid (Fixed, Suggested, Auto or RunTime team/vector sizes) and an
 *  execution space.  They choose how a product is scheduled, never what it
 *  computes: every combination yields the same arrays (the reference's own
 *  test_range_sparse.cpp:512-545 runs all of them against one expected
 *  product).  Here every operator runs on the MI355X engine, whose launch
 *  shapes are set by the library's tuning (dv_set_tuning) and its
 *  autotuner, so the tags are accepted and carried but do not change the
 *  launch.
 */
#ifndef DIVERG_AMD_RANGE_SPARSE_BASE_HPP__
#define DIVERG_AMD_RANGE_SPARSE_BASE_HPP__

#include <chrono>
#include <cstddef>

#include "runtime.hpp"

namespace diverg {
  namespace amd {
    /**
     *  The execution space every operator runs in: the context's HIP stream
     *  on the device named by DIVERG_AMD_DEVICE.  fence() waits for its work.
     */
    struct DeviceSpace {
      using execution_space = DeviceSpace;
      static constexpr const char* name( ) { return "MI355X"; }
      void fence( ) const { check( dv_synchronize( context() ), "fence" ); }
    };

    /** Host-side stand-in for the reference's host execution spaces. */
    struct HostSpace {
      using execution_space = HostSpace;
      static constexpr const char* name( ) { return "Host"; }
      void fence( ) const { }
    };

    /**
     *  Wall-clock timer with the interface of the reference's TTimer
     *  template argument (reset(), seconds()), for callers without one.
     */
    class Timer {
    public:
      Timer( ) : m_start( std::chrono::steady_clock::now() ) { }
      void reset( ) { m_start = std::chrono::steady_clock::now(); }
      double seconds( ) const
      {
        return std::chrono::duration< double >( std::chrono::steady_clock::now() - m_start )
            .count();
      }

    private:
      std::chrono::steady_clock::time_point m_start;
    };
  }  // namespace amd

  /* Accumulator tags (range_sparse_base.hpp:30-47). */
  template< typename TSpec >
  struct Accumulator {
    using spec_type = TSpec;
  };

  struct BTreeTag { };

  template< unsigned int TL1Size >
  struct HBitVectorTag {
    static constexpr const unsigned int L1SizeParam = TL1Size;
  };

  using NoAccumulator = Accumulator< void >;
  using BTreeAccumulator = Accumulator< BTreeTag >;

  template< unsigned int TL1Size = 2048 >
  using HBitVectorAccumulator = Accumulator< HBitVectorTag< TL1Size > >;

  /* Row-partition tags (range_sparse_base.hpp:50-68). */
  template< typename TSpec >
  struct ExecPartition {
    using type = TSpec;
  };

  struct ThreadRangePolicyTag { };
  struct ThreadSequentialTag { };
  struct TeamSequentialTag { };
  struct ThreadParallelTag { };
  struct ThreadRangeParallelTag { };
  struct TeamFlatParallelTag { };

  using ThreadRangePolicyPartition = ExecPartition< ThreadRangePolicyTag >;
  using ThreadSequentialPartition = ExecPartition< ThreadSequentialTag >;
  using TeamSequentialPartition = ExecPartition< TeamSequentialTag >;
  using ThreadParallelPartition = ExecPartition< ThreadParallelTag >;
  using ThreadRangeParallelPartition = ExecPartition< ThreadRangeParallelTag >;
  using TeamFlatParallelPartition = ExecPartition< TeamFlatParallelTag >;

  /* Execution space of an accumulator (range_sparse_base.hpp:71-82): the
   * reference runs BTree on the host; here every product runs on the GPU. */
  template< typename TAccumulator >
  struct AccumulatorExecSpace {
    using type = amd::DeviceSpace;
  };

  /* Default partition of an accumulator (range_sparse_base.hpp:84-97). */
  template< typename TAccumulator >
  struct AccumulatorDefaultPartition;

  template< >
  struct AccumulatorDefaultPartition< BTreeAccumulator > {
    using type = ThreadRangePolicyPartition;
  };

  template< unsigned int TL1Size >
  struct AccumulatorDefaultPartition< HBitVectorAccumulator< TL1Size > > {
    using type = TeamSequentialPartition;
  };

  /* Launch-grid tags (range_sparse_base.hpp:99-105). */
  namespace grid {
    template< int TTeamSize, int TVectorSize, int TTeamWorkSize = 0 >
    struct Fixed { };
    struct Auto { };
    struct Suggested { };
    struct RunTime { };
  }  // namespace grid

  /**
   *  Launch grid of a configuration (ExecGrid, range_sparse_base.hpp:116-290).
   *  Only the RunTime grid holds state: callers set its team, vector and team
   *  work sizes (rcrs_benchmark.cpp:700-720).  The others are empty.
   */
  template< typename TExecSpace, typename TSpec >
  struct ExecGrid { };

  template< typename TExecSpace >
  struct ExecGrid< TExecSpace, grid::RunTime > {
    int m_vector_size = 1;
    int m_team_size = 1;
    int m_team_work_size = 16;

    int vector_size( const int = 0 ) const { return m_vector_size; }
    int vector_size( const std::size_t, const std::size_t ) const { return m_vector_size; }
    int team_size( const int = 0 ) const { return m_team_size; }
    int team_size( const std::size_t, const std::size_t ) const { return m_team_size; }
    int team_work_size( const int = 0 ) const { return m_team_work_size; }
    int team_work_size( const std::size_t, const std::size_t ) const { return m_team_work_size; }
  };

  template< typename TExecSpace, typename TGridSpec >
  using ExecGridType = ExecGrid< TExecSpace, TGridSpec >;

  /**
   *  SparseConfig< TGridSpec, TAccumulator, TPartition, TExecSpace >
   *  (range_sparse_base.hpp:485-496): the same members (part, accm, space,
   *  grid) and member types.
   */
  template< typename TGridSpec,
            typename TAccumulator,
            typename TPartition = typename AccumulatorDefaultPartition< TAccumulator >::type,
            typename TExecSpace = typename AccumulatorExecSpace< TAccumulator >::type >
  struct SparseConfig {
    using partition_type = TPartition;
    using accumulator_type = TAccumulator;
    using execution_space = TExecSpace;
    using grid_type = ExecGridType< execution_space, TGridSpec >;
    partition_type part;
    accumulator_type accm;
    execution_space space;
    grid_type grid;
  };

  using DefaultSparseConfiguration = SparseConfig< grid::Auto, HBitVectorAccumulator<> >;
}  // namespace diverg

#endif  // DIVERG_AMD_RANGE_SPARSE_BASE_HPP__
