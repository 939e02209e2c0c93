/**
 *  @file  diverg/runtime.hpp
 *  @brief Device runtime of the MI355X engine behind DiVerG's runtime API
 *         (runtime.hpp:25-55; src/eti/runtime.cpp:22-49 of DiVerG).
 *
 *  initialize_runtime() opens one dv_context on the HIP device named by
 *  DIVERG_AMD_DEVICE (default 0); every wrapper call of this process uses it.
 *  Wrapper calls made before initialize_runtime() open it on first use.
 */
#ifndef DIVERG_AMD_RUNTIME_HPP__
#define DIVERG_AMD_RUNTIME_HPP__

#include <cstdlib>
#include <iostream>
#include <new>
#include <stdexcept>
#include <string>

#include "diverg_amd.h"

namespace diverg {
  namespace amd {
    inline dv_context*&
    context_slot( )
    {
      static dv_context* ctx = nullptr;
      return ctx;
    }

    inline dv_context*
    context( )
    {
      dv_context*& ctx = context_slot();
      if ( ctx == nullptr ) {
        const char* dev = std::getenv( "DIVERG_AMD_DEVICE" );
        int rc = dv_init( dev ? std::atoi( dev ) : 0, &ctx );
        if ( rc != DV_OK ) {
          ctx = nullptr;
          throw std::runtime_error( std::string( "dv_init failed: " ) + dv_status_string( rc ) );
        }
      }
      return ctx;
    }

    /** Turn a library status into the reference's failure behaviour. */
    inline void
    check( int rc, const char* what )
    {
      if ( rc == DV_OK ) return;
      std::string msg = std::string( what ) + ": " + dv_status_string( rc ) + ": "
                        + dv_last_error( context_slot() );
      if ( rc == DV_ERR_NOMEM ) throw std::bad_alloc();
      throw std::runtime_error( msg );
    }

    /**
     *  What DIVERG_ASSERT does on failure (utils.hpp:46-49, 80-88): print the
     *  failed expression and exit with status 134.
     */
    inline void
    assert_fail( const char* expr, const char* func )
    {
      std::cout << "diverg_amd: " << func << ": Assertion `" << expr << "' failed." << "\n"
                << "Aborted." << std::endl;
      std::exit( 134 );
    }
  }  // namespace amd

  inline void
  initialize_runtime( )
  {
    amd::context();
  }

  inline void
  finalize_runtime( )
  {
    dv_context*& ctx = amd::context_slot();
    if ( ctx ) dv_finalize( ctx );
    ctx = nullptr;
  }

  class ScopedRuntime {
  public:
    explicit ScopedRuntime( bool finalize_on_destroy = true ) : m_finalize( finalize_on_destroy )
    {
      initialize_runtime();
    }
    ~ScopedRuntime( )
    {
      if ( m_finalize ) finalize_runtime();
    }
    ScopedRuntime( ScopedRuntime const& ) = delete;
    ScopedRuntime& operator=( ScopedRuntime const& ) = delete;
    ScopedRuntime( ScopedRuntime&& ) = delete;
    ScopedRuntime& operator=( ScopedRuntime&& ) = delete;

  private:
    bool m_finalize;
  };
}  // namespace diverg

#define DIVERG_AMD_ASSERT( expr ) \
  ( ( expr ) ? static_cast< void >( 0 ) : diverg::amd::assert_fail( #expr, __func__ ) )

#endif  // DIVERG_AMD_RUNTIME_HPP__
