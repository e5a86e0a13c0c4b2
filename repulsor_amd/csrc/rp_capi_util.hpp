#pragma once
// Helpers shared by the C-ABI translation units: argument checks that become status codes,
// the thread-local error message, event timing and row packing.

#include "rp_internal.hpp"

#include <exception>
#include <new>
#include <string>
#include <vector>

namespace rp
{
    void set_error( const std::string & msg );

    // Y = alpha A X + beta Y with the tangent-point metric A on device vectors (defined in rp_capi.hip).
    void metric_apply( rp_ctx * ctx, double alpha, const double * Xd, int64_t ldx, double beta, double * Yd, int64_t ldy,
                       bool nf_flag, bool ff_flag, bool time_nf, bool vf_flag = true, bool inner32 = false );

    struct InvalidArg { std::string msg; int status; };

    inline void require( bool cond, const std::string & msg, int status = RP_EINVAL )
    {
        if( !cond ) { throw InvalidArg{ msg, status }; }
    }

    template<typename F>
    inline int guarded( const char * name, F && f )
    {
        try
        {
            f();
            return RP_OK;
        }
        catch( const InvalidArg & e ) { set_error( std::string(name) + ": " + e.msg ); return e.status; }
        catch( const HipError & e )   { set_error( std::string(name) + ": " + e.msg ); return e.status; }
        catch( const std::bad_alloc & ) { set_error( std::string(name) + ": host allocation failed" ); return RP_ENOMEM; }
        catch( const std::exception & e ) { set_error( std::string(name) + ": " + e.what() ); return RP_EINVAL; }
    }

    struct EventScope
    {
        rp_ctx * ctx; int which;
        EventScope( rp_ctx * c, int w ) : ctx(c), which(w) { RP_HIP( hipEventRecord( ctx->ev[which][0], ctx->stream ) ); }
        void stop() { RP_HIP( hipEventRecord( ctx->ev[which][1], ctx->stream ) ); }
    };

    inline void finish_timing( rp_ctx * ctx, int which )
    {
        float ms = 0;
        RP_HIP( hipEventElapsedTime( &ms, ctx->ev[which][0], ctx->ev[which][1] ) );
        ctx->timings.ms[which] = ms;
    }

    inline void pack_rows( const double * src, int64_t ld, int64_t n, std::vector<double> & dst )
    {
        dst.resize( (size_t)(3 * n) );
        for( int64_t v = 0; v < n; ++v ) { for( int k = 0; k < 3; ++k ) { dst[3 * v + k] = src[ld * v + k]; } }
    }

    inline void unpack_rows( const std::vector<double> & src, int64_t n, double * dst, int64_t ld )
    {
        for( int64_t v = 0; v < n; ++v ) { for( int k = 0; k < 3; ++k ) { dst[ld * v + k] = src[(size_t)(3 * v + k)]; } }
    }
}
