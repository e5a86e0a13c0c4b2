#pragma once
// Common host-side definitions of the MI355X tangent-point library.
//
// The public types follow the reference's template parameters
// (SimplicialMesh<DOM_DIM,AMB_DIM,Real,Int,LInt,SReal,ExtReal>, src/SimplicialMesh.hpp:13-18):
// Real = double, Int = int32, LInt = int64 as in Test_Preconditioner/main.cpp:28-34.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace Repulsor
{
    using Int  = int32_t;
    using LInt = int64_t;

    template<typename T> using cptr = const T *;
    template<typename T> using mptr = T *;
    template<typename T> using cref = const T &;
    template<typename T> using mref = T &;

    // Error reporting mirrors the reference's eprint/wprint: messages go to stderr and the call
    // returns early without throwing across the public API (src/Energies/EnergyDimRestricted.hpp:40-44).
    inline void eprint( const std::string & s ) { std::fprintf(stderr, "ERROR: %s\n", s.c_str()); std::fflush(stderr); }
    inline void wprint( const std::string & s ) { std::fprintf(stderr, "WARNING: %s\n", s.c_str()); std::fflush(stderr); }

    // Scoped wall-clock timer with the reference's tag names (TOOLS_PTIMER in the reference).
    // Timings accumulate in a process-wide table that rpm_timer_report() prints.
    struct TimerTable
    {
        std::mutex mtx;
        std::vector<std::pair<std::string,double>> rows;

        static TimerTable & Get() { static TimerTable t; return t; }

        void Add( const std::string & tag, double sec )
        {
            std::lock_guard<std::mutex> lk(mtx);
            for( auto & r : rows ) { if( r.first == tag ) { r.second += sec; return; } }
            rows.emplace_back( tag, sec );
        }
    };

    class ScopedTimer
    {
        std::string tag;
        std::chrono::steady_clock::time_point t0;
    public:
        explicit ScopedTimer( std::string tag_ ) : tag(std::move(tag_)), t0(std::chrono::steady_clock::now()) {}
        ~ScopedTimer()
        {
            const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            TimerTable::Get().Add( tag, s );
        }
    };

    inline double WallTime()
    {
        return std::chrono::duration<double>( std::chrono::steady_clock::now().time_since_epoch() ).count();
    }

    // Static block partition of [0,n) over `threads` workers, as JobPointers does in the reference
    // (FMM_Traversor.hpp:65-69 uses a cost-balanced variant; for O(n) loops a uniform split suffices).
    template<typename F>
    inline void ParallelFor( const LInt n, const int threads, F && f )
    {
        const int tc = std::max( 1, std::min<int>( threads, (int)std::max<LInt>( 1, n ) ) );
        if( tc <= 1 || n < 1024 )
        {
            for( LInt i = 0; i < n; ++i ) { f(i); }
            return;
        }
        std::vector<std::thread> pool;
        pool.reserve( (size_t)tc );
        for( int t = 0; t < tc; ++t )
        {
            const LInt b = (n * t) / tc;
            const LInt e = (n * (t+1)) / tc;
            pool.emplace_back( [b,e,&f]() { for( LInt i = b; i < e; ++i ) { f(i); } } );
        }
        for( auto & th : pool ) { th.join(); }
    }

    // Runs f(thread) for thread in [0,tc).
    template<typename F>
    inline void ParallelDo( const int tc, F && f )
    {
        if( tc <= 1 ) { f(0); return; }
        std::vector<std::thread> pool;
        for( int t = 0; t < tc; ++t ) { pool.emplace_back( [t,&f]() { f(t); } ); }
        for( auto & th : pool ) { th.join(); }
    }

    inline int DefaultThreadCount()
    {
        const unsigned h = std::thread::hardware_concurrency();
        return (int)std::max( 1u, std::min( h, 64u ) );
    }

    // Packed index of the symmetric (i,j) entry, i <= j, row-major upper triangle:
    // 00,01,02,11,12,22 for AMB_DIM = 3 (the order DNearFarToHulls relies on,
    // src/SimplicialMesh/SimplicialMeshDetails.hpp:1130-1191).
    template<int N>
    constexpr int lin_k( int i, int j )
    {
        if( i > j ) { const int t = i; i = j; j = t; }
        return i * N - (i * (i - 1)) / 2 + (j - i);
    }

    template<int N> constexpr int tri_i( int k ) { int i = 0; while( k >= N - i ) { k -= N - i; ++i; } return i; }
    template<int N> constexpr int tri_j( int k ) { int i = 0; while( k >= N - i ) { k -= N - i; ++i; } return i + k; }

} // namespace Repulsor
