#pragma once
// Dense containers of the public API: Tensor1<T,I> (vector) and Tensor2<T,I> (row-major matrix), with the
// members that user code of the reference calls on its TangentVector_T / CotangentVector_T and
// Solver_RelativeResiduals() (Dim, data, Read, Write, Fill, +=, -=, *=; reference
// src/SimplicialMesh/SimplicialMeshBase.hpp:44-45, Example2/main.cpp:83-131). The storage is host
// memory; the device path receives data() pointers.

#include "Common.hpp"

#include <initializer_list>
#include <type_traits>
#include <ostream>

namespace Repulsor
{
    template<typename T, typename I>
    class Tensor1
    {
        static_assert( std::is_integral<I>::value, "Tensor1: the index type must be an integer type" );
        std::vector<T> a;

    public:
        Tensor1() = default;
        explicit Tensor1( const I n ) : a( (size_t)std::max<I>( I(0), n ) ) {}
        Tensor1( const I n, const T init ) : a( (size_t)std::max<I>( I(0), n ), init ) {}
        Tensor1( std::initializer_list<T> l ) : a( l ) {}

        I Size() const { return (I)a.size(); }
        I Dim( const int k ) const { return k == 0 ? (I)a.size() : I(0); }
        I Rank() const { return I(1); }

        T * data() { return a.data(); }
        const T * data() const { return a.data(); }
        T & operator[]( const I i ) { return a[(size_t)i]; }
        const T & operator[]( const I i ) const { return a[(size_t)i]; }
        T & operator()( const I i ) { return a[(size_t)i]; }
        const T & operator()( const I i ) const { return a[(size_t)i]; }

        typename std::vector<T>::iterator begin() { return a.begin(); }
        typename std::vector<T>::iterator end() { return a.end(); }
        typename std::vector<T>::const_iterator begin() const { return a.begin(); }
        typename std::vector<T>::const_iterator end() const { return a.end(); }
        size_t size() const { return a.size(); }

        void Fill( const T v ) { std::fill( a.begin(), a.end(), v ); }
        void SetZero() { Fill( T(0) ); }
        void Read( const T * p ) { std::copy( p, p + a.size(), a.begin() ); }
        void Write( T * p ) const { std::copy( a.begin(), a.end(), p ); }
    };

    template<typename T, typename I>
    class Tensor2
    {
        static_assert( std::is_integral<I>::value, "Tensor2: the index type must be an integer type" );
        I m = 0, n = 0;
        std::vector<T> a;

    public:
        Tensor2() = default;
        Tensor2( const I rows, const I cols )
        :   m( std::max<I>( I(0), rows ) ), n( std::max<I>( I(0), cols ) ), a( (size_t)m * (size_t)n ) {}
        Tensor2( const I rows, const I cols, const T init )
        :   m( std::max<I>( I(0), rows ) ), n( std::max<I>( I(0), cols ) ), a( (size_t)m * (size_t)n, init ) {}
        // Copies a row-major rows x cols block with leading dimension ld (ld = cols if 0).
        Tensor2( const T * p, const I rows, const I cols, const I ld = 0 ) : Tensor2( rows, cols ) { Read( p, ld ); }

        I Dim( const int k ) const { return k == 0 ? m : ( k == 1 ? n : I(0) ); }
        I Size() const { return (I)a.size(); }
        I Rank() const { return I(2); }

        T * data() { return a.data(); }
        const T * data() const { return a.data(); }
        T * data( const I i ) { return a.data() + (size_t)i * (size_t)n; }
        const T * data( const I i ) const { return a.data() + (size_t)i * (size_t)n; }
        T & operator()( const I i, const I j ) { return a[(size_t)i * (size_t)n + (size_t)j]; }
        const T & operator()( const I i, const I j ) const { return a[(size_t)i * (size_t)n + (size_t)j]; }

        typename std::vector<T>::iterator begin() { return a.begin(); }
        typename std::vector<T>::iterator end() { return a.end(); }
        typename std::vector<T>::const_iterator begin() const { return a.begin(); }
        typename std::vector<T>::const_iterator end() const { return a.end(); }
        size_t size() const { return a.size(); }

        void Fill( const T v ) { std::fill( a.begin(), a.end(), v ); }
        void SetZero() { Fill( T(0) ); }

        // Reads / writes a row-major block with leading dimension ld (ld = Dim(1) if 0).
        void Read( const T * p, const I ld = 0 )
        {
            const size_t l = (size_t)( ld > 0 ? ld : n );
            for( I i = 0; i < m; ++i ) { std::copy( p + (size_t)i * l, p + (size_t)i * l + (size_t)n, a.begin() + (size_t)i * (size_t)n ); }
        }
        void Write( T * p, const I ld = 0 ) const
        {
            const size_t l = (size_t)( ld > 0 ? ld : n );
            for( I i = 0; i < m; ++i ) { std::copy( a.begin() + (size_t)i * (size_t)n, a.begin() + (size_t)( i + 1 ) * (size_t)n, p + (size_t)i * l ); }
        }

        Tensor2 & operator*=( const T s ) { for( auto & x : a ) { x *= s; } return *this; }
        Tensor2 & operator+=( const Tensor2 & b ) { check( b, "+=" ); for( size_t k = 0; k < a.size(); ++k ) { a[k] += b.a[k]; } return *this; }
        Tensor2 & operator-=( const Tensor2 & b ) { check( b, "-=" ); for( size_t k = 0; k < a.size(); ++k ) { a[k] -= b.a[k]; } return *this; }

    private:
        void check( const Tensor2 & b, const char * op ) const
        {
            if( b.m != m || b.n != n ) { throw std::invalid_argument( std::string( "Tensor2::operator" ) + op + ": dimensions differ" ); }
        }
    };

    template<typename T, typename I>
    std::ostream & operator<<( std::ostream & s, const Tensor1<T,I> & t )
    {
        s << "{";
        for( I i = 0; i < t.Size(); ++i ) { s << ( i ? ", " : " " ) << t[i]; }
        return s << " }";
    }

} // namespace Repulsor
