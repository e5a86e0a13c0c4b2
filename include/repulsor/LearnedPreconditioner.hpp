#pragma once
// Node features and model file of the learned preconditioner P = D^{1/2} (P0 + w diag(A)^-1) D^{1/2}
// (an addition of this library; the reference preconditioner P0 is the fallback).
//
// Features per cluster-tree node (all invariant under rotations and uniform scaling of the mesh):
//   0      log2(primitive count) / 16
//   1..3   half-extents of the node's bounding box, sorted decreasingly, over the root's largest
//   4..6   eigenvalues of the node's area-weighted mean projector, sorted decreasingly (the projector
//          is n n^T for triangles and I - t t^T for edges: a flat piece of surface gives 1, 0, 0, a
//          straight piece of curve 1, 1, 0)
//   7      depth / tree depth
//   8      log2(node area / root area) / 16

#include "ClusterTree.hpp"

#include <cstdio>

namespace Repulsor
{
    constexpr int LEARNED_PC_FEATURES = 9;
    constexpr int LEARNED_PC_HIDDEN   = 64;
    inline int LearnedPreconditionerParameterCount( int nfeat = LEARNED_PC_FEATURES ) { return LEARNED_PC_HIDDEN * nfeat + 2 * LEARNED_PC_HIDDEN + 2; }

    // Eigenvalues of a symmetric 3 x 3 matrix (packed 00,01,02,11,12,22), decreasing (trigonometric
    // closed form for the roots of the characteristic polynomial).
    inline void SymmetricEigenvalues3( const double * p, double * ev )
    {
        const double a00 = p[0], a01 = p[1], a02 = p[2], a11 = p[3], a12 = p[4], a22 = p[5];
        const double off = a01 * a01 + a02 * a02 + a12 * a12;
        double e0, e1, e2;
        if( off <= 1e-300 ) { e0 = a00; e1 = a11; e2 = a22; }
        else
        {
            const double q = ( a00 + a11 + a22 ) / 3.0;
            const double b00 = a00 - q, b11 = a11 - q, b22 = a22 - q;
            const double pp = std::sqrt( ( b00 * b00 + b11 * b11 + b22 * b22 + 2.0 * off ) / 6.0 );
            const double det = b00 * ( b11 * b22 - a12 * a12 ) - a01 * ( a01 * b22 - a12 * a02 ) + a02 * ( a01 * a12 - b11 * a02 );
            double r = det / ( 2.0 * pp * pp * pp );
            r = std::max( -1.0, std::min( 1.0, r ) );
            const double phi = std::acos( r ) / 3.0;
            const double two_pi_3 = 2.0943951023931954923;
            e0 = q + 2.0 * pp * std::cos( phi );
            e2 = q + 2.0 * pp * std::cos( phi + two_pi_3 );
            e1 = 3.0 * q - e0 - e2;
        }
        double v[3] = { e0, e1, e2 };
        std::sort( v, v + 3, []( double x, double y ) { return x > y; } );
        for( int k = 0; k < 3; ++k ) { ev[k] = v[k]; }
    }

    // C x LEARNED_PC_FEATURES floats in tree order.
    template<int DOM_DIM>
    std::vector<float> LearnedPreconditionerFeatures( const ClusterTree<DOM_DIM> & S )
    {
        constexpr int F = LEARNED_PC_FEATURES;
        constexpr int CS = ClusterTree<DOM_DIM>::C_SIZE;
        const Int C = S.ClusterCount();
        std::vector<float> f( (size_t)C * F, 0.0f );
        const double * root_box = &S.C_serialized[0];
        const double E = std::max( { root_box[4], root_box[5], root_box[6], 1e-300 } );
        const double A0 = std::max( S.C_far[0], 1e-300 );
        const double depth = std::max( 1, S.TreeDepth() );
        for( Int c = 0; c < C; ++c )
        {
            float * o = &f[(size_t)F * c];
            const double n = (double)( S.C_end[(size_t)c] - S.C_begin[(size_t)c] );
            o[0] = (float)( std::log2( std::max( 1.0, n ) ) / 16.0 );
            double h[3] = { S.C_serialized[(size_t)CS * c + 4], S.C_serialized[(size_t)CS * c + 5], S.C_serialized[(size_t)CS * c + 6] };
            std::sort( h, h + 3, []( double x, double y ) { return x > y; } );
            for( int k = 0; k < 3; ++k ) { o[1 + k] = (float)( h[k] / E ); }
            const double * far = &S.C_far[(size_t)10 * c];
            double ev[3];
            SymmetricEigenvalues3( far + 4, ev );
            for( int k = 0; k < 3; ++k ) { o[4 + k] = (float)ev[k]; }
            o[7] = (float)( S.C_depth[(size_t)c] / depth );
            o[8] = (float)( std::log2( std::max( far[0], 1e-300 ) / A0 ) / 16.0 );
        }
        return f;
    }

    // Model file: magic "RPLP", int32 nfeat, int32 parameter count, then the float32 parameters
    // (W1 row-major 64 x nfeat, b1, w2, b2, log w).
    inline bool ReadLearnedPreconditioner( const std::string & path, std::vector<float> & params, int & nfeat )
    {
        FILE * fp = std::fopen( path.c_str(), "rb" );
        if( !fp ) { return false; }
        char magic[4];
        int32_t hdr[2];
        bool ok = std::fread( magic, 1, 4, fp ) == 4 && std::string( magic, 4 ) == "RPLP" && std::fread( hdr, sizeof(int32_t), 2, fp ) == 2;
        if( ok ) { ok = hdr[0] > 0 && hdr[1] == LearnedPreconditionerParameterCount( hdr[0] ); }
        if( ok )
        {
            params.resize( (size_t)hdr[1] );
            ok = std::fread( params.data(), sizeof(float), params.size(), fp ) == params.size();
            nfeat = hdr[0];
        }
        std::fclose( fp );
        return ok;
    }

} // namespace Repulsor
