#pragma once
// H^1 metric H = w_s K + w_m M of the P1 finite element space on the simplicial mesh and the
// nested-dissection ordering driven by the cluster tree.
//
//   Create_H1Metric            src/SimplicialMesh/H1Metric.hpp:30-169
//   NestedDissectionOrdering   src/SimplicialMesh/NestedDissectionOrdering.hpp:1-208
//
// Per simplex (vertices in input order, df = [X_1 - X_0, ..., X_n - X_0], g = df^T df,
// a = |simplex| = sqrt(det g)/n!):
//   stiffness  K_loc = a * [ sum g^-1, -row sums; -col sums, g^-1 ]
//   mass       M_loc = a * (1 + delta_ij) / (n! (n+2)!) * n!  (= a(1+delta)/12 for triangles, /6 for edges)
// assembled with duplicate entries summed into a symmetric CSR matrix (V x V).

#include "ClusterTree.hpp"

namespace Repulsor
{
    struct CSRMatrix
    {
        Int n = 0;
        std::vector<LInt>   rp;
        std::vector<Int>    ci;
        std::vector<double> v;

        void Multiply3( const double * x, double * y ) const
        {
            for( Int i = 0; i < n; ++i )
            {
                double s[3] = {0,0,0};
                for( LInt k = rp[(size_t)i]; k < rp[(size_t)i + 1]; ++k ) { for( int c = 0; c < 3; ++c ) { s[c] += v[(size_t)k] * x[(size_t)3 * ci[(size_t)k] + c]; } }
                for( int c = 0; c < 3; ++c ) { y[(size_t)3 * i + c] = s[c]; }
            }
        }
    };

    template<int DOM_DIM>
    inline CSRMatrix AssembleH1Metric( cptr<double> V, const LInt nV, cptr<Int> S, const LInt T,
                                       const double stiffness_weight, const double mass_weight )
    {
        ScopedTimer timer("H1Metric::Assemble");
        constexpr int SIZE = DOM_DIM + 1;
        double fact = 1; for( int k = 2; k <= DOM_DIM; ++k ) { fact *= k; }
        double fact2 = 1; for( int k = 2; k <= DOM_DIM + 2; ++k ) { fact2 *= k; }
        const double vol0 = 1.0 / fact;                       // StandardSimplexVolume
        const double m_factor = mass_weight / ( vol0 * fact2 );
        std::vector<std::pair<Int,Int>> keys( (size_t)(T * SIZE * SIZE) );
        std::vector<double> vals( (size_t)(T * SIZE * SIZE) );
        for( LInt t = 0; t < T; ++t )
        {
            const Int * s = S + SIZE * t;
            double df[3][DOM_DIM];
            for( int i = 0; i < 3; ++i ) { for( int j = 0; j < DOM_DIM; ++j ) { df[i][j] = V[3 * (LInt)s[j + 1] + i] - V[3 * (LInt)s[0] + i]; } }
            double g[DOM_DIM][DOM_DIM];
            for( int r = 0; r < DOM_DIM; ++r ) { for( int c = 0; c < DOM_DIM; ++c ) { double a = 0; for( int i = 0; i < 3; ++i ) { a += df[i][r] * df[i][c]; } g[r][c] = a; } }
            double ginv[DOM_DIM][DOM_DIM];
            double det;
            if constexpr ( DOM_DIM == 2 )
            {
                det = g[0][0] * g[1][1] - g[0][1] * g[1][0];
                ginv[0][0] = g[1][1] / det; ginv[1][1] = g[0][0] / det;
                ginv[0][1] = -g[0][1] / det; ginv[1][0] = -g[1][0] / det;
            }
            else
            {
                det = g[0][0];
                ginv[0][0] = 1.0 / g[0][0];
            }
            const double a = vol0 * std::sqrt( det );
            double val[SIZE][SIZE] = {};
            const double l_factor = a * stiffness_weight;
            for( int i = 0; i < DOM_DIM; ++i )
            {
                for( int j = 0; j < DOM_DIM; ++j )
                {
                    const double a_ij = l_factor * ginv[i][j];
                    val[0][0]         += a_ij;
                    val[i + 1][0]     -= a_ij;
                    val[0][j + 1]     -= a_ij;
                    val[i + 1][j + 1]  = a_ij;
                }
            }
            for( int i = 0; i < SIZE; ++i ) { for( int j = 0; j < SIZE; ++j ) { val[i][j] += a * m_factor * ( 1.0 + (i == j ? 1.0 : 0.0) ); } }
            for( int i = 0; i < SIZE; ++i )
            {
                for( int j = 0; j < SIZE; ++j )
                {
                    const size_t k = (size_t)((t * SIZE + i) * SIZE + j);
                    keys[k] = { s[i], s[j] };
                    vals[k] = val[i][j];
                }
            }
        }
        // bucket the entries by row (keeping input order), then per row sort stably by column and sum
        // duplicates in input order
        std::vector<LInt> bptr( (size_t)nV + 1, 0 );
        for( auto & k : keys ) { ++bptr[(size_t)k.first + 1]; }
        for( LInt i = 0; i < nV; ++i ) { bptr[(size_t)i + 1] += bptr[(size_t)i]; }
        std::vector<LInt> idx( keys.size() );
        {
            std::vector<LInt> pos( bptr.begin(), bptr.end() - 1 );
            for( size_t k = 0; k < keys.size(); ++k ) { idx[(size_t)pos[(size_t)keys[k].first]++] = (LInt)k; }
        }
        CSRMatrix H;
        H.n = (Int)nV;
        H.rp.assign( (size_t)nV + 1, 0 );
        std::vector<std::vector<std::pair<Int,double>>> rows( (size_t)nV );
        ParallelFor( nV, 8, [&]( LInt i )
        {
            auto b = idx.begin() + bptr[(size_t)i], e = idx.begin() + bptr[(size_t)i + 1];
            std::stable_sort( b, e, [&]( LInt x, LInt y ) { return keys[(size_t)x].second < keys[(size_t)y].second; } );
            auto & out = rows[(size_t)i];
            for( auto it = b; it != e; )
            {
                const Int col = keys[(size_t)*it].second;
                double sum = 0;
                while( it != e && keys[(size_t)*it].second == col ) { sum += vals[(size_t)*it]; ++it; }
                out.emplace_back( col, sum );
            }
        });
        for( LInt i = 0; i < nV; ++i ) { H.rp[(size_t)i + 1] = H.rp[(size_t)i] + (LInt)rows[(size_t)i].size(); }
        H.ci.resize( (size_t)H.rp.back() );
        H.v.resize( (size_t)H.rp.back() );
        for( LInt i = 0; i < nV; ++i )
        {
            LInt q = H.rp[(size_t)i];
            for( auto & cv : rows[(size_t)i] ) { H.ci[(size_t)q] = cv.first; H.v[(size_t)q] = cv.second; ++q; }
        }
        return H;
    }

    // Cluster-tree nested dissection: for every cluster whose children split its vertex range into
    // vertices touched only by the left child's primitives, only by the right child's, and by both,
    // the range is reordered stably as [left, right, separator] and the children are refined further.
    template<int DOM_DIM>
    inline std::vector<Int> NestedDissectionOrdering( const ClusterTree<DOM_DIM> & tree, const LInt nV )
    {
        ScopedTimer timer("NestedDissectionOrdering");
        constexpr int SIZE = DOM_DIM + 1;
        const Int m = (Int)nV;
        const Int T = tree.PrimitiveCount();
        const Int C = tree.ClusterCount();
        std::vector<Int> perm0( (size_t)m ), perm1( (size_t)m );
        for( Int i = 0; i < m; ++i ) { perm0[(size_t)i] = i; }
        std::vector<Int> Vb( (size_t)C, 0 ), Ve( (size_t)C, 0 );
        Vb[0] = 0; Ve[0] = m;
        // vertex -> incident primitives (tree positions): the pattern of the low-order post-operator
        std::vector<LInt> vp( (size_t)m + 1, 0 );
        for( Int i = 0; i < T; ++i ) { for( int l = 0; l < SIZE; ++l ) { ++vp[(size_t)tree.pre_cols[(size_t)(SIZE * i + l)] + 1]; } }
        for( Int v = 0; v < m; ++v ) { vp[(size_t)v + 1] += vp[(size_t)v]; }
        std::vector<Int> vprim( (size_t)vp.back() );
        {
            std::vector<LInt> pos( vp.begin(), vp.end() - 1 );
            for( Int i = 0; i < T; ++i ) { for( int l = 0; l < SIZE; ++l ) { vprim[(size_t)pos[(size_t)tree.pre_cols[(size_t)(SIZE * i + l)]]++] = i; } }
        }
        std::vector<Int> queue0 { 0 }, queue1;
        std::vector<signed char> side( (size_t)T, 0 );   // bit 0: in a left child, bit 1: in a right child
        std::vector<signed char> type( (size_t)m, -1 );
        while( !queue0.empty() )
        {
            queue1.clear();
            perm1 = perm0;
            std::fill( side.begin(), side.end(), 0 );
            for( Int C0 : queue0 )
            {
                const Int L = tree.C_left[(size_t)C0], R = tree.C_right[(size_t)C0];
                if( L < 0 ) { continue; }
                for( Int i = tree.C_begin[(size_t)L]; i < tree.C_end[(size_t)L]; ++i ) { side[(size_t)i] |= 1; }
                for( Int i = tree.C_begin[(size_t)R]; i < tree.C_end[(size_t)R]; ++i ) { side[(size_t)i] |= 2; }
            }
            for( Int C0 : queue0 )
            {
                const Int L = tree.C_left[(size_t)C0], R = tree.C_right[(size_t)C0];
                if( L < 0 ) { continue; }
                const Int ib = Vb[(size_t)C0], ie = Ve[(size_t)C0];
                Int ctr[3] = {0,0,0};
                for( Int i = ib; i < ie; ++i )
                {
                    const Int v = perm0[(size_t)i];
                    int t = 0;
                    for( LInt k = vp[(size_t)v]; k < vp[(size_t)v + 1]; ++k ) { t |= side[(size_t)vprim[(size_t)k]]; }
                    type[(size_t)i] = (signed char)( t - 1 );
                    if( t > 0 ) { ++ctr[t - 1]; }
                }
                Int pos[4];
                pos[0] = ib; pos[1] = pos[0] + ctr[0]; pos[2] = pos[1] + ctr[1]; pos[3] = pos[2] + ctr[2];
                if( ctr[0] > 0 && ctr[1] > 0 && pos[3] == ie )
                {
                    Vb[(size_t)L] = pos[0]; Ve[(size_t)L] = pos[1];
                    Vb[(size_t)R] = pos[1]; Ve[(size_t)R] = pos[2];
                    queue1.push_back( L );
                    queue1.push_back( R );
                    for( Int i = ib; i < ie; ++i )
                    {
                        const int t = type[(size_t)i];
                        if( t >= 0 ) { perm1[(size_t)pos[t]++] = perm0[(size_t)i]; }
                    }
                }
            }
            std::swap( queue0, queue1 );
            std::swap( perm0, perm1 );
        }
        return perm0;
    }

} // namespace Repulsor
