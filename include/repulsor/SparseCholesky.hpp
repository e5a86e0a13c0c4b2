#pragma once
// Supernodal multifrontal Cholesky factorisation H = L L^T of a symmetric positive definite sparse
// matrix (the H^1 metric), with a fill-reducing ordering given by the caller. It replaces the
// Tensors library's Sparse::CholeskyDecomposition used by the reference's H1Solver
// (src/SimplicialMesh/H1Metric.hpp:208-257), which is absent from the reference checkout.
//
// Algorithm (textbook, e.g. Davis, "Direct Methods for Sparse Linear Systems", 2006):
//   elimination tree (Liu) -> postorder -> column counts by row-subtree traversal ->
//   fundamental supernodes, amalgamated along chains (a column whose only child is the previous
//   column joins its supernode, explicit zeros where the structures differ) -> supernode row sets by
//   merging the children's update rows -> multifrontal numeric factorisation (dense partial
//   Cholesky of each front, extend-add of the children's update matrices).
// For the triangular solves on the GPU every supernode keeps the inverse of its dense diagonal
// block, so both sweeps are matrix-vector products per supernode with no intra-block recurrence.

#include "Common.hpp"

#include <atomic>

namespace Repulsor
{
    struct SupernodalFactor
    {
        Int n = 0;
        std::vector<Int> perm;        // new -> old
        std::vector<Int> iperm;       // old -> new
        // supernodes (postorder; children before parents)
        Int ns = 0;
        std::vector<Int>  sn_begin;   // first column (new numbering), size ns+1
        std::vector<LInt> row_ptr;    // ns+1: offsets into rows
        std::vector<Int>  rows;       // row sets; the first w entries are the supernode's own columns
        std::vector<LInt> val_ptr;    // ns+1: offsets into vals (column-major r x w blocks)
        std::vector<double> vals;     // L values; the w x w diagonal block holds inv(L_ss) (lower)
        std::vector<Int>  sn_parent;  // -1 for roots
        std::vector<Int>  sn_level;   // height: leaves 0
        Int levels = 0;
        LInt nnz_L = 0;
        double flops = 0;

        Int Width( Int s ) const { return sn_begin[(size_t)s + 1] - sn_begin[(size_t)s]; }
        Int RowCount( Int s ) const { return (Int)( row_ptr[(size_t)s + 1] - row_ptr[(size_t)s] ); }

        // Host solve H x = b for 3 right-hand sides (row-major n x 3), used by the CPU tests.
        void Solve( const double * b, double * x ) const
        {
            std::vector<double> y( (size_t)n * 3 );
            for( Int i = 0; i < n; ++i ) { for( int k = 0; k < 3; ++k ) { y[(size_t)3 * i + k] = b[(size_t)3 * perm[(size_t)i] + k]; } }
            // forward: y_s = inv(L_ss) y_s ; y_rows -= L_off y_s
            for( Int s = 0; s < ns; ++s )
            {
                const Int f = sn_begin[(size_t)s], w = Width(s), r = RowCount(s);
                const double * Ls = &vals[(size_t)val_ptr[(size_t)s]];
                const Int * R = &rows[(size_t)row_ptr[(size_t)s]];
                double t[3 * 256];
                std::vector<double> tv;
                double * tt = t;
                if( w > 256 ) { tv.resize( (size_t)3 * w ); tt = tv.data(); }
                for( Int a = 0; a < w; ++a )
                {
                    double acc[3] = {0,0,0};
                    for( Int c = 0; c <= a; ++c ) { const double l = Ls[(size_t)c * r + a]; for( int k = 0; k < 3; ++k ) { acc[k] += l * y[(size_t)3 * (f + c) + k]; } }
                    for( int k = 0; k < 3; ++k ) { tt[3 * a + k] = acc[k]; }
                }
                for( Int a = 0; a < w; ++a ) { for( int k = 0; k < 3; ++k ) { y[(size_t)3 * (f + a) + k] = tt[3 * a + k]; } }
                for( Int q = w; q < r; ++q )
                {
                    double acc[3] = {0,0,0};
                    for( Int c = 0; c < w; ++c ) { const double l = Ls[(size_t)c * r + q]; for( int k = 0; k < 3; ++k ) { acc[k] += l * y[(size_t)3 * (f + c) + k]; } }
                    for( int k = 0; k < 3; ++k ) { y[(size_t)3 * R[q] + k] -= acc[k]; }
                }
            }
            // backward: x_s = inv(L_ss)^T ( y_s - L_off^T x_rows )
            for( Int s = ns; s-- > 0; )
            {
                const Int f = sn_begin[(size_t)s], w = Width(s), r = RowCount(s);
                const double * Ls = &vals[(size_t)val_ptr[(size_t)s]];
                const Int * R = &rows[(size_t)row_ptr[(size_t)s]];
                std::vector<double> t( (size_t)3 * w );
                for( Int c = 0; c < w; ++c )
                {
                    double acc[3] = { y[(size_t)3 * (f + c)], y[(size_t)3 * (f + c) + 1], y[(size_t)3 * (f + c) + 2] };
                    for( Int q = w; q < r; ++q ) { const double l = Ls[(size_t)c * r + q]; for( int k = 0; k < 3; ++k ) { acc[k] -= l * y[(size_t)3 * R[q] + k]; } }
                    for( int k = 0; k < 3; ++k ) { t[(size_t)3 * c + k] = acc[k]; }
                }
                for( Int c = 0; c < w; ++c )
                {
                    double acc[3] = {0,0,0};
                    for( Int a = c; a < w; ++a ) { const double l = Ls[(size_t)c * r + a]; for( int k = 0; k < 3; ++k ) { acc[k] += l * t[(size_t)3 * a + k]; } }
                    for( int k = 0; k < 3; ++k ) { y[(size_t)3 * (f + c) + k] = acc[k]; }
                }
            }
            for( Int i = 0; i < n; ++i ) { for( int k = 0; k < 3; ++k ) { x[(size_t)3 * perm[(size_t)i] + k] = y[(size_t)3 * i + k]; } }
        }
    };

    // A: symmetric CSR (full pattern, with diagonal), perm0: fill-reducing ordering new -> old.
    // Numeric phase (reference: Solver_T::NumericFactorization, src/SimplicialMesh/H1Metric.hpp:256): the
    // values of F for the matrix (rp, ci, av) on the structure F already holds (ordering, supernodes, row
    // sets). A matrix of the same pattern with new values (the H^1 metric of a moved mesh) is refactorised
    // without redoing the symbolic analysis; the result equals a fresh SparseCholeskyFactor bit for bit.
    inline void SparseCholeskyNumeric( SupernodalFactor & F, const std::vector<LInt> & rp, const std::vector<Int> & ci,
                                       const std::vector<double> & av, const int threads = 1 )
    {
        ScopedTimer timer("SparseCholesky::Numeric");
        const Int n = F.n, ns = F.ns;
        const std::vector<Int> & sb = F.sn_begin;
        std::vector<std::vector<Int>> kids( (size_t)ns );
        for( Int s = 0; s < ns; ++s ) { if( F.sn_parent[(size_t)s] >= 0 ) { kids[(size_t)F.sn_parent[(size_t)s]].push_back( s ); } }
        if( F.vals.size() != (size_t)F.val_ptr.back() ) { F.vals.assign( (size_t)F.val_ptr.back(), 0.0 ); }
        // ------------------------------------------------ numeric multifrontal factorisation
        // Fronts are processed as soon as all their children are done (a ready queue shared by the
        // worker threads); update matrices (lower triangles, column-major dense) are kept until the
        // parent consumes them. Inside a front the first w columns are factorised by blocked
        // right-looking Cholesky: panel of NB columns, then the trailing update with the panel
        // copied row-major so that every entry update is one contiguous dot product.
        std::vector<std::vector<double>> upd( (size_t)ns );
        std::vector<std::atomic<Int>> pending( (size_t)ns );
        for( Int s = 0; s < ns; ++s ) { pending[(size_t)s].store( (Int)kids[(size_t)s].size() ); }
        std::mutex qm;
        std::vector<Int> ready;
        for( Int s = 0; s < ns; ++s ) { if( kids[(size_t)s].empty() ) { ready.push_back( s ); } }
        std::reverse( ready.begin(), ready.end() );
        std::atomic<Int> done{0};
        std::atomic<bool> failed{false};
        std::vector<double> flops_t( (size_t)std::max( 1, threads ), 0.0 );
        constexpr Int NB = 32;

        auto front = [&]( const Int s, std::vector<Int> & pos, double & fl )
        {
            const Int * R = F.rows.data() + F.row_ptr[(size_t)s];
            const Int r = F.RowCount( s );
            const Int f = sb[(size_t)s], l = sb[(size_t)s + 1];
            const Int w = l - f;
            for( Int q = 0; q < r; ++q ) { pos[(size_t)R[(size_t)q]] = q; }
            std::vector<double> Fm( (size_t)r * r, 0.0 );   // column-major, lower part used
            for( Int j = f; j < l; ++j )
            {
                const Int oj = F.perm[(size_t)j];
                const Int cj = j - f;
                for( LInt k = rp[(size_t)oj]; k < rp[(size_t)oj + 1]; ++k )
                {
                    const Int i = F.iperm[(size_t)ci[(size_t)k]];
                    if( i >= j ) { Fm[(size_t)cj * r + pos[(size_t)i]] += av[(size_t)k]; }
                }
            }
            for( Int c : kids[(size_t)s] )
            {
                const Int * Rc = F.rows.data() + F.row_ptr[(size_t)c];
                const Int wc = sb[(size_t)c + 1] - sb[(size_t)c];
                const Int uc = F.RowCount( c ) - wc;
                const std::vector<double> & U = upd[(size_t)c];
                for( Int b = 0; b < uc; ++b )
                {
                    const Int pb = pos[(size_t)Rc[(size_t)(wc + b)]];
                    double * col = &Fm[(size_t)pb * r];
                    const double * uu = &U[(size_t)b * uc];
                    for( Int a = b; a < uc; ++a ) { col[pos[(size_t)Rc[(size_t)(wc + a)]]] += uu[a]; }
                }
                std::vector<double>().swap( upd[(size_t)c] );
            }
            std::vector<double> pan;   // row-major panel copy r x NB
            for( Int kb = 0; kb < w; kb += NB )
            {
                const Int nb = std::min( NB, w - kb );
                for( Int k = kb; k < kb + nb; ++k )
                {
                    double * ck = &Fm[(size_t)k * r];
                    const double d = ck[k];
                    if( !(d > 0) ) { failed = true; return; }
                    const double ldk = std::sqrt( d );
                    ck[k] = ldk;
                    const double inv = 1.0 / ldk;
                    for( Int i = k + 1; i < r; ++i ) { ck[i] *= inv; }
                    for( Int jj = k + 1; jj < kb + nb; ++jj )
                    {
                        const double ljk = ck[jj];
                        double * cj = &Fm[(size_t)jj * r];
                        for( Int i = jj; i < r; ++i ) { cj[i] -= ck[i] * ljk; }
                    }
                }
                const Int t0 = kb + nb;
                if( t0 >= r ) { continue; }
                // the panel copied k-major (row i of the panel at [k][i - t0], padded by 4 zero rows), so that
                // the trailing update runs 4 x 4 register blocks of entries: each entry's dot product over
                // the panel still sums k = 0, 1, ... into its own accumulator (the same sums as one entry at
                // a time), but 16 independent sums proceed together instead of one dependent chain
                const Int np = r - t0 + 4;
                pan.assign( (size_t)np * NB, 0.0 );
                for( Int k = 0; k < nb; ++k )
                {
                    const double * ck = &Fm[(size_t)(kb + k) * r];
                    double * pk = &pan[(size_t)k * np];
                    for( Int i = t0; i < r; ++i ) { pk[i - t0] = ck[i]; }
                }
                auto upd_cols = [&]( Int jb, Int je )
                {
                    for( Int j0 = jb; j0 < je; j0 += 4 )
                    {
                        const Int jn = std::min<Int>( 4, je - j0 );
                        for( Int i0 = j0; i0 < r; i0 += 4 )
                        {
                            double acc[4][4] = { { 0.0, 0.0, 0.0, 0.0 }, { 0.0, 0.0, 0.0, 0.0 }, { 0.0, 0.0, 0.0, 0.0 }, { 0.0, 0.0, 0.0, 0.0 } };
                            for( Int k = 0; k < NB; ++k )
                            {
                                const double * pk = &pan[(size_t)k * np];
                                const double a0 = pk[i0 - t0], a1 = pk[i0 - t0 + 1], a2 = pk[i0 - t0 + 2], a3 = pk[i0 - t0 + 3];
                                const double b0 = pk[j0 - t0], b1 = pk[j0 - t0 + 1], b2 = pk[j0 - t0 + 2], b3 = pk[j0 - t0 + 3];
                                acc[0][0] += a0 * b0; acc[0][1] += a0 * b1; acc[0][2] += a0 * b2; acc[0][3] += a0 * b3;
                                acc[1][0] += a1 * b0; acc[1][1] += a1 * b1; acc[1][2] += a1 * b2; acc[1][3] += a1 * b3;
                                acc[2][0] += a2 * b0; acc[2][1] += a2 * b1; acc[2][2] += a2 * b2; acc[2][3] += a2 * b3;
                                acc[3][0] += a3 * b0; acc[3][1] += a3 * b1; acc[3][2] += a3 * b2; acc[3][3] += a3 * b3;
                            }
                            const Int in = std::min<Int>( 4, r - i0 );
                            for( Int v = 0; v < jn; ++v )
                            {
                                double * cj = &Fm[(size_t)( j0 + v ) * r];
                                for( Int u = 0; u < in; ++u ) { if( i0 + u >= j0 + v ) { cj[i0 + u] -= acc[u][v]; } }
                            }
                        }
                    }
                };
                const Int ncols = r - t0;
                if( (double)ncols * ncols * NB > 4e7 && threads > 1 )
                {
                    // split the columns of the trailing triangle into chunks of equal work
                    const int tc = threads;
                    std::vector<Int> cut( (size_t)tc + 1, r );
                    cut[0] = t0;
                    const double tot = 0.5 * (double)ncols * ncols;
                    for( int q = 1; q < tc; ++q )
                    {
                        const double target = tot * (1.0 - (double)q / tc);   // work of columns >= cut is ~ (r-cut)^2/2
                        cut[(size_t)q] = std::max( cut[(size_t)q - 1], (Int)( r - std::sqrt( 2.0 * target ) ) );
                    }
                    ParallelDo( tc, [&]( int q ) { upd_cols( cut[(size_t)q], cut[(size_t)q + 1] ); } );
                }
                else
                {
                    upd_cols( t0, r );
                }
                fl += 2.0 * (double)nb * ncols * ncols * 0.5 + (double)nb * nb * r;
            }
            double * Ls = &F.vals[(size_t)F.val_ptr[(size_t)s]];
            for( Int c = 0; c < w; ++c ) { for( Int q = 0; q < r; ++q ) { Ls[(size_t)c * r + q] = (q >= c) ? Fm[(size_t)c * r + q] : 0.0; } }
            {
                // column-major inverse of the lower diagonal block, column c by forward substitution on
                // e_c in column (axpy) order; wide blocks split their columns over the threads in chunks
                // of equal work (column c costs about (w - c)^2 / 2)
                std::vector<double> inv( (size_t)w * w, 0.0 );
                auto inv_cols = [&]( Int cb, Int ce )
                {
                    for( Int c = cb; c < ce; ++c )
                    {
                        double * x = &inv[(size_t)c * w];
                        x[c] = 1.0;
                        for( Int b = c; b < w; ++b )
                        {
                            const double * lb = &Ls[(size_t)b * r];
                            const double xb = x[b] / lb[b];
                            x[b] = xb;
                            for( Int a = b + 1; a < w; ++a ) { x[a] -= lb[a] * xb; }
                        }
                    }
                };
                if( w > 256 && threads > 1 )
                {
                    const int tc = threads;
                    std::vector<Int> cut( (size_t)tc + 1, w );
                    cut[0] = 0;
                    const double tot = (double)w * w * w / 6.0;
                    for( int q = 1; q < tc; ++q )
                    {
                        const double target = tot * ( 1.0 - (double)q / tc );   // work of the columns >= cut is ~ (w - cut)^3 / 6
                        cut[(size_t)q] = std::max( cut[(size_t)q - 1], (Int)( w - std::cbrt( 6.0 * target ) ) );
                    }
                    ParallelDo( tc, [&]( int q ) { inv_cols( cut[(size_t)q], cut[(size_t)q + 1] ); } );
                }
                else { inv_cols( 0, w ); }
                for( Int c = 0; c < w; ++c ) { for( Int a = 0; a < w; ++a ) { Ls[(size_t)c * r + a] = inv[(size_t)c * w + a]; } }
            }
            const Int u = r - w;
            if( u > 0 && F.sn_parent[(size_t)s] >= 0 )
            {
                std::vector<double> & U = upd[(size_t)s];
                U.assign( (size_t)u * u, 0.0 );
                for( Int b = 0; b < u; ++b ) { for( Int a = b; a < u; ++a ) { U[(size_t)b * u + a] = Fm[(size_t)(w + b) * r + (w + a)]; } }
            }
            for( Int q = 0; q < r; ++q ) { pos[(size_t)R[(size_t)q]] = -1; }
        };

        const int tc = std::max( 1, threads );
        ParallelDo( tc, [&]( int t )
        {
            std::vector<Int> pos( (size_t)n, -1 );
            while( done.load() < ns && !failed.load() )
            {
                Int s = -1;
                {
                    std::lock_guard<std::mutex> lk( qm );
                    if( !ready.empty() ) { s = ready.back(); ready.pop_back(); }
                }
                if( s < 0 ) { std::this_thread::yield(); continue; }
                front( s, pos, flops_t[(size_t)t] );
                const Int p = F.sn_parent[(size_t)s];
                if( p >= 0 && pending[(size_t)p].fetch_sub( 1 ) == 1 )
                {
                    std::lock_guard<std::mutex> lk( qm );
                    ready.push_back( p );
                }
                done.fetch_add( 1 );
            }
        });
        if( failed.load() ) { throw std::runtime_error( "SparseCholesky: matrix is not positive definite" ); }
        double flops = 0;
        for( double x : flops_t ) { flops += x; }
        F.flops = flops;
    }

    inline SupernodalFactor SparseCholeskyFactor( const Int n, const std::vector<LInt> & rp, const std::vector<Int> & ci,
                                                  const std::vector<double> & av, const std::vector<Int> & perm0,
                                                  const int threads = 1, const Int relax_cols = 16, const Int max_width = 2048,
                                                  const bool chain_merge = true, const Int bottom_levels = 4 )
    {
        ScopedTimer timer("SparseCholesky::Factor");
        SupernodalFactor F;
        F.n = n;
        // ------------------------------------------------ elimination tree of P A P^T (perm0)
        std::vector<Int> ip0( (size_t)n );
        for( Int i = 0; i < n; ++i ) { ip0[(size_t)perm0[(size_t)i]] = i; }
        auto etree = [&]( const std::vector<Int> & perm, const std::vector<Int> & iperm )
        {
            std::vector<Int> parent( (size_t)n, -1 ), anc( (size_t)n, -1 );
            for( Int j = 0; j < n; ++j )
            {
                const Int oj = perm[(size_t)j];
                for( LInt k = rp[(size_t)oj]; k < rp[(size_t)oj + 1]; ++k )
                {
                    Int i = iperm[(size_t)ci[(size_t)k]];
                    if( i >= j ) { continue; }
                    while( i != -1 && i < j )
                    {
                        const Int next = anc[(size_t)i];
                        anc[(size_t)i] = j;
                        if( next == -1 ) { parent[(size_t)i] = j; }
                        i = next;
                    }
                }
            }
            return parent;
        };
        std::vector<Int> par0 = etree( perm0, ip0 );
        // ------------------------------------------------ postorder
        std::vector<Int> head( (size_t)n, -1 ), next( (size_t)n, -1 );
        for( Int j = n; j-- > 0; ) { const Int p = par0[(size_t)j]; if( p >= 0 ) { next[(size_t)j] = head[(size_t)p]; head[(size_t)p] = j; } }
        std::vector<Int> post; post.reserve( (size_t)n );
        {
            std::vector<Int> stack;
            for( Int j = 0; j < n; ++j )
            {
                if( par0[(size_t)j] != -1 ) { continue; }
                stack.push_back( j );
                while( !stack.empty() )
                {
                    const Int p = stack.back();
                    const Int c = head[(size_t)p];
                    if( c == -1 ) { stack.pop_back(); post.push_back( p ); }
                    else { head[(size_t)p] = next[(size_t)c]; stack.push_back( c ); }
                }
            }
        }
        F.perm.resize( (size_t)n ); F.iperm.resize( (size_t)n );
        for( Int k = 0; k < n; ++k ) { F.perm[(size_t)k] = perm0[(size_t)post[(size_t)k]]; }
        for( Int k = 0; k < n; ++k ) { F.iperm[(size_t)F.perm[(size_t)k]] = k; }
        std::vector<Int> parent = etree( F.perm, F.iperm );
        // ------------------------------------------------ column counts (row-subtree traversal)
        std::vector<Int> cc( (size_t)n, 1 ), mark( (size_t)n, -1 );
        for( Int i = 0; i < n; ++i )
        {
            mark[(size_t)i] = i;
            const Int oi = F.perm[(size_t)i];
            for( LInt k = rp[(size_t)oi]; k < rp[(size_t)oi + 1]; ++k )
            {
                Int j = F.iperm[(size_t)ci[(size_t)k]];
                while( j < i && mark[(size_t)j] != i )
                {
                    ++cc[(size_t)j];
                    mark[(size_t)j] = i;
                    j = parent[(size_t)j];
                }
            }
        }
        // ------------------------------------------------ supernodes
        std::vector<Int> nchild( (size_t)n, 0 );
        for( Int j = 0; j < n; ++j ) { if( parent[(size_t)j] >= 0 ) { ++nchild[(size_t)parent[(size_t)j]]; } }
        std::vector<Int> sb; sb.push_back( 0 );
        for( Int j = 1; j < n; ++j )
        {
            const bool fundamental = parent[(size_t)(j - 1)] == j && cc[(size_t)(j - 1)] == cc[(size_t)j] + 1 && nchild[(size_t)j] == 1;
            const Int width = j - sb.back();
            // relaxed: also merge a short chain link when the column count grows by at most 2
            const bool relaxed = parent[(size_t)(j - 1)] == j && nchild[(size_t)j] == 1 && width < relax_cols && cc[(size_t)(j - 1)] <= cc[(size_t)j] + 3;
            // chain: a column whose only child is the previous column joins it whatever the fill; a chain
            // of supernodes is a chain of dependent elimination levels, one launch per level and sweep
            // on the GPU, and merging them costs about 1 % more stored values on nested-dissection H^1
            // factors while cutting the levels by a third
            const bool chain = chain_merge && parent[(size_t)(j - 1)] == j && nchild[(size_t)j] == 1;
            if( !( fundamental || relaxed || chain ) || width >= max_width ) { sb.push_back( j ); }
        }
        sb.push_back( n );
        if( bottom_levels > 0 )
        {
            // the maximal subtrees of supernodes at heights below bottom_levels become one supernode each
            // (a subtree is a contiguous column range in postorder; explicit zeros where the structures
            // differ): the tree loses bottom_levels - 1 levels at its widest, lowest part. Measured on the
            // GPU solves at 1.31M triangles: 2 (3 % more stored values) 0.749 -> 0.726 ms per solve; with
            // the batched loads and per-sweep lane-group choice of the GPU solves, 3 (21 -> 20 levels, 7 %
            // more stored values) 0.605 -> 0.582 ms; with the single-entry front pulls, 4 (19 levels, 68.4M
            // instead of 60.3M stored values) 15 us less per solve, while 5 and 6 (18 and 17 levels, 85M and
            // 117M values, lowest supernodes up to 182 and 304 rows) doubled the solve time
            const Int m = (Int)sb.size() - 1;
            std::vector<Int> sof( (size_t)n );
            for( Int t = 0; t < m; ++t ) { for( Int j = sb[(size_t)t]; j < sb[(size_t)t + 1]; ++j ) { sof[(size_t)j] = t; } }
            std::vector<Int> par( (size_t)m, -1 ), ht( (size_t)m, 0 ), sz( (size_t)m, 1 );
            std::vector<LInt> wsum( (size_t)m, 0 );
            for( Int t = 0; t < m; ++t )
            {
                const Int p = parent[(size_t)( sb[(size_t)t + 1] - 1 )];
                par[(size_t)t] = p >= 0 ? sof[(size_t)p] : -1;
                wsum[(size_t)t] += sb[(size_t)t + 1] - sb[(size_t)t];
                if( par[(size_t)t] >= 0 )
                {
                    ht[(size_t)par[(size_t)t]] = std::max( ht[(size_t)par[(size_t)t]], ht[(size_t)t] + 1 );
                    sz[(size_t)par[(size_t)t]] += sz[(size_t)t];
                    wsum[(size_t)par[(size_t)t]] += wsum[(size_t)t];
                }
            }
            // rootof: the root of the maximal low subtree holding t (parents come after children)
            std::vector<Int> rootof( (size_t)m, -1 );
            for( Int t = m - 1; t >= 0; --t )
            {
                if( ht[(size_t)t] >= bottom_levels ) { continue; }
                const Int p = par[(size_t)t];
                rootof[(size_t)t] = ( p < 0 || ht[(size_t)p] >= bottom_levels ) ? t : rootof[(size_t)p];
            }
            std::vector<Int> nb;
            for( Int t = 0; t < m; ++t )
            {
                const Int r0 = rootof[(size_t)t];
                const bool merged = r0 >= 0 && wsum[(size_t)r0] <= max_width;
                if( !merged ) { nb.push_back( sb[(size_t)t] ); }
                else if( r0 == t ) { nb.push_back( sb[(size_t)( t - sz[(size_t)t] + 1 )] ); }
            }
            nb.push_back( n );
            sb.swap( nb );
        }
        const Int ns = (Int)sb.size() - 1;
        F.ns = ns;
        F.sn_begin = sb;
        std::vector<Int> sn_of( (size_t)n );
        for( Int s = 0; s < ns; ++s ) { for( Int j = sb[(size_t)s]; j < sb[(size_t)s + 1]; ++j ) { sn_of[(size_t)j] = s; } }
        F.sn_parent.assign( (size_t)ns, -1 );
        for( Int s = 0; s < ns; ++s )
        {
            const Int last = sb[(size_t)s + 1] - 1;
            const Int p = parent[(size_t)last];
            F.sn_parent[(size_t)s] = (p >= 0) ? sn_of[(size_t)p] : -1;
        }
        // children lists
        std::vector<std::vector<Int>> kids( (size_t)ns );
        for( Int s = 0; s < ns; ++s ) { if( F.sn_parent[(size_t)s] >= 0 ) { kids[(size_t)F.sn_parent[(size_t)s]].push_back( s ); } }
        // row sets: own columns, then sorted union of A's rows below and the children's update rows
        F.row_ptr.assign( (size_t)ns + 1, 0 );
        std::vector<std::vector<Int>> RS( (size_t)ns );
        {
            std::vector<Int> flag( (size_t)n, -1 );
            for( Int s = 0; s < ns; ++s )
            {
                const Int f = sb[(size_t)s], l = sb[(size_t)s + 1];
                std::vector<Int> below;
                for( Int j = f; j < l; ++j )
                {
                    const Int oj = F.perm[(size_t)j];
                    for( LInt k = rp[(size_t)oj]; k < rp[(size_t)oj + 1]; ++k )
                    {
                        const Int i = F.iperm[(size_t)ci[(size_t)k]];
                        if( i >= l && flag[(size_t)i] != s ) { flag[(size_t)i] = s; below.push_back( i ); }
                    }
                }
                for( Int c : kids[(size_t)s] )
                {
                    const std::vector<Int> & Rc = RS[(size_t)c];
                    const Int wc = sb[(size_t)c + 1] - sb[(size_t)c];
                    for( size_t q = (size_t)wc; q < Rc.size(); ++q )
                    {
                        const Int i = Rc[q];
                        if( i >= l && flag[(size_t)i] != s ) { flag[(size_t)i] = s; below.push_back( i ); }
                    }
                }
                std::sort( below.begin(), below.end() );
                std::vector<Int> & R = RS[(size_t)s];
                R.reserve( (size_t)(l - f) + below.size() );
                for( Int j = f; j < l; ++j ) { R.push_back( j ); }
                R.insert( R.end(), below.begin(), below.end() );
                F.row_ptr[(size_t)s + 1] = F.row_ptr[(size_t)s] + (LInt)R.size();
            }
        }
        F.rows.resize( (size_t)F.row_ptr.back() );
        F.val_ptr.assign( (size_t)ns + 1, 0 );
        for( Int s = 0; s < ns; ++s )
        {
            std::copy( RS[(size_t)s].begin(), RS[(size_t)s].end(), F.rows.begin() + F.row_ptr[(size_t)s] );
            const LInt r = (LInt)RS[(size_t)s].size(), w = sb[(size_t)s + 1] - sb[(size_t)s];
            F.val_ptr[(size_t)s + 1] = F.val_ptr[(size_t)s] + r * w;
            F.nnz_L += w * (w + 1) / 2 + (r - w) * w;
        }
        // levels (height in the supernode tree)
        F.sn_level.assign( (size_t)ns, 0 );
        for( Int s = 0; s < ns; ++s )
        {
            const Int p = F.sn_parent[(size_t)s];
            if( p >= 0 ) { F.sn_level[(size_t)p] = std::max( F.sn_level[(size_t)p], F.sn_level[(size_t)s] + 1 ); }
            F.levels = std::max( F.levels, F.sn_level[(size_t)s] + 1 );
        }
        SparseCholeskyNumeric( F, rp, ci, av, threads );
        return F;
    }

} // namespace Repulsor
