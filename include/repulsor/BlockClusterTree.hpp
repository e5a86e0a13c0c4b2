#pragma once
// Symmetric block cluster tree (S == T) over a ClusterTree.
//
// Restates the admissibility traversal of the reference:
//   BlockClusterTree::ComputeBlocks            src/BlockClusterTree.hpp:234-330
//   ClusterTreePairTraversor::Traverse_DepthFirst src/ClusterTreePairTraversor/Traverse_DepthFirst.hpp:3-267
//   BlockSplit_Kernel::{AdmissableQ,ComputeLeaf,ComputeAdmissable} src/BlockClusterTree/BlockSplit_Kernel.hpp:164-251
//   AABB MAC: max(r_S^2, r_T^2) < theta_far^2 * dist(AABB_S, AABB_T)^2   src/GJK.hpp:1175-1183, src/BoundingVolumes/AABB.hpp:407-430
// Leaf-leaf primitive pairs (i < j) are Near if the primitives share a vertex or pass the primitive MAC
// with theta_near; otherwise Inter (intersecting) or VeryNear.
//
// Storage. The near field is kept as the list of non-admissible leaf-leaf blocks plus the (rare)
// pairs inside them that are not Near; the primitive-level CSR of the reference (BinaryMatrixCSR,
// full symmetric, sorted columns, no diagonal) is generated from it on demand. The far field is a
// full symmetric CSR over clusters.

#include "ClusterTree.hpp"
#include "PrimitiveDistance.hpp"

namespace Repulsor
{
    struct SymmetricCSR
    {
        Int rows = 0;
        std::vector<LInt> rp;    // rows+1
        std::vector<Int>  ci;    // nnz
        std::vector<LInt> diag;  // first position with ci >= row (start of the strict upper part)

        LInt NonzeroCount() const { return rp.empty() ? 0 : rp.back(); }

        // Builds the full symmetric pattern from upper pairs (i < j); duplicates are removed.
        static SymmetricCSR FromUpperPairs( const Int n, std::vector<std::pair<Int,Int>> pairs, const int threads )
        {
            (void)threads;
            SymmetricCSR A;
            A.rows = n;
            std::sort( pairs.begin(), pairs.end() );
            pairs.erase( std::unique( pairs.begin(), pairs.end() ), pairs.end() );
            std::vector<LInt> cnt( (size_t)n + 1, 0 );
            for( auto & p : pairs ) { ++cnt[(size_t)p.first + 1]; ++cnt[(size_t)p.second + 1]; }
            for( Int i = 0; i < n; ++i ) { cnt[(size_t)i + 1] += cnt[(size_t)i]; }
            A.rp = cnt;
            A.ci.resize( (size_t)A.rp.back() );
            std::vector<LInt> pos( A.rp.begin(), A.rp.end() - 1 );
            // Lower part first (columns < row) in increasing column order, then the upper part.
            // Pairs are sorted by (first,second): iterating them in order visits, for a fixed row r,
            // the lower entries (r as second) by increasing first, and the upper entries (r as first)
            // by increasing second.
            for( auto & p : pairs ) { A.ci[(size_t)pos[(size_t)p.second]++] = p.first; }
            A.diag.resize( (size_t)n );
            for( Int i = 0; i < n; ++i ) { A.diag[(size_t)i] = pos[(size_t)i]; }
            for( auto & p : pairs ) { A.ci[(size_t)pos[(size_t)p.first]++] = p.second; }
            return A;
        }
    };

    // Near-field work description for the GPU: per leaf I ("row leaf") the concatenated primitive
    // ranges of all leaves J that share a non-admissible block with I (both orientations, J sorted),
    // with a bit mask per column: bit ii set <=> (begin_I + ii, column primitive) is a Near pair.
    struct NearTiles
    {
        Int leaf_count = 0;
        std::vector<Int>      leaf_begin, leaf_end;   // primitive ranges of the leaves (tree order)
        std::vector<LInt>     col_ptr;                // leaf_count + 1
        std::vector<Int>      col_prim;               // column primitive index
        std::vector<uint32_t> col_mask;               // row bits
        std::vector<LInt>     val_ptr;                // leaf_count + 1: offsets of n_I * M_I value slots
        LInt masked_pairs = 0;                        // Near pairs counted with multiplicity 2
    };

    // Block cluster tree lists built on the device (rp_build_topology).
    struct BuiltBlockClusterTree
    {
        std::vector<LInt> far_rp;
        std::vector<Int>  far_ci;
        std::vector<std::pair<Int,Int>> leaf_blocks, verynear_pairs, inter_pairs;   // sorted, upper
        std::vector<LInt> col_ptr;
        std::vector<Int>  col_prim;
        std::vector<uint32_t> col_mask;
    };

    template<int DOM_DIM>
    class BlockClusterTree
    {
    public:
        using Tree_T = ClusterTree<DOM_DIM>;
        static constexpr int SIZE = DOM_DIM + 1;

    protected:
        const Tree_T & S;
        BlockClusterTreeSettings settings;
        double far_theta2, near_theta2, intersection_theta2;
        int threads;

    public:
        SymmetricCSR far;                                     // C x C
        std::vector<std::pair<Int,Int>> leaf_blocks;          // upper (I <= J) leaf indices
        std::vector<std::pair<Int,Int>> verynear_pairs;       // upper primitive pairs inside leaf blocks
        std::vector<std::pair<Int,Int>> inter_pairs;          // upper primitive pairs inside leaf blocks
        LInt near_nnz = 0;                                    // full symmetric count
        NearTiles tiles;

        BlockClusterTree( const Tree_T & S_, const BlockClusterTreeSettings & settings_, const int threads_ )
        :   S( S_ ), settings( settings_ )
        ,   far_theta2 ( settings_.far_field_separation_parameter  * settings_.far_field_separation_parameter  )
        ,   near_theta2( settings_.near_field_separation_parameter * settings_.near_field_separation_parameter )
        // intersection_theta2 is initialised from near_theta2 (BlockSplit_Kernel.hpp:42).
        ,   intersection_theta2( near_theta2 )
        ,   threads( std::max( 1, threads_ ) )
        {
            if( S.PrimitiveCount() > 0 ) { ComputeBlocks(); }
        }

        // Takes over the lists of a device build (same settings); derives the counts and the tile value
        // offsets as ComputeBlocks and BuildTiles do.
        BlockClusterTree( const Tree_T & S_, const BlockClusterTreeSettings & settings_, const int threads_, BuiltBlockClusterTree && B )
        :   S( S_ ), settings( settings_ )
        ,   far_theta2 ( settings_.far_field_separation_parameter  * settings_.far_field_separation_parameter  )
        ,   near_theta2( settings_.near_field_separation_parameter * settings_.near_field_separation_parameter )
        ,   intersection_theta2( near_theta2 )
        ,   threads( std::max( 1, threads_ ) )
        {
            ScopedTimer timer("BlockClusterTree::Adopt");
            const Int C = S.ClusterCount(), L = S.LeafClusterCount();
            if( (Int)B.far_rp.size() != C + 1 || (Int)B.col_ptr.size() != L + 1 )
            {
                throw std::runtime_error( "BlockClusterTree: inconsistent device-built lists" );
            }
            far.rows = C;
            far.rp = std::move( B.far_rp );
            far.ci = std::move( B.far_ci );
            far.diag.resize( (size_t)C );
            for( Int i = 0; i < C; ++i )
            {
                far.diag[(size_t)i] = std::lower_bound( far.ci.begin() + far.rp[(size_t)i], far.ci.begin() + far.rp[(size_t)i + 1], i ) - far.ci.begin();
            }
            leaf_blocks = std::move( B.leaf_blocks );
            verynear_pairs = std::move( B.verynear_pairs );
            inter_pairs = std::move( B.inter_pairs );
            LInt pairs = 0;
            for( auto & b : leaf_blocks )
            {
                const LInt ni = S.leaf_cluster_ptr[(size_t)b.first + 1]  - S.leaf_cluster_ptr[(size_t)b.first];
                const LInt nj = S.leaf_cluster_ptr[(size_t)b.second + 1] - S.leaf_cluster_ptr[(size_t)b.second];
                pairs += (b.first == b.second) ? (ni * (ni - 1)) / 2 : ni * nj;
            }
            near_nnz = 2 * ( pairs - (LInt)verynear_pairs.size() - (LInt)inter_pairs.size() );
            NearTiles & Tl = tiles;
            Tl.leaf_count = L;
            Tl.leaf_begin.resize( (size_t)L );
            Tl.leaf_end.resize( (size_t)L );
            for( Int l = 0; l < L; ++l ) { Tl.leaf_begin[(size_t)l] = S.leaf_cluster_ptr[(size_t)l]; Tl.leaf_end[(size_t)l] = S.leaf_cluster_ptr[(size_t)l+1]; }
            Tl.col_ptr = std::move( B.col_ptr );
            Tl.col_prim = std::move( B.col_prim );
            Tl.col_mask = std::move( B.col_mask );
            Tl.val_ptr.assign( (size_t)L + 1, 0 );
            Tl.masked_pairs = 0;
            for( Int l = 0; l < L; ++l )
            {
                const LInt m = Tl.col_ptr[(size_t)l + 1] - Tl.col_ptr[(size_t)l];
                Tl.val_ptr[(size_t)l + 1] = Tl.val_ptr[(size_t)l] + m * (Tl.leaf_end[(size_t)l] - Tl.leaf_begin[(size_t)l]);
                for( LInt k = Tl.col_ptr[(size_t)l]; k < Tl.col_ptr[(size_t)l + 1]; ++k ) { Tl.masked_pairs += __builtin_popcount( Tl.col_mask[(size_t)k] ); }
            }
        }

        const Tree_T & GetS() const { return S; }
        const Tree_T & GetT() const { return S; }
        double FarFieldSeparationParameter()  const { return std::sqrt(far_theta2); }
        double NearFieldSeparationParameter() const { return std::sqrt(near_theta2); }
        const BlockClusterTreeSettings & Settings() const { return settings; }

        LInt NearFieldInteractionCount()     const { return near_nnz; }
        LInt FarFieldInteractionCount()      const { return far.NonzeroCount(); }
        LInt VeryNearFieldInteractionCount() const { return 2 * (LInt)verynear_pairs.size(); }
        LInt PrimitiveIntersectionCount()    const { return 2 * (LInt)inter_pairs.size(); }

        std::string Stats() const
        {
            char buf[512];
            std::snprintf( buf, sizeof(buf),
                "BlockClusterTree<%d,3>: T=%d C=%d leaves=%d depth=%d | near nnz=%lld far nnz=%lld very near nnz=%lld inter nnz=%lld leaf blocks=%lld",
                DOM_DIM, S.PrimitiveCount(), S.ClusterCount(), S.LeafClusterCount(), S.TreeDepth(),
                (long long)NearFieldInteractionCount(), (long long)FarFieldInteractionCount(),
                (long long)VeryNearFieldInteractionCount(), (long long)PrimitiveIntersectionCount(),
                (long long)leaf_blocks.size() );
            return std::string(buf);
        }

    protected:
        struct ThreadOut
        {
            std::vector<std::pair<Int,Int>> far, blocks, vf, inter;
        };

        bool AdmissableQ( const Int i, const Int j ) const
        {
            const double * P = &S.C_serialized[ (size_t)Tree_T::C_SIZE * i ];
            const double * Q = &S.C_serialized[ (size_t)Tree_T::C_SIZE * j ];
            double d2 = 0;
            for( int k = 0; k < 3; ++k )
            {
                const double x = std::max( 0.0,
                    std::max( P[1+k] - P[4+k], Q[1+k] - Q[4+k] ) - std::min( P[1+k] + P[4+k], Q[1+k] + Q[4+k] ) );
                d2 += x * x;
            }
            return std::max( P[0], Q[0] ) < far_theta2 * d2;
        }

        bool AdjacentQ( const Int i, const Int j ) const
        {
            const Int * a = &S.pre_cols[ (size_t)SIZE * i ];
            const Int * b = &S.pre_cols[ (size_t)SIZE * j ];
            for( int k = 0; k < SIZE; ++k ) { for( int l = 0; l < SIZE; ++l ) { if( a[k] == b[l] ) { return true; } } }
            return false;
        }

        // Classification of the primitive pair (i,j), i < j (BlockSplit_Kernel::ComputeLeaf).
        // Returns 0 = Near, 1 = VeryNear, 2 = Inter.
        int ClassifyPair( const Int i, const Int j ) const
        {
            if( AdjacentQ( i, j ) ) { return 0; }
            const double * P = &S.P_serialized[ (size_t)Tree_T::P_SIZE * i ];
            const double * Q = &S.P_serialized[ (size_t)Tree_T::P_SIZE * j ];
            const double r2max = std::max( P[0], Q[0] );
            const double r2min = std::min( P[0], Q[0] );
            // Cheap sufficient test: distance of the circumscribed balls around the interior points.
            double c2 = 0;
            for( int k = 0; k < 3; ++k ) { const double d = P[1+k] - Q[1+k]; c2 += d * d; }
            const double gap = std::sqrt(c2) - std::sqrt(P[0]) - std::sqrt(Q[0]);
            if( gap > 0 && r2max < near_theta2 * gap * gap ) { return 0; }
            const double d2 = SimplexSquaredDistance( P + 4, SIZE, Q + 4, SIZE );
            if( r2max < near_theta2 * d2 ) { return 0; }
            // GJK IntersectingQ with TOL^2 = eps^2 min(r^2), eps = sqrt(DBL_EPSILON) (src/GJK.hpp:79,643).
            const double eps2 = std::numeric_limits<double>::epsilon();
            return ( intersection_theta2 * d2 <= eps2 * r2min ) ? 2 : 1;
        }

        void LeafLeaf( const Int ci, const Int cj, ThreadOut & out ) const
        {
            const Int li = S.leaf_cluster_lookup[(size_t)ci];
            const Int lj = S.leaf_cluster_lookup[(size_t)cj];
            out.blocks.emplace_back( std::min(li,lj), std::max(li,lj) );
            const Int ib = S.C_begin[(size_t)ci], ie = S.C_end[(size_t)ci];
            const Int jb = S.C_begin[(size_t)cj], je = S.C_end[(size_t)cj];
            auto visit = [&]( Int a, Int b )
            {
                const int c = ClassifyPair( a, b );
                if( c == 1 ) { out.vf.emplace_back( a, b ); }
                else if( c == 2 ) { out.inter.emplace_back( a, b ); }
            };
            if( ci == cj )
            {
                for( Int ii = ib; ii < ie; ++ii ) { for( Int jj = ii + 1; jj < ie; ++jj ) { visit( ii, jj ); } }
            }
            else
            {
                for( Int ii = ib; ii < ie; ++ii )
                {
                    for( Int jj = jb; jj < je; ++jj ) { if( ii < jj ) { visit( ii, jj ); } else { visit( jj, ii ); } }
                }
            }
        }

        // One step of Traverse_DepthFirst for the cluster pair (i,j); children are handed to push().
        template<typename Push>
        void Step( const Int i, const Int j, Push && push, ThreadOut & out ) const
        {
            const bool admissable = (i != j) && AdmissableQ( i, j );
            if( admissable )
            {
                out.far.emplace_back( std::min(i,j), std::max(i,j) );
                return;
            }
            const Int L_i = S.C_left[(size_t)i], L_j = S.C_left[(size_t)j];
            if( (L_i >= 0) || (L_j >= 0) )
            {
                const Int R_i = S.C_right[(size_t)i], R_j = S.C_right[(size_t)j];
                const double score_i = (L_i >= 0) ? S.C_serialized[(size_t)Tree_T::C_SIZE * i] : 0.0;
                const double score_j = (L_j >= 0) ? S.C_serialized[(size_t)Tree_T::C_SIZE * j] : 0.0;
                if( score_i == score_j && score_i > 0 )
                {
                    if( i == j ) { push( L_i, R_j ); push( R_i, R_j ); push( L_i, L_j ); }
                    else         { push( R_i, R_j ); push( L_i, R_j ); push( R_i, L_j ); push( L_i, L_j ); }
                }
                else if( score_i > score_j ) { push( R_i, j ); push( L_i, j ); }
                else                         { push( i, R_j ); push( i, L_j ); }
            }
            else
            {
                LeafLeaf( i, j, out );
            }
        }

        void ComputeBlocks()
        {
            ScopedTimer timer("BlockClusterTree::ComputeBlocks");
            // Breadth-first expansion into enough independent tasks, then parallel depth-first scans
            // (ClusterTreePairTraversor::Traverse_Parallel, src/ClusterTreePairTraversor.hpp:126-146).
            ThreadOut seed_out;
            std::vector<std::pair<Int,Int>> queue { {0,0} };
            const size_t target = (size_t)16 * (size_t)threads;
            while( !queue.empty() && queue.size() < target )
            {
                std::vector<std::pair<Int,Int>> next;
                for( auto & ij : queue )
                {
                    Step( ij.first, ij.second, [&]( Int a, Int b ) { next.emplace_back(a,b); }, seed_out );
                }
                if( next.empty() ) { queue.clear(); break; }
                queue.swap( next );
            }
            std::vector<ThreadOut> outs( (size_t)threads );
            std::atomic<size_t> counter{0};
            ParallelDo( threads, [&]( int t )
            {
                std::vector<std::pair<Int,Int>> stack;
                while( true )
                {
                    const size_t k = counter.fetch_add(1);
                    if( k >= queue.size() ) { break; }
                    stack.clear();
                    stack.push_back( queue[k] );
                    while( !stack.empty() )
                    {
                        const auto ij = stack.back(); stack.pop_back();
                        Step( ij.first, ij.second, [&]( Int a, Int b ) { stack.emplace_back(a,b); }, outs[(size_t)t] );
                    }
                }
            });
            outs.push_back( std::move(seed_out) );

            std::vector<std::pair<Int,Int>> far_pairs;
            for( auto & o : outs )
            {
                far_pairs.insert( far_pairs.end(), o.far.begin(), o.far.end() );
                leaf_blocks.insert( leaf_blocks.end(), o.blocks.begin(), o.blocks.end() );
                verynear_pairs.insert( verynear_pairs.end(), o.vf.begin(), o.vf.end() );
                inter_pairs.insert( inter_pairs.end(), o.inter.begin(), o.inter.end() );
                o = ThreadOut();
            }
            std::sort( leaf_blocks.begin(), leaf_blocks.end() );
            leaf_blocks.erase( std::unique( leaf_blocks.begin(), leaf_blocks.end() ), leaf_blocks.end() );
            std::sort( verynear_pairs.begin(), verynear_pairs.end() );
            std::sort( inter_pairs.begin(), inter_pairs.end() );
            far = SymmetricCSR::FromUpperPairs( S.ClusterCount(), std::move(far_pairs), threads );

            LInt pairs = 0;
            for( auto & b : leaf_blocks )
            {
                const LInt ni = S.leaf_cluster_ptr[(size_t)b.first + 1]  - S.leaf_cluster_ptr[(size_t)b.first];
                const LInt nj = S.leaf_cluster_ptr[(size_t)b.second + 1] - S.leaf_cluster_ptr[(size_t)b.second];
                pairs += (b.first == b.second) ? (ni * (ni - 1)) / 2 : ni * nj;
            }
            near_nnz = 2 * ( pairs - (LInt)verynear_pairs.size() - (LInt)inter_pairs.size() );
            BuildTiles();
        }

        bool ExcludedQ( const Int i, const Int j ) const
        {
            const auto key = std::make_pair( std::min(i,j), std::max(i,j) );
            return std::binary_search( verynear_pairs.begin(), verynear_pairs.end(), key )
                || std::binary_search( inter_pairs.begin(), inter_pairs.end(), key );
        }

        void BuildTiles()
        {
            ScopedTimer timer("BlockClusterTree::BuildTiles");
            const Int L = S.LeafClusterCount();
            NearTiles & Tl = tiles;
            Tl.leaf_count = L;
            Tl.leaf_begin.resize( (size_t)L );
            Tl.leaf_end.resize( (size_t)L );
            for( Int l = 0; l < L; ++l ) { Tl.leaf_begin[(size_t)l] = S.leaf_cluster_ptr[(size_t)l]; Tl.leaf_end[(size_t)l] = S.leaf_cluster_ptr[(size_t)l+1]; }
            // Full symmetric leaf-block graph.
            std::vector<LInt> cnt( (size_t)L + 1, 0 );
            for( auto & b : leaf_blocks ) { ++cnt[(size_t)b.first + 1]; if( b.first != b.second ) { ++cnt[(size_t)b.second + 1]; } }
            for( Int l = 0; l < L; ++l ) { cnt[(size_t)l + 1] += cnt[(size_t)l]; }
            std::vector<Int> nbr( (size_t)cnt.back() );
            {
                std::vector<LInt> pos( cnt.begin(), cnt.end() - 1 );
                for( auto & b : leaf_blocks ) { nbr[(size_t)pos[(size_t)b.first]++] = b.second; if( b.first != b.second ) { nbr[(size_t)pos[(size_t)b.second]++] = b.first; } }
                ParallelFor( L, threads, [&]( LInt l ) { std::sort( nbr.begin() + cnt[(size_t)l], nbr.begin() + cnt[(size_t)l+1] ); } );
            }
            Tl.col_ptr.assign( (size_t)L + 1, 0 );
            Tl.val_ptr.assign( (size_t)L + 1, 0 );
            for( Int l = 0; l < L; ++l )
            {
                LInt m = 0;
                for( LInt k = cnt[(size_t)l]; k < cnt[(size_t)l+1]; ++k ) { const Int J = nbr[(size_t)k]; m += Tl.leaf_end[(size_t)J] - Tl.leaf_begin[(size_t)J]; }
                Tl.col_ptr[(size_t)l + 1] = Tl.col_ptr[(size_t)l] + m;
                Tl.val_ptr[(size_t)l + 1] = Tl.val_ptr[(size_t)l] + m * (Tl.leaf_end[(size_t)l] - Tl.leaf_begin[(size_t)l]);
            }
            Tl.col_prim.resize( (size_t)Tl.col_ptr.back() );
            Tl.col_mask.resize( (size_t)Tl.col_ptr.back() );
            const bool any_excluded = !verynear_pairs.empty() || !inter_pairs.empty();
            std::vector<LInt> masked( (size_t)L, 0 );
            ParallelFor( L, threads, [&]( LInt l )
            {
                const Int ib = Tl.leaf_begin[(size_t)l], ie = Tl.leaf_end[(size_t)l];
                LInt pos = Tl.col_ptr[(size_t)l];
                LInt cntl = 0;
                for( LInt k = cnt[(size_t)l]; k < cnt[(size_t)l+1]; ++k )
                {
                    const Int J = nbr[(size_t)k];
                    for( Int j = Tl.leaf_begin[(size_t)J]; j < Tl.leaf_end[(size_t)J]; ++j )
                    {
                        uint32_t mask = 0;
                        for( Int i = ib; i < ie; ++i )
                        {
                            bool on = (i != j);
                            if( on && any_excluded ) { on = !ExcludedQ( i, j ); }
                            if( on ) { mask |= (1u << (i - ib)); ++cntl; }
                        }
                        Tl.col_prim[(size_t)pos] = j;
                        Tl.col_mask[(size_t)pos] = mask;
                        ++pos;
                    }
                }
                masked[(size_t)l] = cntl;
            });
            Tl.masked_pairs = 0;
            for( auto c : masked ) { Tl.masked_pairs += c; }
        }

    public:
        // Reference-layout near pattern (BinaryMatrixCSR, full symmetric, sorted, no diagonal).
        // With row_end >= 0 only rows [0, row_end) are filled (the others are empty).
        SymmetricCSR NearCSR( const Int row_end = -1 ) const
        {
            const Int T = S.PrimitiveCount();
            const Int R = (row_end < 0) ? T : std::min( row_end, T );
            SymmetricCSR A;
            A.rows = T;
            A.rp.assign( (size_t)T + 1, 0 );
            const NearTiles & Tl = tiles;
            for( Int l = 0; l < Tl.leaf_count; ++l )
            {
                for( Int i = Tl.leaf_begin[(size_t)l]; i < std::min( R, Tl.leaf_end[(size_t)l] ); ++i )
                {
                    LInt c = 0;
                    for( LInt k = Tl.col_ptr[(size_t)l]; k < Tl.col_ptr[(size_t)l+1]; ++k ) { c += (Tl.col_mask[(size_t)k] >> (i - Tl.leaf_begin[(size_t)l])) & 1u; }
                    A.rp[(size_t)i + 1] = c;
                }
            }
            for( Int i = 0; i < T; ++i ) { A.rp[(size_t)i + 1] += A.rp[(size_t)i]; }
            A.ci.resize( (size_t)A.rp.back() );
            A.diag.assign( (size_t)T, 0 );
            for( Int i = R; i < T; ++i ) { A.diag[(size_t)i] = A.rp[(size_t)i]; }
            for( Int l = 0; l < Tl.leaf_count; ++l )
            {
                for( Int i = Tl.leaf_begin[(size_t)l]; i < std::min( R, Tl.leaf_end[(size_t)l] ); ++i )
                {
                    LInt pos = A.rp[(size_t)i];
                    A.diag[(size_t)i] = -1;
                    for( LInt k = Tl.col_ptr[(size_t)l]; k < Tl.col_ptr[(size_t)l+1]; ++k )
                    {
                        if( (Tl.col_mask[(size_t)k] >> (i - Tl.leaf_begin[(size_t)l])) & 1u )
                        {
                            const Int j = Tl.col_prim[(size_t)k];
                            if( j > i && A.diag[(size_t)i] < 0 ) { A.diag[(size_t)i] = pos; }
                            A.ci[(size_t)pos++] = j;
                        }
                    }
                    if( A.diag[(size_t)i] < 0 ) { A.diag[(size_t)i] = pos; }
                }
            }
            return A;
        }

        SymmetricCSR VeryNearCSR() const
        {
            return SymmetricCSR::FromUpperPairs( S.PrimitiveCount(), verynear_pairs, threads );
        }
    };

} // namespace Repulsor
