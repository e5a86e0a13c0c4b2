#pragma once
// Binary AABB cluster tree over the primitives, numbered depth-first.
//
// Restates the host-side construction of the reference's ClusterTree
// (src/ClusterTree.hpp:115-180, src/ClusterTree/Split_Thread.hpp:3-166,
//  src/ClusterTree/Serialize_Thread.hpp:3-156, src/BoundingVolumes/AABB_LongestAxisSplit.hpp:134-253,
//  src/BoundingVolumes/AABB.hpp:155-237, src/Primitives/Polytope.hpp FromIndexList/BoxMinMax,
//  src/ClusterTree/ComputeClusterData_Thread.hpp:3-129, src/ClusterTree.hpp:319-338,410-639).
//
// Layouts (all row-major):
//   primitive serialized  [r^2, interior point (3), hull (n+1)x3]  (Polytope data layout)
//   cluster AABB          [r^2, center (3), half edge lengths (3)] (AABB data layout)
//   P_near  T x NEAR_DIM, P_far T x 10, C_far C x 10  in tree order.
//   mixed pre-operator per primitive i (tree order): columns = sorted simplex of primitive
//   P_ordering[i]; values [4][n+1]: row 0 = a_i/(n+1) (averaging), rows 1..3 = a_i * DiffOp.

#include "Common.hpp"
#include "Settings.hpp"
#include "SimplexData.hpp"

namespace Repulsor
{
    // Tree arrays built on the device (rp_build_topology): depth-first numbering, as ComputeClusters
    // produces them.
    struct BuiltClusterTree
    {
        std::vector<Int> P_ordering, C_left, C_right, C_begin, C_end, C_depth;
        std::vector<double> C_serialized;   // C x 7
    };

    template<int DOM_DIM>
    class ClusterTree
    {
    public:
        static constexpr int AMB_DIM   = 3;
        static constexpr int SIZE      = DOM_DIM + 1;
        static constexpr int NEAR_DIM  = SimplexData<DOM_DIM>::NEAR_DIM;
        static constexpr int FAR_DIM   = SimplexData<DOM_DIM>::FAR_DIM;
        static constexpr int P_SIZE    = 1 + AMB_DIM + SIZE * AMB_DIM;   // primitive serialized size
        static constexpr int C_SIZE    = 1 + AMB_DIM + AMB_DIM;          // AABB serialized size
        static constexpr int PRE_ROWS  = AMB_DIM + 1;

    protected:
        struct Node
        {
            Int begin, end, depth;
            Int left  = -1;   // index into nodes
            Int right = -1;
            Int desc_count = 1;
            Int desc_leaf_count = 1;
            double box[C_SIZE] = {};
        };

        ClusterTreeSettings settings;
        int threads = 1;

        Int primitive_count = 0;

    public:
        // Tree arrays (depth-first numbering).
        std::vector<Int> C_left, C_right, C_begin, C_end, C_depth, C_next, C_parent;
        std::vector<Int> leaf_clusters, leaf_cluster_lookup, leaf_cluster_ptr;
        std::vector<double> C_serialized;   // C x C_SIZE

        std::vector<Int> P_ordering, P_inverse_ordering;
        std::vector<double> P_serialized;   // T x P_SIZE (tree order)

        std::vector<double> P_near, P_far, C_far;

        // Mixed pre-operator (fixed at construction, like the reference's pre_post_initialized guard).
        std::vector<Int>    pre_cols;   // T x SIZE
        std::vector<double> pre_vals;   // T x PRE_ROWS x SIZE

        Int depth = 0;

    public:
        ClusterTree() = default;

        // V_frozen: coordinates used for clustering; P_near_orig/P_far_orig: simplex data of the
        // current coordinates in original order; DiffOp values T x (3*SIZE) (rows k, sorted columns).
        ClusterTree(
            cptr<double> V_frozen, cptr<Int> S, const LInt T,
            cptr<double> P_near_orig, cptr<double> P_far_orig, cptr<double> DiffOp_vals, cptr<Int> sorted_simplices,
            const ClusterTreeSettings & settings_, const int threads_, const BuiltClusterTree * built = nullptr )
        :   settings( settings_ ), threads( std::max(1, threads_) ), primitive_count( (Int)T )
        {
            ScopedTimer timer("ClusterTree::ClusterTree");
            // Primitive serialized data (Polytope::FromIndexList on frozen coordinates).
            P_serialized.resize( (size_t)T * P_SIZE );
            ParallelFor( T, threads, [&]( LInt i )
            {
                double * p = &P_serialized[ (size_t)P_SIZE * i ];
                double * center = p + 1;
                double * hull   = p + 1 + AMB_DIM;
                for( int j = 0; j < SIZE; ++j )
                {
                    for( int k = 0; k < AMB_DIM; ++k ) { hull[AMB_DIM * j + k] = V_frozen[ (LInt)AMB_DIM * S[(LInt)SIZE * i + j] + k ]; }
                }
                const double w = 1.0 / SIZE;
                for( int k = 0; k < AMB_DIM; ++k )
                {
                    center[k] = hull[k];
                    for( int j = 1; j < SIZE; ++j ) { center[k] += hull[AMB_DIM * j + k]; }
                    center[k] *= w;
                }
                double r2 = 0;
                for( int j = 0; j < SIZE; ++j )
                {
                    double d = hull[AMB_DIM * j] - center[0];
                    double sq = d * d;
                    for( int k = 1; k < AMB_DIM; ++k ) { d = hull[AMB_DIM * j + k] - center[k]; sq += d * d; }
                    r2 = std::max( r2, sq );
                }
                p[0] = r2;
            });

            P_ordering.resize( (size_t)T );
            for( LInt i = 0; i < T; ++i ) { P_ordering[(size_t)i] = (Int)i; }

            if( built ) { AdoptClusters( *built ); } else { ComputeClusters(); }

            // Tree-ordered primitive data (ComputePrimitiveData, src/ClusterTree.hpp:319-338).
            P_near.resize( (size_t)T * NEAR_DIM );
            P_far .resize( (size_t)T * FAR_DIM  );
            ComputePrimitiveData( P_near_orig, P_far_orig );
            C_far.assign( (size_t)ClusterCount() * FAR_DIM, 0.0 );
            ComputeClusterData();

            // Mixed pre-operator (ComputePrePost + RequireMixedPrePost, src/ClusterTree.hpp:410-639).
            pre_cols.resize( (size_t)T * SIZE );
            pre_vals.resize( (size_t)T * PRE_ROWS * SIZE );
            ParallelFor( T, threads, [&]( LInt i )
            {
                const Int j = P_ordering[(size_t)i];
                const double a = P_near[ (size_t)NEAR_DIM * i ];
                for( int l = 0; l < SIZE; ++l ) { pre_cols[(size_t)SIZE * i + l] = sorted_simplices[(LInt)SIZE * j + l]; }
                double * pv = &pre_vals[ (size_t)PRE_ROWS * SIZE * i ];
                for( int l = 0; l < SIZE; ++l ) { pv[l] = a * (1.0 / SIZE); }
                for( int k = 0; k < AMB_DIM; ++k )
                {
                    for( int l = 0; l < SIZE; ++l ) { pv[SIZE * (1 + k) + l] = a * DiffOp_vals[ (LInt)AMB_DIM * SIZE * j + SIZE * k + l ]; }
                }
            });
        }

        Int PrimitiveCount()   const { return primitive_count; }
        Int ClusterCount()     const { return (Int)C_left.size(); }
        Int LeafClusterCount() const { return (Int)leaf_clusters.size(); }
        Int SplitThreshold()   const { return settings.split_threshold; }
        Int TreeDepth()        const { return depth; }
        int ThreadCount()      const { return threads; }
        int NearDim()          const { return NEAR_DIM; }
        int FarDim()           const { return FAR_DIM; }

        // Updates only the computational data, as ClusterTree::SemiStaticUpdate does
        // (src/ClusterTree.hpp:983-993); tree, BCT and pre/post operators stay frozen.
        void SemiStaticUpdate( cptr<double> P_near_orig, cptr<double> P_far_orig )
        {
            ScopedTimer timer("ClusterTree::SemiStaticUpdate");
            ComputePrimitiveData( P_near_orig, P_far_orig );
            ComputeClusterData();
        }

    protected:
        void BoxFromPrimitives( double * box, const Int begin, const Int end ) const
        {
            double lower[AMB_DIM], upper[AMB_DIM];
            for( int k = 0; k < AMB_DIM; ++k ) { lower[k] = std::numeric_limits<double>::max(); upper[k] = std::numeric_limits<double>::lowest(); }
            for( Int i = begin; i < end; ++i )
            {
                const double * hull = &P_serialized[ (size_t)P_SIZE * i + 1 + AMB_DIM ];
                for( int j = 0; j < SIZE; ++j )
                {
                    for( int k = 0; k < AMB_DIM; ++k )
                    {
                        const double x = hull[AMB_DIM * j + k];
                        lower[k] = std::min( lower[k], x );
                        upper[k] = std::max( upper[k], x );
                    }
                }
            }
            double r2 = 0;
            for( int k = 0; k < AMB_DIM; ++k )
            {
                const double diff = 0.5 * (upper[k] - lower[k]);
                r2 += diff * diff;
                lower[k] += diff;
                upper[k]  = diff;
            }
            box[0] = r2;
            for( int k = 0; k < AMB_DIM; ++k ) { box[1 + k] = lower[k]; box[1 + AMB_DIM + k] = upper[k]; }
        }

        void SwapPrimitives( const Int i, const Int j )
        {
            if( i == j ) { return; }
            std::swap( P_ordering[(size_t)i], P_ordering[(size_t)j] );
            double * a = &P_serialized[ (size_t)P_SIZE * i ];
            double * b = &P_serialized[ (size_t)P_SIZE * j ];
            for( int k = 0; k < P_SIZE; ++k ) { std::swap( a[k], b[k] ); }
        }

        // AABB_LongestAxisSplit::split (src/BoundingVolumes/AABB_LongestAxisSplit.hpp:134-253).
        Int SplitRange( const double * box, const Int begin, const Int end,
                        std::vector<double> & score, std::vector<Int> & perm, std::vector<Int> & inv_perm )
        {
            int split_dir = 0;
            double L_max = box[1 + AMB_DIM];
            for( int k = 1; k < AMB_DIM; ++k )
            {
                const double L_k = box[1 + AMB_DIM + k];
                if( L_k > L_max ) { L_max = L_k; split_dir = k; }
            }
            const double mid = box[1 + split_dir];
            Int split_index = begin;
            for( Int i = begin; i < end; ++i )
            {
                const double x = P_serialized[ (size_t)P_SIZE * i + 1 + split_dir ];
                if( x < mid )
                {
                    SwapPrimitives( split_index, i );
                    ++split_index;
                }
            }
            if( split_index == begin || split_index == end )
            {
                // Median split fallback. The reference partitions with std::nth_element, which leaves the
                // order inside the two halves to the standard library; this restatement sorts the range
                // stably by the interior coordinate, one of the orders nth_element may produce, so that
                // the tree is defined independently of the library (and the device build reproduces it).
                split_index = begin + (end - begin) / 2;
                for( Int i = begin; i < end; ++i )
                {
                    score[(size_t)i] = P_serialized[ (size_t)P_SIZE * i + 1 + split_dir ];
                    perm [(size_t)i] = i;
                }
                const double * sc = score.data();
                std::stable_sort( perm.data() + begin, perm.data() + end,
                    [sc]( const Int i, const Int j ) { return sc[i] < sc[j]; } );
                for( Int i = begin; i < end; ++i ) { inv_perm[(size_t)perm[(size_t)i]] = i; }
                for( Int i = begin; i < end; ++i )
                {
                    Int next = i;
                    while( inv_perm[(size_t)next] >= 0 )
                    {
                        const Int temp = inv_perm[(size_t)next];
                        SwapPrimitives( i, temp );
                        inv_perm[(size_t)next] = -1;
                        next = temp;
                    }
                }
            }
            return split_index;
        }

        void ComputeClusters()
        {
            ScopedTimer timer("ClusterTree::ComputeClusters");
            const Int T = primitive_count;
            std::vector<Node> nodes;
            nodes.reserve( (size_t)std::max<Int>( 1, 2 * T ) );
            nodes.push_back( Node{ 0, T, 0 } );
            BoxFromPrimitives( nodes[0].box, 0, T );

            std::vector<double> score( (size_t)T );
            std::vector<Int> perm( (size_t)T ), inv_perm( (size_t)T );

            // Depth-first split; the result does not depend on the visiting order because every
            // split only permutes the primitives inside its own range.
            std::vector<Int> stack;
            stack.push_back(0);
            while( !stack.empty() )
            {
                const Int c = stack.back(); stack.pop_back();
                const Int begin = nodes[(size_t)c].begin, end = nodes[(size_t)c].end;
                if( end - begin > settings.split_threshold )
                {
                    double box[C_SIZE];
                    for( int k = 0; k < C_SIZE; ++k ) { box[k] = nodes[(size_t)c].box[k]; }
                    const Int s = SplitRange( box, begin, end, score, perm, inv_perm );
                    const Int d = nodes[(size_t)c].depth + 1;
                    Node L{ begin, s, d };
                    Node R{ s, end, d };
                    BoxFromPrimitives( L.box, begin, s );
                    BoxFromPrimitives( R.box, s, end );
                    const Int li = (Int)nodes.size();
                    nodes.push_back( L );
                    nodes.push_back( R );
                    nodes[(size_t)c].left  = li;
                    nodes[(size_t)c].right = li + 1;
                    stack.push_back( li + 1 );
                    stack.push_back( li );
                }
            }
            // Descendant counts (children have larger indices than parents).
            for( size_t c = nodes.size(); c-- > 0; )
            {
                Node & N = nodes[c];
                if( N.left >= 0 )
                {
                    N.desc_count      = 1 + nodes[(size_t)N.left].desc_count + nodes[(size_t)N.right].desc_count;
                    N.desc_leaf_count = nodes[(size_t)N.left].desc_leaf_count + nodes[(size_t)N.right].desc_leaf_count;
                }
            }
            // Serialize in depth-first order (Serialize_Cluster, src/ClusterTree/Serialize_Thread.hpp:107-148).
            const Int C = nodes[0].desc_count;
            const Int L = nodes[0].desc_leaf_count;
            C_left.assign( (size_t)C, -1 ); C_right.assign( (size_t)C, -1 );
            C_begin.assign( (size_t)C, 0 ); C_end.assign( (size_t)C, 0 );
            C_depth.assign( (size_t)C, 0 ); C_next.assign( (size_t)C, 0 ); C_parent.assign( (size_t)C, -1 );
            C_serialized.assign( (size_t)C * C_SIZE, 0.0 );
            leaf_clusters.assign( (size_t)L, 0 );
            leaf_cluster_lookup.assign( (size_t)C, -1 );
            struct Item { Int node; Int g; Int leafs_before; };
            std::vector<Item> st;
            st.push_back( { 0, 0, 0 } );
            depth = 0;
            while( !st.empty() )
            {
                const Item it = st.back(); st.pop_back();
                const Node & N = nodes[(size_t)it.node];
                const Int g = it.g;
                C_begin[(size_t)g] = N.begin;
                C_end  [(size_t)g] = N.end;
                C_depth[(size_t)g] = N.depth;
                C_next [(size_t)g] = g + N.desc_count;
                depth = std::max( depth, N.depth );
                for( int k = 0; k < C_SIZE; ++k ) { C_serialized[(size_t)C_SIZE * g + k] = N.box[k]; }
                if( N.left >= 0 )
                {
                    const Int gl = g + 1;
                    const Int gr = g + 1 + nodes[(size_t)N.left].desc_count;
                    C_left [(size_t)g] = gl;
                    C_right[(size_t)g] = gr;
                    C_parent[(size_t)gl] = g;
                    C_parent[(size_t)gr] = g;
                    st.push_back( { N.right, gr, it.leafs_before + nodes[(size_t)N.left].desc_leaf_count } );
                    st.push_back( { N.left,  gl, it.leafs_before } );
                }
                else
                {
                    leaf_clusters[(size_t)it.leafs_before] = g;
                    leaf_cluster_lookup[(size_t)g] = it.leafs_before;
                }
            }
            P_inverse_ordering.assign( (size_t)primitive_count, 0 );
            for( Int i = 0; i < primitive_count; ++i ) { P_inverse_ordering[(size_t)P_ordering[(size_t)i]] = i; }
            leaf_cluster_ptr.assign( (size_t)L + 1, 0 );
            for( Int i = 0; i < L; ++i ) { leaf_cluster_ptr[(size_t)i + 1] = C_end[(size_t)leaf_clusters[(size_t)i]]; }
        }

        // Takes over a tree built elsewhere (rp_build_topology) and derives the same members as
        // ComputeClusters: tree-ordered serialized primitives, parents, subtree ends, leaf lists.
        void AdoptClusters( const BuiltClusterTree & B )
        {
            ScopedTimer timer("ClusterTree::AdoptClusters");
            const Int T = primitive_count;
            const Int C = (Int)B.C_left.size();
            if( (Int)B.P_ordering.size() != T || C < 1 || (Int)B.C_serialized.size() != C_SIZE * C )
            {
                throw std::runtime_error( "AdoptClusters: inconsistent tree arrays" );
            }
            P_ordering = B.P_ordering;
            {
                std::vector<double> ser( P_serialized.size() );
                ParallelFor( T, threads, [&]( LInt i )
                {
                    std::memcpy( &ser[(size_t)P_SIZE * i], &P_serialized[(size_t)P_SIZE * P_ordering[(size_t)i]], sizeof(double) * P_SIZE );
                });
                P_serialized.swap( ser );
            }
            C_left = B.C_left; C_right = B.C_right; C_begin = B.C_begin; C_end = B.C_end; C_depth = B.C_depth;
            C_serialized = B.C_serialized;
            C_parent.assign( (size_t)C, -1 );
            C_next.assign( (size_t)C, 0 );
            std::vector<Int> desc( (size_t)C, 1 );
            for( Int c = C; c-- > 0; )
            {
                if( C_left[(size_t)c] >= 0 )
                {
                    desc[(size_t)c] = 1 + desc[(size_t)C_left[(size_t)c]] + desc[(size_t)C_right[(size_t)c]];
                    C_parent[(size_t)C_left[(size_t)c]] = c;
                    C_parent[(size_t)C_right[(size_t)c]] = c;
                }
            }
            depth = 0;
            leaf_clusters.clear();
            leaf_cluster_lookup.assign( (size_t)C, -1 );
            for( Int c = 0; c < C; ++c )
            {
                C_next[(size_t)c] = c + desc[(size_t)c];
                depth = std::max( depth, C_depth[(size_t)c] );
                if( C_left[(size_t)c] < 0 )
                {
                    leaf_cluster_lookup[(size_t)c] = (Int)leaf_clusters.size();
                    leaf_clusters.push_back( c );
                }
            }
            const Int L = (Int)leaf_clusters.size();
            P_inverse_ordering.assign( (size_t)T, 0 );
            for( Int i = 0; i < T; ++i ) { P_inverse_ordering[(size_t)P_ordering[(size_t)i]] = i; }
            leaf_cluster_ptr.assign( (size_t)L + 1, 0 );
            for( Int i = 0; i < L; ++i ) { leaf_cluster_ptr[(size_t)i + 1] = C_end[(size_t)leaf_clusters[(size_t)i]]; }
        }

        void ComputePrimitiveData( cptr<double> P_near_orig, cptr<double> P_far_orig )
        {
            ParallelFor( primitive_count, threads, [&]( LInt i )
            {
                const LInt j = P_ordering[(size_t)i];
                std::memcpy( &P_near[(size_t)NEAR_DIM * i], &P_near_orig[NEAR_DIM * j], sizeof(double) * NEAR_DIM );
                std::memcpy( &P_far [(size_t)FAR_DIM  * i], &P_far_orig [FAR_DIM  * j], sizeof(double) * FAR_DIM  );
            });
        }

        // ComputeClusterData_Step (src/ClusterTree/ComputeClusterData_Thread.hpp:50-129), applied in
        // reverse depth-first order so that children are done before parents.
        void ComputeClusterData()
        {
            const bool singletons = settings.split_threshold <= 1;
            for( Int C = ClusterCount(); C-- > 0; )
            {
                const Int L = C_left[(size_t)C];
                const Int R = C_right[(size_t)C];
                double * C_C = &C_far[(size_t)FAR_DIM * C];
                if( L >= 0 )
                {
                    const double * C_L = &C_far[(size_t)FAR_DIM * L];
                    const double * C_R = &C_far[(size_t)FAR_DIM * R];
                    double L_weight = C_L[0];
                    double R_weight = C_R[0];
                    const double C_mass = L_weight + R_weight;
                    C_C[0] = C_mass;
                    const double C_invmass = 1.0 / C_mass;
                    L_weight *= C_invmass;
                    R_weight *= C_invmass;
                    for( int k = 1; k < FAR_DIM; ++k ) { C_C[k] = L_weight * C_L[k] + R_weight * C_R[k]; }
                }
                else if( singletons )
                {
                    std::memcpy( C_C, &P_far[(size_t)FAR_DIM * C_begin[(size_t)C]], sizeof(double) * FAR_DIM );
                }
                else
                {
                    const Int begin = C_begin[(size_t)C], end = C_end[(size_t)C];
                    {
                        const double * P = &P_far[(size_t)FAR_DIM * begin];
                        const double a = P[0];
                        C_C[0] = a;
                        for( int j = 1; j < FAR_DIM; ++j ) { C_C[j] = a * P[j]; }
                    }
                    for( Int i = begin + 1; i < end; ++i )
                    {
                        const double * P = &P_far[(size_t)FAR_DIM * i];
                        const double a = P[0];
                        C_C[0] += a;
                        for( int j = 1; j < FAR_DIM; ++j ) { C_C[j] += a * P[j]; }
                    }
                    const double C_invmass = 1.0 / C_C[0];
                    for( int j = 1; j < FAR_DIM; ++j ) { C_C[j] *= C_invmass; }
                }
            }
        }
    };

} // namespace Repulsor
