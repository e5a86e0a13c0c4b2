#pragma once
// SimplicialMesh_Factory<MeshBase_T, MinDomDim, MaxDomDim, MinAmbDim, MaxAmbDim> (reference
// src/SimplicialMesh/SimplicialMesh_Factory.hpp:6-240): Make builds a SimplicialMesh of the dimensions
// given at run time behind a SimplicialMeshBase pointer; Make_FromFile reads the reference's text format
//
//     dom_dim <n>  amb_dim <m>  vertex_count <V>  simplex_count <T>
//     V lines of m coordinates, then T lines of n + 1 vertex indices (0-based)
//
// (whitespace-separated; the four keywords are read as tokens and not checked, as the reference does).
// Dimensions outside the factory's range, or not built here (curves and surfaces in R^3), give an error
// message and nullptr.

#include "SimplicialMesh.hpp"

#include <filesystem>
#include <fstream>

namespace Repulsor
{
    template<typename MeshBase_T, int ...N> class SimplicialMesh_Factory {};

    template<typename Real, typename Int, typename LInt, typename SReal, typename ExtReal,
             int MinDomDim_, int MaxDomDim_, int MinAmbDim_, int MaxAmbDim_>
    class SimplicialMesh_Factory<SimplicialMeshBase<Real,Int,LInt,SReal,ExtReal>, MinDomDim_, MaxDomDim_, MinAmbDim_, MaxAmbDim_>
    {
        static_assert( MinDomDim_ <= MaxDomDim_, "MinDomDim_ <= MaxDomDim_ required." );
        static_assert( MinAmbDim_ <= MaxAmbDim_, "MinAmbDim_ <= MaxAmbDim_ required." );

    public:
        using MeshBase_T = SimplicialMeshBase<Real,Int,LInt,SReal,ExtReal>;

        static constexpr int MinAmbDim = std::max( 1, MinAmbDim_ );
        static constexpr int MaxAmbDim = std::max( 1, MaxAmbDim_ );
        static constexpr int MinDomDim = std::max( 0, MinDomDim_ );
        static constexpr int MaxDomDim = std::min( MaxAmbDim - 1, MaxDomDim_ );

        SimplicialMesh_Factory() = default;

        // vertex_coords: vertex_count x amb_dim (amb_dim x vertex_count with vertex_coords_ColMajorQ);
        // simplices: simplex_count x simplex_size (transposed with simplices_ColMajorQ).
        template<typename ExtInt>
        [[nodiscard]] std::unique_ptr<MeshBase_T> Make(
            const ExtReal * const vertex_coords, const Int vertex_count, const Int amb_dim, const bool vertex_coords_ColMajorQ,
            const ExtInt  * const simplices,     const Int simplex_count, const Int simplex_size, const bool simplices_ColMajorQ,
            const Int thread_count = 1 ) const
        {
            const Int dom_dim = simplex_size - 1;
            if( dom_dim < MinDomDim || dom_dim > MaxDomDim ) { eprint( ClassName() + "::Make: dom_dim is out of range." ); return Error( dom_dim, amb_dim ); }
            if( amb_dim < MinAmbDim || amb_dim > MaxAmbDim ) { eprint( ClassName() + "::Make: amb_dim is out of range." ); return Error( dom_dim, amb_dim ); }
            if( dom_dim >= amb_dim || amb_dim != 3 ) { return Error( dom_dim, amb_dim ); }
            try
            {
                if constexpr ( MinDomDim <= 2 && 2 <= MaxDomDim )
                {
                    if( dom_dim == 2 )
                    {
                        return std::unique_ptr<MeshBase_T>( new SimplicialMesh<2,3,Real,Int,LInt,SReal,ExtReal>(
                            vertex_coords, vertex_count, vertex_coords_ColMajorQ, simplices, simplex_count, simplices_ColMajorQ, thread_count ) );
                    }
                }
                if constexpr ( MinDomDim <= 1 && 1 <= MaxDomDim )
                {
                    if( dom_dim == 1 )
                    {
                        return std::unique_ptr<MeshBase_T>( new SimplicialMesh<1,3,Real,Int,LInt,SReal,ExtReal>(
                            vertex_coords, vertex_count, vertex_coords_ColMajorQ, simplices, simplex_count, simplices_ColMajorQ, thread_count ) );
                    }
                }
            }
            catch( const std::exception & e ) { eprint( ClassName() + "::Make: " + e.what() ); return nullptr; }
            return Error( dom_dim, amb_dim );
        }

        [[nodiscard]] std::unique_ptr<MeshBase_T> Make_FromFile( const std::filesystem::path & file, const Int thread_count ) const
        {
            ScopedTimer timer( ClassName() + "::Make_FromFile" );
            std::ifstream s( file );
            if( !s.good() ) { eprint( "Make_FromFile: File " + file.string() + " could not be opened." ); return nullptr; }
            std::string tok;
            long long dom_dim = -1, amb_dim = -1, vertex_count = -1, simplex_count = -1;
            s >> tok >> dom_dim >> tok >> amb_dim >> tok >> vertex_count >> tok >> simplex_count;
            if( !s || dom_dim < 0 || amb_dim < 1 || vertex_count < 1 || simplex_count < 1
                || vertex_count > std::numeric_limits<Int>::max() || simplex_count > std::numeric_limits<Int>::max() )
            {
                eprint( "Make_FromFile: File " + file.string() + " has no valid header." );
                return nullptr;
            }
            const Int simplex_size = (Int)dom_dim + 1;
            std::vector<ExtReal> V( (size_t)( vertex_count * amb_dim ) );
            std::vector<Int> S( (size_t)( simplex_count * simplex_size ) );
            for( auto & x : V ) { s >> x; }
            for( auto & x : S ) { s >> x; }
            if( !s ) { eprint( "Make_FromFile: File " + file.string() + " ended before all coordinates and simplices were read." ); return nullptr; }
            return Make( V.data(), (Int)vertex_count, (Int)amb_dim, false, S.data(), (Int)simplex_count, simplex_size, false, thread_count );
        }

        std::string ClassName() const
        {
            return "SimplicialMesh_Factory<" + std::to_string( MinDomDim ) + "," + std::to_string( MaxDomDim ) + ","
                 + std::to_string( MinAmbDim ) + "," + std::to_string( MaxAmbDim ) + ">";
        }

    private:
        std::unique_ptr<MeshBase_T> Error( const Int dom_dim, const Int amb_dim ) const
        {
            eprint( ClassName() + " cannot create SimplicialMesh with domain dimension " + std::to_string( dom_dim )
                  + " and ambient dimension " + std::to_string( amb_dim ) + "." );
            return nullptr;
        }
    };

} // namespace Repulsor
