#pragma once
// TangentPointEnergy0 with the reference's public surface
// (src/Energies/TangentPointEnergy0.hpp:35-86, src/Energies/EnergyBase.hpp:40-121):
//   Value(M), Differential(M) (cached), Differential(M, alpha, beta, Y, ldY), Differential(M, Y, addto).
// The near/far kernels, derivative collection and chain rule run on the GPU through librepulsor_hip.so.

#include "SimplicialMesh.hpp"

namespace Repulsor
{
    template<typename Mesh_T>
    class TangentPointEnergy0
    {
    public:
        static constexpr int DOM_DIM = Mesh_T::DOM_DIM;

    protected:
        double q, p;
        Precision precision = Precision::FP64;

    public:
        // (p): q = p, p = 2p  (src/Energies/TangentPointEnergy0.hpp:35-39)
        explicit TangentPointEnergy0( const double p_ ) : q( p_ ), p( 2 * p_ ) {}
        TangentPointEnergy0( const double q_, const double p_ ) : q( q_ ), p( p_ ) {}

        void SetPrecision( Precision prec ) { precision = prec; }
        Precision GetPrecision() const { return precision; }
        double Q() const { return q; }
        double P() const { return p; }

        std::string ClassName() const
        {
            char buf[128];
            std::snprintf( buf, sizeof(buf), "TangentPointEnergy0<%d,3>(%g,%g)%s", DOM_DIM, q, p, PrecisionTag( precision ) );
            return std::string(buf);
        }

        // Energy value; cached in the mesh until the next SemiStaticUpdate.
        double Value( const Mesh_T & M ) const
        {
            const std::string tag = ClassName() + "::Value";
            auto it = M.scalar_cache.find( tag );
            if( it != M.scalar_cache.end() ) { return it->second; }
            double en = 0;
            if( !compute( M, RP_ENERGY, &en ) ) { return 0; }
            M.scalar_cache[tag] = en;
            return en;
        }

        // Cached differential (V x 3); also stores the value as the reference does (EnergyBase.hpp:62-87).
        const std::vector<double> & Differential( const Mesh_T & M ) const
        {
            const std::string tag = ClassName() + "::Differential";
            auto it = M.vector_cache.find( tag );
            if( it != M.vector_cache.end() ) { return it->second; }
            std::vector<double> diff( (size_t)(3 * M.VertexCount()), 0.0 );
            const double en = Differential( M, 1.0, 0.0, diff.data(), 3 );
            M.scalar_cache.emplace( ClassName() + "::Value", en );
            return M.vector_cache.emplace( tag, std::move(diff) ).first->second;
        }

        // Y = alpha * dE + beta * Y; returns the energy (EnergyBase.hpp:90-104). No caching.
        double Differential( const Mesh_T & M, const double alpha, const double beta, double * Y, const LInt ldY ) const
        {
            ScopedTimer timer( "TangentPointEnergy0::Differential (pointer)" );
            double en = 0;
            if( !compute( M, RP_ENERGY | RP_DIFF, &en ) ) { return 0; }
            try
            {
                if( !Check( rp_assemble_vertex_gradient( M.Device(), alpha, beta, Y, ldY, 0 ), ClassName() + "::Differential" ) ) { return 0; }
            }
            catch( const std::exception & e ) { eprint( ClassName() + "::Differential: " + e.what() ); return 0; }
            return en;
        }

        // Y = dE (+ Y if addto) with ldY = 3 (EnergyBase.hpp:107-121).
        double Differential( const Mesh_T & M, double * Y, const bool addto = false ) const
        {
            return Differential( M, 1.0, addto ? 1.0 : 0.0, Y, 3 );
        }

        // Energy and differential into a device buffer (no host round trip).
        double DifferentialDevice( const Mesh_T & M, const double alpha, const double beta, double * Y_dev, const LInt ldY ) const
        {
            double en = 0;
            if( !compute( M, RP_ENERGY | RP_DIFF, &en ) ) { return 0; }
            if( !Check( rp_assemble_vertex_gradient( M.Device(), alpha, beta, Y_dev, ldY, RP_DEVICE_PTR ), ClassName() + "::DifferentialDevice" ) ) { return 0; }
            return en;
        }

        // Moment derivatives of the last differential (ClusterTree::PrimitiveDNearFieldData/DFarFieldData).
        bool MomentDerivatives( const Mesh_T & M, double * P_D_near, double * P_D_far ) const
        {
            return Check( rp_download_derivatives( M.Device(), P_D_near, P_D_far ), ClassName() + "::MomentDerivatives" );
        }

    protected:
        bool compute( const Mesh_T & M, unsigned flags, double * en ) const
        {
            try
            {
                rp_ctx * c = M.Device();
                return Check( rp_tp0_compute( c, flags, q, p, (int)precision, en ), ClassName() );
            }
            catch( const std::exception & e ) { eprint( ClassName() + ": " + e.what() ); return false; }
        }
    };

} // namespace Repulsor
